"""Benchmark of the periodic-box P_l(k) path on one MI355X.

Workload (the configuration the headline metric is quoted on): 134,217,728
particles (a 512^3 lattice displaced by Zel'dovich displacements of a seeded
Gaussian random field with P(k) = 2e4 (k/0.1)^-1.5, f = 0.5, aH = 100) in a
2000 Mpc/h box, float32 positions/velocities, unit weights; 1024^3 grid,
TSC + interlacing, density and line-of-sight momentum fields (reference
Momentum = 1: grids delta, delta_int, p, p_int), spectra P00, P11 and the
P01 cross, multipoles l = 0, 2, 4, dk = 0.005 h/Mpc (NK = 321).

A step = one pass of the hot path with the catalogue resident in HBM:
mass assignment (sort + tile deposit + mean subtraction), r2c FFT of the
four grids, fused binning -> raw bins on the host.  `value` is the FP32
form's ms per step; the FP64 plain form is timed beside it.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "ms per P_0,2,4(k) at 1024³ TSC+interlace, FP32 vs FP64 GPU; mocks/s batched"
HBM_PEAK_GBS = 8000.0
COPY_BYTES = 1 << 30
TRAFFIC_FILE = "c2_fp32_traffic.md"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def algorithmic_bytes(kernel, n_p, ng_reals, G, s):
    """Compulsory HBM bytes per launch of each kernel (sort scratch excluded):
    particles 16 B (x, y, z, v_los in FP32; 32 B in FP64), a grid N_g*s."""
    bp = 4 * s
    grid = ng_reals * s
    return {
        "deposit": n_p * bp + G * grid,        # read sorted particles, write G grids once
        "fft_z": 2 * G * grid,                 # read + write every grid
        "fft_y": 2 * G * grid,
        "fft_x": 2 * G * grid,
        "fft_x_bin": G * grid,                 # x pass fused with the binning: read every grid once
        "bin": G * grid,                       # read every grid once
        "bucket_count": n_p * bp,
        "bucket_scatter": 2 * n_p * bp,
    }.get(kernel)


def run_form(prec, parts, n_p, args, geom):
    from capow_amd import api
    nx = geom["n"]
    p = api.make_params(nx=nx, ny=nx, nz=nx, xmin=0.0, xmax=geom["L"], ymin=0.0, ymax=geom["L"],
                        zmin=0.0, zmax=geom["L"], interp_order=3, do_interlacing=1, los=0,
                        momentum=1, nk=geom["nk"], maxk=geom["nk"] * geom["dk"], precision=prec,
                        extra_spectra=api.SPEC_AUTO_A | api.SPEC_AUTO_B)
    ps = api.PowerSpectrum(p)
    ps.set_particles(parts)
    import torch
    # the project's autotuner picks the sort, deposit and FFT settings on
    # this catalogue before the warmup (untimed; its report is recorded)
    tune = None if args.no_autotune else ps.autotune(reps=2)
    for _ in range(args.warmup):
        res = ps.power()
    torch.cuda.synchronize()
    steps = args.steps if prec == api.FP32 else max(1, args.fp64_steps)
    # timed region: per-kernel HIP events on the context's stream, recorded
    # around every launch of the K steps (the step time is the host clock
    # around the K steps, which includes the event records)
    ps.set_timing(True)
    ps.reset_kernel_times()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        res = ps.power()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    ktimes = ps.kernel_times()
    ps.set_timing(False)
    out = dict(ms=ms, kernels=ktimes, res=res, stats=ps.stats, params=p, steps=steps,
               autotune=(tune.strip().splitlines()[-1] if tune else None))
    ps.close()
    return out


def run_batch(args):
    """Batched mocks: `args.mocks` catalogues in a 500 Mpc/h box, particle
    counts lognormal (median 5e5, sigma_ln 0.6, clipped to [1e5, 3e6]), each
    half Zel'dovich lattice (P(k) as above, per-mock seed) and half uniform;
    grids alternate 128^3 (NK = 80) and 192^3 (NK = 120), CIC + interlacing,
    density P_0,2,4(k), dk = 0.01, FP32 grids with FP64 bins.  All catalogues
    are resident on the GPU before the timed run."""
    import torch
    from capow_amd import api, mocks
    Lb = 500.0
    counts = mocks.lognormal_counts(args.mocks, seed=2024)
    cfgs = [dict(nx=g, ny=g, nz=g, xmax=Lb, ymax=Lb, zmax=Lb, interp_order=2, do_interlacing=1,
                 momentum=0, nk=nk, maxk=nk * 0.01, precision=api.FP32)
            for g, nk in ((128, 80), (192, 120))]
    which = [i % 2 for i in range(args.mocks)]
    t0 = time.perf_counter()
    cats = [mocks.mixed_mock_torch(int(n), Lb, seed=1000 + i) for i, n in enumerate(counts)]
    torch.cuda.synchronize()
    ntot = int(sum(c.shape[0] for c in cats))
    log("batch: %d mocks, %d particles generated in %.1f s" % (len(cats), ntot, time.perf_counter() - t0))
    b = api.Batch(cfgs, max_particles=int(counts.max()), nstreams=args.mock_streams)
    if args.no_autotune:
        b.set_stack(64)
    b.run(cats, which)                    # warmup: one untimed pass over every catalogue
    # the autotuner picks the per-set settings, the streams in use and the
    # mode: stacks of 16 or 64 catalogues (one launch per stage per stack)
    # or one catalogue per launch
    tune = None if args.no_autotune else b.autotune(cats, which, reps=2)
    chosen = [ln.strip() for ln in tune.splitlines() if ln.startswith("chosen")] if tune else None
    if tune:
        log("batch autotune:\n" + tune.rstrip())
    stack = 64 if args.no_autotune else int(
        [c for c in chosen if c.startswith("chosen: stack=")][0].split("=")[1])

    def timed_run():
        # one untimed run (a graph: eager + capture), then the median of 3
        b.run(cats, which)
        ts = []
        for _ in range(3):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            r = b.run(cats, which)
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - t0)
        return r, sorted(ts)[1]

    (res, stats, ms), wall = timed_run()
    # graph mode: the same list captured into one HIP graph, then one timed
    # replay, checked against the eager run's bins
    b.set_graph(True)
    (gres, _, _), gwall = timed_run()
    b.set_graph(False)
    same = all(np.array_equal(a[c]["pk"], g[c]["pk"]) for a, g in zip(res, gres) for c in a["codes"])
    # the other mode beside it (stacks <-> per-catalogue streams), bins
    # checked bitwise across the modes
    owall, same_modes = None, None
    if not args.no_other_mode:
        b.set_stack(0 if stack else 64)
        (ores, _, _), owall = timed_run()
        same_modes = all(np.array_equal(a[c]["pk"], g[c]["pk"])
                         for a, g in zip(res, ores) for c in a["codes"])
    b.close()
    del cats
    torch.cuda.empty_cache()
    # algorithmic HBM bytes: per particle 16 B read by the sort's count,
    # 16 + 16 B by its scatter, 16 B by the deposit; per cell and grid (A,
    # A_int) 4 B deposit write, 4 + 4 B z pass, 4 + 4 B y pass, 4 B fused
    # x pass + binning
    cells = sum(float(cfgs[q]["nx"]) ** 3 for q in which)
    algo = 64.0 * ntot + 2 * 24.0 * cells
    gbs = algo / wall / 1e9
    mode = "stacks of %d" % stack if stack else "one catalogue per launch"
    out = dict(mocks=len(counts), mocks_per_s=round(len(counts) / wall, 1), mode=mode,
               autotune=chosen,
               graph_mocks_per_s=round(len(counts) / gwall, 1), graph_bins_equal_eager=bool(same),
               other_mode=("one catalogue per launch" if stack else "stacks of 64"),
               other_mode_mocks_per_s=round(len(counts) / owall, 1) if owall else None,
               bins_equal_across_modes=same_modes,
               wall_s=round(wall, 4), p50_ms=round(float(np.percentile(ms, 50)), 3),
               p99_ms=round(float(np.percentile(ms, 99)), 3), particles=ntot,
               mparticles_per_s=round(ntot / wall / 1e6, 1), streams=args.mock_streams,
               roofline={"bound": "hbm", "achieved": round(gbs, 1), "peak": 8000.0, "unit": "GB/s",
                         "frac": round(gbs / 8000.0, 3), "algo_bytes": algo,
                         "note": "whole-batch algorithmic bytes over the batch's wall time"},
               grids="128^3 (NK 80) / 192^3 (NK 120) alternating", box=Lb,
               interp="CIC+interlacing", dtype="f32 grids, f64 bins",
               per_mock_ms=("stacks: GPU time of the mock's stack; one catalogue per launch: "
                            "GPU time from the mock's first to last stage on its stream"))
    log("batch (%s): %.1f mocks/s (graph replay %.1f, bins equal %s; %s %s, equal %s), "
        "%.0f GB/s algorithmic"
        % (mode, out["mocks_per_s"], out["graph_mocks_per_s"], same, out["other_mode"],
           out["other_mode_mocks_per_s"], same_modes, gbs))
    return out


def run_adjoint(args):
    """Gradient configuration: 16,777,216 Zel'dovich particles (256^3
    lattice, L = 1000 Mpc/h), 512^3 grid, TSC + interlacing, density P_0,2,4,
    FP32 form: one forward (capow_power, keeping its sort and transform for
    the backward: capow_keep_forward) and one backward right after it
    (capow_backward: the pending x pass, binning backward, inverse FFT,
    gather), G = dL/d(raw bins) fixed; beside it the backward that recomputes
    the assignment and FFT (particles set again before it)."""
    import torch
    from capow_amd import api, mocks
    Lb, n = 1000.0, 512
    parts = mocks.zeldovich_torch(256, Lb, seed=1, device="cuda", dtype=torch.float32)
    w = torch.ones(parts.shape[0], device="cuda", dtype=torch.float32)
    p = api.make_params(nx=n, ny=n, nz=n, xmax=Lb, ymax=Lb, zmax=Lb, interp_order=3,
                        do_interlacing=1, momentum=0, nk=160, maxk=1.6, precision=api.FP32)
    ps = api.PowerSpectrum(p)
    ps.keep_forward(True)
    ps.set_particles(parts, w)
    G = np.random.default_rng(1).normal(size=(1, 5, 160))
    ps.power()
    ps.backward(G)                     # warmup
    torch.cuda.synchronize()
    reps = 3
    fwd = bwd = 0.0
    reused = True
    for _ in range(reps):
        ps.set_particles(parts, w)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ps.power()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        gx, gw = ps.backward(G)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        reused = reused and ps.backward_reused()
        fwd += (t1 - t0) * 1e3 / reps
        bwd += (t2 - t1) * 1e3 / reps
    ok = bool(torch.isfinite(gx).all().item() and torch.isfinite(gw).all().item())
    # the recomputing backward (particles set again: the forward is not kept)
    t0 = time.perf_counter()
    for _ in range(reps):
        ps.set_particles(parts, w)
        gx2, gw2 = ps.backward(G)
    torch.cuda.synchronize()
    bwd_re = (time.perf_counter() - t0) * 1e3 / reps
    same = bool(torch.equal(gx, gx2) and torch.equal(gw, gw2))
    # per-kernel HIP-event times of one more forward + backward (not in the
    # figures above)
    ps.set_timing(True)
    ps.reset_kernel_times()
    ps.set_particles(parts, w)
    ps.power()
    ps.reset_kernel_times()
    ps.backward(G)
    kt = {k: round(v[0], 4) for k, v in ps.kernel_times().items()}
    ps.set_timing(False)
    # gather: sorted rows and weights, the two gradient grids (density, main
    # and interlaced) with each
    # tile's halo, the index and the gradient rows (algorithmic: the grids
    # once, without the halo)
    gbytes = 2 * n * n * (n + 2) * 4 + parts.shape[0] * (16 + 4 + 4 + 16 + 4)
    if "gather" in kt:
        kt["gather_gbs"] = round(gbytes / (kt["gather"] * 1e-3) / 1e9, 1)
    ps.close()
    del parts, w, gx, gw
    torch.cuda.empty_cache()
    out = dict(workload="16,777,216 Zel'dovich particles, 512^3, TSC+interlacing, P_0,2,4, FP32",
               forward_ms=round(fwd, 3), backward_ms=round(bwd, 3),
               backward_over_forward=round(bwd / fwd, 3),
               forward_includes="sort (recording the rows' input indices), deposit, z/y passes, "
                                "fused x pass + binning",
               backward_includes="the forward's sort and transform reused: x pass, binning "
                                 "backward, inverse FFT, gather",
               backward_reused_forward=bool(reused),
               backward_recompute_ms=round(bwd_re, 3), recompute_gradients_equal=same,
               backward_kernels_ms=kt, gradients_finite=ok)
    log("adjoint: forward %.2f ms, backward %.2f ms (reused %s; recomputing %.2f ms, equal %s)"
        % (fwd, bwd, reused, bwd_re, same))
    return out


def run_survey(args, parts, geom):
    """Survey (Yamamoto) estimator: the workload's Zel'dovich catalogue
    deposited on the 1024^3 grid (TSC + interlacing, density), observer at
    (-L/2, -L/2, -L/2) outside the box, P_0..4 with odd multipoles: 35
    weighted transforms per call (1 + 3 + 6 + 10 + 15), FP32 form."""
    import torch
    from capow_amd import api
    n, L = geom["n"], geom["L"]
    p = api.make_params(nx=n, ny=n, nz=n, xmax=L, ymax=L, zmax=L, interp_order=3,
                        do_interlacing=1, momentum=0, nk=geom["nk"], maxk=geom["nk"] * geom["dk"],
                        odd_multipoles=1, precision=api.FP32)
    ps = api.PowerSpectrum(p)
    ps.set_particles(parts)
    ps.assign()
    origin = [-L / 2, -L / 2, -L / 2]
    ps.survey_power(origin)            # warmup (allocates the survey buffers)
    torch.cuda.synchronize()
    reps = 2
    ps.set_timing(True)
    ps.reset_kernel_times()
    t0 = time.perf_counter()
    for _ in range(reps):
        r = ps.survey_power(origin)
    ms = (time.perf_counter() - t0) * 1e3 / reps
    kt = ps.kernel_times()
    ps.close()
    torch.cuda.empty_cache()
    G, s = 2, 4
    real = n * n * n * G * s                       # real cells of the used grids
    cplx = n * n * (n // 2 + 1) * 2 * s            # one complex half-grid
    kern = {}
    for k, (t, la) in kt.items():
        avg = t / la
        b = None
        if k == "survey_weight":
            b = 2 * real
        elif k == "survey_bin":
            b = (G + 1) * cplx
        elif k.startswith("fft_"):
            b = 2 * G * cplx
        row = dict(ms_per_call=round(t / reps, 3), launches_per_call=la // reps,
                   ms_per_launch=round(avg, 4))
        if b:
            row["alg_bytes"] = b
            row["frac_hbm"] = round(b / (avg * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
        kern[k] = row
    log("survey: %.1f ms per call (35 weighted transforms)" % ms)
    for k, v in sorted(kern.items(), key=lambda kv: -kv[1]["ms_per_call"]):
        log("    %-16s %9.3f ms/call  %9.4f ms/launch  %s" % (k, v["ms_per_call"], v["ms_per_launch"],
                                                          v.get("frac_hbm", "")))
    return dict(workload=("survey estimator: the workload's catalogue on %d^3, TSC+interlacing, "
                          "density, observer (-L/2,-L/2,-L/2), P_0..4 incl. odd, FP32" % n),
                ms_per_call=round(ms, 2), transforms_per_call=35,
                nmodes_total=int(r["nmodes"].sum()), kernels=kern)


def run_survey_ingest(n_data=250000, n_rand=750000, nside=256):
    """Survey ingest on the host (capow_read_survey for a data and a randoms
    file in (RA, Dec, z), the n(z) estimate, FKP weights, the sums, the
    packing) and the survey deposit on the GPU (data and -alpha randoms,
    TSC + interlacing, FP32) of a seeded full-sky catalogue, z in
    [0.05, 0.6], Omega_m 0.31, on a nside^3 grid of side 4000 Mpc/h."""
    import ctypes as C
    import tempfile
    import torch
    from capow_amd import api
    L = api.lib()
    rng = np.random.default_rng(2024)
    tmp = tempfile.mkdtemp(prefix="capow_survey_")
    paths = []
    for n, ncol in ((n_data, 12), (n_rand, 8)):
        cols = np.zeros((n, ncol))
        cols[:, 0] = rng.uniform(0, 360, n)
        cols[:, 1] = np.degrees(np.arcsin(rng.uniform(-1, 1, n)))
        cols[:, 2] = rng.uniform(0.05, 0.6, n)
        cols[:, 3] = rng.uniform(0.5, 1.5, n)
        cols[:, 4] = 1e-4
        cols[:, 6] = rng.normal(0, 0.05, n)
        cols[:, 7] = rng.uniform(0.01, 0.1, n)
        path = os.path.join(tmp, "cat%d.txt" % ncol)
        with open(path, "w") as f:
            f.write("\n".join(" ".join("%.9g" % v for v in row) for row in cols.tolist()) + "\n")
        paths.append(path)
    mb = sum(os.path.getsize(q) for q in paths) / 1e6
    g = api.make_params(nx=nside, ny=nside, nz=nside, xmin=-2000.0, xmax=2000.0, ymin=-2000.0,
                        ymax=2000.0, zmin=-2000.0, zmax=2000.0, interp_order=3, do_interlacing=1,
                        momentum=0, nk=40, maxk=0.2, omega_m=0.31, precision=api.FP32,
                        edge_mode=api.EDGE_SURVEY)
    sp = api.SurveyParams(coord_type=1, momentum=0, nbar_column=-1, fkp_column=-1, omega_m=0.31,
                          redmin_in=0.05, redmax_in=0.6, sky_area=41253.0, fkp_pk=1e4,
                          fkp_pk_mom=1e9, fkp_pk_cross=1e6)
    t = {}
    t0 = time.perf_counter()
    cosmo = C.c_void_p()
    assert L.capow_cosmo_create(0.31, C.byref(cosmo)) == 0
    t["zr_table_ms"] = (time.perf_counter() - t0) * 1e3
    zr = np.array([1e30, -1e30])
    objs = []
    for k, path in enumerate(paths):
        rows, nk_, nr_ = C.POINTER(api.SurveyObj)(), C.c_int64(), C.c_int64()
        t0 = time.perf_counter()
        rc = L.capow_read_survey(path.encode(), k, C.byref(g), C.byref(sp), cosmo, 10 ** 8,
                                 C.byref(rows), C.byref(nk_), C.byref(nr_),
                                 zr.ctypes.data_as(C.POINTER(C.c_double)), None, None, 0, None)
        t["read_%s_ms" % ("randoms" if k else "data")] = (time.perf_counter() - t0) * 1e3
        assert rc == 0, rc
        objs.append((rows, nk_.value))
    (d, nd), (r, nr) = objs
    nb = L.capow_survey_nbar_bins(C.byref(sp))
    zn, nbar = np.zeros(nb + 2), np.zeros(nb + 2)
    dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
    t0 = time.perf_counter()
    assert L.capow_survey_nbar(C.byref(sp), cosmo, dp(zr), d, nd, r, nr, dp(zn), dp(nbar), nb + 2) == 0
    t["nbar_ms"] = (time.perf_counter() - t0) * 1e3
    t0 = time.perf_counter()
    assert L.capow_survey_fkp(C.byref(sp), d, nd) == 0 and L.capow_survey_fkp(C.byref(sp), r, nr) == 0
    tot = api.SurveyTotals()
    assert L.capow_survey_sums(C.byref(sp), d, nd, r, nr, C.byref(tot)) == 0
    xyzv = np.zeros((nd + nr, 4), dtype=np.float32)
    w = np.zeros(nd + nr, dtype=np.float32)
    npk = C.c_int64()
    assert L.capow_survey_pack(C.byref(sp), d, nd, r, nr, tot.alpha, api.FP32, xyzv.ctypes.data,
                               w.ctypes.data, C.byref(npk)) == 0
    t["fkp_sums_pack_ms"] = (time.perf_counter() - t0) * 1e3
    host_ms = sum(t.values())
    ps = api.PowerSpectrum(g)
    xg = torch.from_numpy(xyzv).cuda()
    wg = torch.from_numpy(w).cuda()
    ps.set_particles(xg, wg)
    ps.assign()                                          # warm-up
    torch.cuda.synchronize()
    reps = 5
    t0 = time.perf_counter()
    for _ in range(reps):
        ps.set_particles(xg, wg)
        ps.assign()
    torch.cuda.synchronize()
    dep_ms = (time.perf_counter() - t0) * 1e3 / reps
    ps.close()
    L.capow_free_survey(d)
    L.capow_free_survey(r)
    L.capow_cosmo_destroy(cosmo)
    for q in paths:
        os.remove(q)
    os.rmdir(tmp)
    res = {k: round(v, 2) for k, v in t.items()}
    # the readers and the n(z) step use OMP_NUM_THREADS, else the CPUs this
    # process may run on, at most 16 (the same rule as the library)
    env = os.environ.get("OMP_NUM_THREADS", "")
    nthreads = int(env) if env.isdigit() and int(env) > 0 else len(os.sched_getaffinity(0))
    nthreads = min(nthreads, 16)
    res.update(workload=("survey ingest: %d data + %d randoms, full sky (RA, Dec, z), z 0.05-0.6, "
                         "%.0f MB of text, up to %d host threads; deposit of data and -alpha "
                         "randoms on %d^3, TSC+interlacing, FP32" % (nd, nr, mb, nthreads, nside)),
               host_threads=nthreads,
               host_ms=round(host_ms, 1), host_mb_per_s=round(mb / (host_ms * 1e-3), 1),
               deposit_ms=round(dep_ms, 3), alpha=tot.alpha)
    log("survey ingest: host %.0f ms (%.0f MB/s), deposit %.2f ms" % (host_ms, res["host_mb_per_s"],
                                                                      dep_ms))
    return res


def run_bin_variants(parts, geom):
    """Binning of the workload's grids (1024^3, TSC + interlacing, P00/P11/P01)
    with log k bins (OutputLog, 100 bins from 10^-2.3 to 10^0.2) and with
    (k, mu) bins (Output2D, NMU = 10 and 4), FP32 form: the binning kernels'
    HIP-event times of one timed step after one warmup step."""
    import torch
    from capow_amd import api
    n, L = geom["n"], geom["L"]
    out = {}
    for name, kw in (("log_k", dict(output_log=1, mink=-2.3, maxk=0.2, nk=100)),
                     ("k_mu_nmu10", dict(output_2d=1, nmu=10)),
                     ("k_mu_nmu4", dict(output_2d=1, nmu=4))):
        base = dict(nx=n, ny=n, nz=n, xmax=L, ymax=L, zmax=L, interp_order=3, do_interlacing=1,
                    los=0, momentum=1, nk=geom["nk"], maxk=geom["nk"] * geom["dk"],
                    precision=api.FP32, extra_spectra=api.SPEC_AUTO_A | api.SPEC_AUTO_B)
        base.update(kw)
        ps = api.PowerSpectrum(api.make_params(**base))
        ps.set_particles(parts)
        ps.power()
        ps.set_timing(True)
        ps.reset_kernel_times()
        ps.power()
        kt = ps.kernel_times()
        ps.close()
        torch.cuda.empty_cache()
        bins = {k: round(v[0] / v[1], 3) for k, v in kt.items() if k.startswith("bin")}
        out[name] = dict(kernels_ms=bins, bin_total_ms=round(sum(v[0] for k, v in kt.items()
                                                                  if k.startswith("bin")), 3))
        log("binning %s: %s" % (name, bins))
    return out


def profile_pass(parts, geom):
    """One FP32 step of the workload, then a hipMemcpy of COPY_BYTES (the
    counter calibration), for rocprofv3 --pmc runs."""
    import torch
    from capow_amd import api
    nx = geom["n"]
    p = api.make_params(nx=nx, ny=nx, nz=nx, xmax=geom["L"], ymax=geom["L"], zmax=geom["L"],
                        interp_order=3, do_interlacing=1, los=0, momentum=1, nk=geom["nk"],
                        maxk=geom["nk"] * geom["dk"], precision=api.FP32,
                        extra_spectra=api.SPEC_AUTO_A | api.SPEC_AUTO_B)
    ps = api.PowerSpectrum(p)
    ps.set_particles(parts)
    ps.power()
    torch.cuda.synchronize()
    ps.close()
    src = torch.ones(COPY_BYTES // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    dst.copy_(src)
    torch.cuda.synchronize()
    log("profile pass done")


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed counter summary
    (FETCH_SIZE + WRITE_SIZE, calibrated), or None."""
    path = os.path.join(ROOT, "profiles", TRAFFIC_FILE)
    try:
        text = open(path).read()
    except OSError:
        return None
    m = re.search(r"```json\s*(\{.*?\})\s*```", text, re.S)
    if not m:
        return None
    return json.loads(m.group(1)).get("traffic_bytes", {}).get(kernel)


def cpu_baseline(geom_small, nl_small):
    """The FP64 oracle on a bounded sample of the same workload (about 10 s
    of one core): a 256^3 grid and a 200^3 Zel'dovich lattice (8,000,000
    particles, same box and spectrum), TSC + interlacing, Momentum = 1
    grids, three spectra.  Stage times are scaled to 1024^3 / 134M
    particles: assignment by the particle ratio, FFT by N log N, binning by
    the mode ratio."""
    import ctypes as C
    from capow_amd import mocks
    from oracle import pyoracle as O
    L = geom_small["L"]
    n = geom_small["n"]
    pos, vel = mocks.zeldovich(nl_small, L, 1)
    pos = pos.astype(np.float32).astype(np.float64)
    vel = vel.astype(np.float32).astype(np.float64)
    p = O.params(nx=n, ny=n, nz=n, xmax=L, ymax=L, zmax=L, interp_order=3, do_interlacing=1,
                 momentum=1, nk=geom_small["nk"], maxk=geom_small["nk"] * geom_small["dk"])
    t0 = time.perf_counter()
    rc, grids, st = O.assign(p, pos, vel)
    alpha, shot, norm = O.scalars(p, st)
    O.mean_subtract(p, alpha, grids)
    t1 = time.perf_counter()
    for g in grids:
        O.fft_r2c_3d(g, n, n, n)
    t2 = time.perf_counter()
    for code in (1, 0, 2):
        p.momentum = code
        # the oracle bins one spectrum per pass, as the reference does
        O.compute_power(p, grids if code == 1 else [grids[0], grids[1], None, None]
                        if code == 0 else [grids[2], grids[3], None, None])
    t3 = time.perf_counter()
    scale_p = (512 ** 3) / float(nl_small ** 3)   # 134,217,728 / sample particles
    scale_g = (1024 / n) ** 3
    fft_scale = scale_g * (np.log2(1024 ** 3) / np.log2(n ** 3))
    est = (t1 - t0) * scale_p + (t2 - t1) * fft_scale + (t3 - t2) * scale_g
    del C
    return dict(value=est * 1e3, unit="ms", cores=1, kind="port",
                sample=("FP64 oracle, 1 thread: %d^3 grid, %d particles (Zel'dovich %d^3 "
                        "lattice, L=2000), TSC+interlace, Momentum=1 grids, P01+P00+P11; "
                        "measured assign %.2f s, 4 FFTs %.2f s, 3 binnings %.2f s; scaled to "
                        "1024^3/134M (x%.0f particles, x%.0f FFT, x%.0f modes)" %
                        (n, nl_small ** 3, nl_small, t1 - t0, t2 - t1, t3 - t2, scale_p,
                         fft_scale, scale_g)))


def host_cpu():
    """The host CPU model and the CPUs this process may run on."""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return dict(model=model, nproc=os.cpu_count(), affinity=len(os.sched_getaffinity(0)))


C0_N, C0_L, C0_NP, C0_NK, C0_SEED = 256, 1000.0, 1_000_000, 80, 20


def c0_params(mod, prec=None):
    kw = dict(nx=C0_N, ny=C0_N, nz=C0_N, xmax=C0_L, ymax=C0_L, zmax=C0_L, interp_order=2,
              do_interlacing=0, los=0, momentum=0, nk=C0_NK, maxk=C0_NK * 0.01)
    if prec is not None:
        kw["precision"] = prec
    return mod.make_params(**kw) if hasattr(mod, "make_params") else mod.params(**kw)


def run_c0_gpu(steps=20):
    """The reference's CPU-runnable configuration on the GPU: 1,000,000
    uniform particles, 256^3, CIC, P_0, dk = 0.01, FP64 and FP32 forms;
    ms per step (inputs resident)."""
    import torch
    from capow_amd import api, mocks
    pos, vel = mocks.uniform(C0_NP, C0_L, seed=C0_SEED)
    out = {}
    for name, prec, dt in (("fp64", api.FP64, torch.float64), ("fp32", api.FP32, torch.float32)):
        x = torch.tensor(mocks.xyzv(pos, vel, 0), device="cuda", dtype=dt)
        ps = api.PowerSpectrum(c0_params(api, prec))
        ps.set_particles(x)
        for _ in range(3):
            ps.power()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            ps.power()
        torch.cuda.synchronize()
        out[name + "_ms"] = round((time.perf_counter() - t0) * 1e3 / steps, 3)
        ps.close()
    return out


def cpu_c0():
    """The FP64 oracle (one thread) at the reference's CPU configuration
    exactly: stage times of assignment + mean subtraction, the r2c FFT and
    the binning."""
    from capow_amd import mocks
    from oracle import pyoracle as O
    pos, vel = mocks.uniform(C0_NP, C0_L, seed=C0_SEED)
    p = c0_params(O)
    t0 = time.perf_counter()
    rc, grids, st = O.assign(p, pos, vel)
    alpha, shot, norm = O.scalars(p, st)
    O.mean_subtract(p, alpha, grids)
    t1 = time.perf_counter()
    O.fft_r2c_3d(grids[0], C0_N, C0_N, C0_N)
    t2 = time.perf_counter()
    O.compute_power(p, grids)
    t3 = time.perf_counter()
    return dict(assign_ms=round((t1 - t0) * 1e3, 1), fft_ms=round((t2 - t1) * 1e3, 1),
                bin_ms=round((t3 - t2) * 1e3, 1), total_ms=round((t3 - t0) * 1e3, 1))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--fp64-steps", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--grid", type=int, default=1024, help="grid size (1024 = the workload)")
    ap.add_argument("--lattice", type=int, default=512)
    ap.add_argument("--mocks", type=int, default=512, help="batched mocks (0: skip)")
    ap.add_argument("--mock-streams", type=int, default=8)
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default settings (no capow_autotune before the timed steps)")
    ap.add_argument("--no-survey", action="store_true")
    ap.add_argument("--batch-only", action="store_true",
                    help="run the batched mocks alone and print their figures (not the bench line)")
    ap.add_argument("--no-other-mode", action="store_true",
                    help="batched mocks: time the chosen mode only (not the other one beside it)")
    ap.add_argument("--no-bin-variants", action="store_true")
    ap.add_argument("--profile-pass", action="store_true",
                    help="for rocprofv3 counter runs: one FP32 step and a %d-byte calibration "
                         "copy, no timing" % COPY_BYTES)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: the periodic P_l(k) path runs on a single MI355X "
              "(multi-GPU sharding is a design note, not built); run with --gpus 1.",
              file=sys.stderr)
        sys.exit(2)
    import torch
    from capow_amd import api, mocks
    api.lib()
    if args.batch_only:
        # (for profiling the batched mocks alone: prints their figures only)
        print(json.dumps({"batched_mocks": run_batch(args)}), flush=True)
        return
    L = 2000.0
    n = args.grid
    dk = 0.005
    nk = int(np.floor(np.pi * n / L / dk))
    geom = dict(n=n, L=L, dk=dk, nk=nk)
    log("generating Zel'dovich catalogue %d^3 on the GPU" % args.lattice)
    t0 = time.perf_counter()
    parts32 = mocks.zeldovich_torch(args.lattice, L, seed=1, device="cuda", dtype=torch.float32)
    torch.cuda.synchronize()
    n_p = parts32.shape[0]
    log("catalogue: %d particles in %.1f s" % (n_p, time.perf_counter() - t0))

    if args.profile_pass:
        profile_pass(parts32, geom)
        return

    log("FP32 form")
    f32 = run_form(api.FP32, parts32, n_p, args, geom)
    log("  %.2f ms/step" % f32["ms"])
    for k, (ms, la) in sorted(f32["kernels"].items(), key=lambda kv: -kv[1][0]):
        log("    %-16s %9.3f ms/launch  (%d launches)" % (k, ms / la, la))
    f64 = None
    if not args.no_fp64:
        parts64 = parts32.to(torch.float64)
        log("FP64 plain form")
        f64 = run_form(api.FP64, parts64, n_p, args, geom)
        log("  %.2f ms/step" % f64["ms"])
        for k, (ms, la) in sorted(f64["kernels"].items(), key=lambda kv: -kv[1][0]):
            log("    %-16s %9.3f ms/launch  (%d launches)" % (k, ms / la, la))
        del parts64
        torch.cuda.empty_cache()

    # roofline of the dominant kernel (FP32 form)
    ng = n * n * 2 * (n // 2 + 1)
    G = 4
    kt = f32["kernels"]
    rows = {}
    for name, (ms, launches) in kt.items():
        b = algorithmic_bytes(name, n_p, ng, G, 4)
        if b is None:
            rows[name] = dict(ms=ms, launches=launches, ms_per_launch=ms / launches)
            continue
        avg = ms / launches
        rows[name] = dict(ms=ms, launches=launches, ms_per_launch=avg, alg_bytes=b,
                          gbs=b / (avg * 1e-3) / 1e9)
    dom = max((k for k in rows if "alg_bytes" in rows[k]), key=lambda k: rows[k]["ms"])
    d = rows[dom]
    avg_ms = d["ms"] / d["launches"]
    tb = measured_traffic(dom)
    roofline = dict(bound="hbm", kernel=dom, achieved=round(d["gbs"], 1), peak=HBM_PEAK_GBS,
                    unit="GB/s", frac=round(d["gbs"] / HBM_PEAK_GBS, 4),
                    traffic=(round(tb) if tb else None),
                    avg_launch_ms=round(avg_ms, 4), alg_bytes_per_launch=d["alg_bytes"])
    assign_ms = sum(kt[k][0] for k in ("bucket_count", "stats_final", "scan", "bucket_scatter",
                                        "deposit") if k in kt) / f32["steps"]
    assign_bytes = n_p * 16 + G * ng * 4

    survey = None if args.no_survey else run_survey(args, parts32, geom)
    survey_ingest = None if args.no_survey else run_survey_ingest()
    bin_variants = None if args.no_bin_variants else run_bin_variants(parts32, geom)
    del parts32
    torch.cuda.empty_cache()
    batch = run_batch(args) if args.mocks > 0 else None
    adjoint = None if args.no_adjoint else run_adjoint(args)

    cpu = None
    c0 = run_c0_gpu()
    log("C0 on the GPU: FP64 %.3f ms, FP32 %.3f ms per step" % (c0["fp64_ms"], c0["fp32_ms"]))
    if not args.no_cpu_baseline:
        log("CPU baseline (oracle sample)")
        small = dict(n=256, L=L, dk=dk, nk=int(np.floor(np.pi * 256 / L / dk)))
        cpu = cpu_baseline(small, 200)
        log("  est %.0f ms" % cpu["value"])
        oc0 = cpu_c0()
        log("  C0 measured: %.0f ms (assign %.0f, FFT %.0f, binning %.0f)" %
            (oc0["total_ms"], oc0["assign_ms"], oc0["fft_ms"], oc0["bin_ms"]))
        cpu["value_is"] = ("an estimate for the workload: the sample's measured stage times "
                           "scaled to 1024^3 / 134M particles")
        cpu["host"] = host_cpu()
        cpu["c0_measured"] = dict(
            config=("C0: 1,000,000 uniform particles (seed %d), L=1000, 256^3, CIC, no "
                    "interlacing, P_0, dk=0.01, NK=80, FP64" % C0_SEED),
            oracle_ms=oc0["total_ms"], oracle_stages_ms=oc0, cores=1,
            gpu_fp64_ms=c0["fp64_ms"], gpu_fp32_ms=c0["fp32_ms"])

    res = f32["res"]
    out = {
        "metric": METRIC,
        "value": round(f32["ms"], 3),
        "unit": "ms",
        "n_gpus": 1,
        "steps": f32["steps"],
        "warmup": args.warmup,
        "ms_per_step": round(f32["ms"], 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded Zel'dovich catalogue generated on the GPU)",
        "config": {
            "workload": ("C2: %d particles (%d^3 lattice, Zel'dovich), L=%.0f Mpc/h, %d^3 grid, "
                         "TSC+interlacing, density+momentum grids (4), P00/P11/P01, l=0,2,4, "
                         "dk=%.3f, NK=%d" % (n_p, args.lattice, L, n, dk, nk)),
            "particles": n_p, "grid": n, "interp": "TSC", "interlacing": True,
            "spectra": res["codes"], "nk": nk,
        },
        "fp64_ms_per_step": round(f64["ms"], 3) if f64 else None,
        "autotune": {"fp32": f32["autotune"], "fp64": f64["autotune"] if f64 else None},
        "fp32_speedup_vs_fp64": round(f64["ms"] / f32["ms"], 3) if f64 else None,
        "mparticles_per_s": round(n_p / (f32["ms"] * 1e-3) / 1e6, 1),
        "assign_ms": round(assign_ms, 3),
        "assign_frac_hbm": round(assign_bytes / (assign_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        "kernels_fp32": {k: {kk: (round(vv, 4) if isinstance(vv, float) else vv)
                             for kk, vv in v.items()} for k, v in rows.items()},
        "kernels_fp64": ({k: dict(ms=round(v[0], 4), launches=v[1], ms_per_launch=round(v[0] / v[1], 4))
                          for k, v in f64["kernels"].items()} if f64 else None),
        "roofline": roofline,
        "batched_mocks": batch,
        "adjoint": adjoint,
        "survey": survey,
        "survey_ingest": survey_ingest,
        "bin_variants": bin_variants,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
