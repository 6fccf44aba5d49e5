"""The benchmark configurations at their own sizes, through the C ABI.

* C0 exactly: 1,000,000 uniform float64 particles in a 1000 Mpc/h box,
  256^3, CIC, no interlacing, P_0 only, dk = 0.01 (NK = 80), FP64: mode
  counts bit-exact with the oracle, |dP_l| <= 1e-10 (|P_0| + shot/norm).
* C1: 16,777,216 Zel'dovich particles (256^3 lattice, L = 1000), 512^3, TSC
  + interlacing, P00, P11 and P01 with l = 0..4, NK = 160: the FP64 form
  against the oracle's bins (a fixture written by
  tests/golden/make_c1_oracle.py, the oracle taking minutes at this size)
  within 1e-10, counts bit-exact; the FP32 form within 1e-4 for k <
  k_Nyq/2, its per-l, per-bin error table printed (and written to
  $CAPOW_TABLE_DIR/c1_fp32_error_table.md when that is set).
* C3: four mocks (128^3 and 192^3, L = 500, CIC + interlacing, lognormal
  counts in [1e5, 3e6], half Zel'dovich, half uniform) through capow_batch
  (FP32) and single FP64 contexts, against the oracle.
* C4: the loss L = sum_{l,k} (P_l - P_l^target)^2 / sigma^2(k),
  sigma^2 = 2 P_0^2 / N_modes of the target (a second seed), over P00, P11
  and P01.  Its gradient at 512^3 with the C1 catalogue against FP64
  central differences of the GPU FP64 forward on 1,000 random particles
  (one random direction in (x, y, z, v_los, w) each); and capow_backward
  on a 64^3 grid with 262,144 particles (64^3 lattice, L = 250) against the
  FP64 reference gradient.
* The lattice known answer at 512^3 and 1024^3 (134,217,728 particles):
  a 2-cell lattice has no power below k_Nyq, so P_0 = -V/N and P_2 = P_4 =
  0 there; with the reference's edge rules (CIC + interlacing) and with
  periodic edges (TSC + interlacing, the benchmark kernels).
"""
import os
import sys

import numpy as np
import pytest

from capow_amd import api, mocks

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _oracle_params(O, p):
    return O.params(nx=p.nx, ny=p.ny, nz=p.nz, xmin=p.xmin, xmax=p.xmax, ymin=p.ymin,
                    ymax=p.ymax, zmin=p.zmin, zmax=p.zmax, interp_order=p.interp_order,
                    do_interlacing=p.do_interlacing, los=p.los, momentum=p.momentum, nk=p.nk,
                    nmu=p.nmu, output_2d=p.output_2d, output_log=p.output_log,
                    odd_multipoles=p.odd_multipoles, mink=p.mink, maxk=p.maxk,
                    omega_m=p.omega_m, redshift=p.redshift)


def _fin(p, raw5, nmodes, shot, norm):
    pk5, nm, _, _, _ = api.finalize(p, shot, norm, np.asarray(raw5, np.float64).ravel(),
                                    np.asarray(nmodes, np.int64))
    return pk5.reshape(5, -1), nm


# ---------------------------------------------------------------- C0

def test_c0_exact_fp64_against_oracle(oracle):
    say("C0: 1,000,000 uniform particles, 256^3 CIC, FP64")
    L, n, nk = 1000.0, 256, 80
    pos, vel = mocks.uniform(1_000_000, L, seed=20)
    p = api.make_params(nx=n, ny=n, nz=n, xmax=L, ymax=L, zmax=L, interp_order=2,
                        do_interlacing=0, los=0, momentum=0, nk=nk, maxk=nk * 0.01,
                        precision=api.FP64)
    ps = api.PowerSpectrum(p)
    ps.set_particles(mocks.xyzv(pos, vel, 0))
    res = ps.power()
    shot, norm = api.scalars(p, ps.stats, 0)
    ps.close()
    rc, ob, st, alpha, oshot, onorm = oracle.run(_oracle_params(oracle, p), pos, vel)
    assert rc == 0
    assert np.array_equal(res["nmodes"], ob.nmodes)
    assert shot == oshot and norm == pytest.approx(onorm, rel=1e-15)
    gpk, gnm = _fin(p, res[0]["pk"], res["nmodes"], shot, norm)
    opk, onm = _fin(p, np.stack([ob.pk[l] for l in range(5)]), ob.nmodes, oshot, onorm)
    assert np.array_equal(gnm, onm)
    scale = np.abs(opk[0]) + oshot / onorm
    for l in (0, 2, 4):
        rel = (np.abs(gpk[l] - opk[l]) / scale).max()
        say("  l", l, "max rel", rel)
        assert rel <= 1e-10


# ---------------------------------------------------------------- C1

C1_NL, C1_L, C1_N, C1_NK, C1_DK = 256, 1000.0, 512, 160, 0.01


def _c1_params(prec, momentum=1):
    return api.make_params(nx=C1_N, ny=C1_N, nz=C1_N, xmax=C1_L, ymax=C1_L, zmax=C1_L,
                           interp_order=3, do_interlacing=1, los=0, momentum=momentum,
                           nk=C1_NK, maxk=C1_NK * C1_DK, odd_multipoles=1, precision=prec,
                           extra_spectra=api.SPEC_AUTO_A | api.SPEC_AUTO_B)


@pytest.fixture(scope="module")
def c1_catalogue():
    say("C1: building the 256^3 Zel'dovich catalogue")
    pos, vel = mocks.zeldovich(C1_NL, C1_L, 1)
    return pos, vel


@pytest.mark.timeout(600)
def test_c1_fp64_and_fp32_against_oracle(c1_catalogue):
    ref = np.load(os.path.join(GOLDEN, "c1_oracle.npz"), allow_pickle=False)
    pos, vel = c1_catalogue
    chk = np.array([pos.sum(), (pos * pos).sum(), vel[:, 2].sum(), (vel[:, 2] ** 2).sum()])
    np.testing.assert_allclose(chk, ref["checksums"], rtol=1e-12)
    xyzv = mocks.xyzv(pos, vel, 0)
    fin = {}
    for prec in (api.FP64, api.FP32):
        say("C1: GPU", "fp64" if prec == api.FP64 else "fp32")
        p = _c1_params(prec)
        ps = api.PowerSpectrum(p)
        ps.set_particles(xyzv if prec == api.FP64 else xyzv.astype(np.float32))
        res = ps.power()
        st = ps.stats
        ps.close()
        assert list(res["codes"]) == [1, 0, 2]
        fin[prec] = {}
        for code in (0, 2, 1):
            assert np.array_equal(res["nmodes"], ref["nmodes_%d" % code]), code
            shot, norm = api.scalars(p, st, code)
            s = ref["scalars_%d" % code]
            assert st.nread == s[1]
            # FP32 positions: a particle rounded onto the box edge may be
            # rejected (part of the FP32 form's error)
            assert st.ngrid == s[0] if prec == api.FP64 else abs(st.ngrid - s[0]) <= 1e-6 * s[0]
            if prec == api.FP64:
                # moments summed in another order: the shot agrees to rounding
                # of sum |w v^n| (the P01 shot, sum v, nearly cancels)
                vscale = np.abs(vel[:, 2]).sum() if code == 1 else abs(s[3])
                assert abs(shot - s[3]) <= 1e-12 * vscale
                assert norm == pytest.approx(s[4], rel=1e-15)
            fin[prec][code] = _fin(p, res[code]["pk"], res["nmodes"], shot, norm)[0]
    p = _c1_params(api.FP64)
    o = {}
    for code in (0, 2, 1):
        s = ref["scalars_%d" % code]
        o[code] = _fin(p, ref["raw_%d" % code], ref["nmodes_%d" % code], s[3], s[4])[0]
    sc = {0: np.abs(o[0][0]) + ref["scalars_0"][3] / ref["scalars_0"][4],
          2: np.abs(o[2][0]) + ref["scalars_2"][3] / ref["scalars_2"][4]}
    sc[1] = np.sqrt(sc[0] * sc[2])
    kc = (np.arange(C1_NK) + 0.5) * C1_DK
    sel = kc < (np.pi * C1_N / C1_L) / 2
    lines = ["# C1 FP32 form: |P_l^FP32 - P_l^oracle| / (|P_0| + shot/norm) per k bin",
             "# (cross P01: geometric mean of the P00 and P11 scales); k < k_Nyq/2 = %.3f" %
             ((np.pi * C1_N / C1_L) / 2), "",
             "| spectrum | l | max rel (k < k_Nyq/2) | bin of max | median | FP64 max rel |",
             "|---|---|---|---|---|---|"]
    per_bin = ["", "| k | " + " | ".join("%s l=%d" % (nm, l) for nm in ("P00", "P11", "P01")
                                         for l in (0, 2, 4)) + " |",
               "|---|" + "---|" * 9]
    rows = {}
    for code, name in ((0, "P00"), (2, "P11"), (1, "P01")):
        for l in (0, 1, 2, 3, 4):
            r64 = np.abs(fin[api.FP64][code][l] - o[code][l]) / sc[code]
            r32 = np.abs(fin[api.FP32][code][l] - o[code][l]) / sc[code]
            assert r64.max() <= 1e-10, (name, l, r64.max())
            m = r32[sel]
            lines.append("| %s | %d | %.2e | %d | %.2e | %.2e |" % (name, l, m.max(), int(np.argmax(m)),
                                                                  np.median(m), r64.max()))
            rows[(name, l)] = r32
            say("  %s l=%d fp64 %.2e fp32 %.2e" % (name, l, r64.max(), m.max()))
            assert m.max() <= 1e-4, (name, l, m.max())
    for k in range(C1_NK):
        if not sel[k]:
            break
        per_bin.append("| %.3f | " % kc[k] + " | ".join("%.1e" % rows[(nm, l)][k]
                                                        for nm in ("P00", "P11", "P01")
                                                        for l in (0, 2, 4)) + " |")
    out = os.environ.get("CAPOW_TABLE_DIR")
    if out:
        os.makedirs(out, exist_ok=True)
        with open(os.path.join(out, "c1_fp32_error_table.md"), "w") as f:
            f.write("\n".join(lines + per_bin) + "\n")


# ---------------------------------------------------------------- C3

@pytest.mark.timeout(300)
def test_c3_mocks_against_oracle(oracle):
    import torch
    say("C3: four mocks at 128^3 and 192^3")
    Lb = 500.0
    cfgs = []
    for n, nk in ((128, 80), (192, 120)):
        cfgs.append(dict(nx=n, ny=n, nz=n, xmax=Lb, ymax=Lb, zmax=Lb, interp_order=2,
                         do_interlacing=1, los=0, momentum=0, nk=nk, maxk=nk * 0.01))
    counts = mocks.lognormal_counts(4, seed=33)
    counts[0] = 3_000_000      # the clip's upper end
    which = [0, 1, 0, 1]
    cats = [mocks.mixed_mock_torch(int(c), Lb, seed=500 + i) for i, c in enumerate(counts)]
    b = api.Batch([api.make_params(precision=api.FP32, **c) for c in cfgs],
                  max_particles=int(max(counts)), nstreams=2)
    res32, st32, ms = b.run(cats, which)
    b.close()
    for i, (cat, q) in enumerate(zip(cats, which)):
        host = cat.double().cpu().numpy()
        pos = host[:, :3]
        vel = np.zeros_like(pos)
        vel[:, 2] = host[:, 3]
        p64 = api.make_params(precision=api.FP64, **cfgs[q])
        ps = api.PowerSpectrum(p64)
        ps.set_particles(host)
        r64 = ps.power()
        st64 = ps.stats
        ps.close()
        rc, ob, ost, oa, oshot, onorm = oracle.run(_oracle_params(oracle, p64), pos, vel)
        assert rc == 0
        opk, onm = _fin(p64, np.stack([ob.pk[l] for l in range(5)]), ob.nmodes, oshot, onorm)
        scale = np.abs(opk[0]) + oshot / onorm
        n = cfgs[q]["nx"]
        kc = (np.arange(cfgs[q]["nk"]) + 0.5) * 0.01
        sel = kc < (np.pi * n / Lb) / 2
        for tag, r, st, tol, s_ in (("fp64", r64, st64, 1e-10, slice(None)),
                                    ("fp32", res32[i], st32[i], 1e-4, sel)):
            assert np.array_equal(r["nmodes"], ob.nmodes), (i, tag)
            shot, norm = api.scalars(p64, st, 0)
            gpk, _ = _fin(p64, r[0]["pk"], r["nmodes"], shot, norm)
            for l in (0, 2, 4):
                rel = (np.abs(gpk[l] - opk[l]) / scale)[s_].max()
                say("  mock %d (n=%d, %d^3) %s l=%d max rel %.2e" % (i, counts[i], n, tag, l, rel))
                assert rel <= tol, (i, tag, l, rel)
    del torch


# ---------------------------------------------------------------- C4

def _loss_parts(P, Pt, nmt):
    """sum over P00, P11 (sigma^2 = 2 P_0^2 / N) and P01 (sigma^2 =
    P00_0 P11_0 / N) of (P_l - P_l^t)^2 / sigma^2, l = 0..4, bins with
    modes (target's last filled bin excluded: it is zeroed)."""
    import torch
    ok = (nmt > 0) & (Pt[0][0] != 0)
    n = torch.clamp(nmt, min=1)
    s00 = torch.where(ok, 2 * Pt[0][0] ** 2 / n, torch.ones_like(n))
    s11 = torch.where(ok, 2 * Pt[2][0] ** 2 / n, torch.ones_like(n))
    s01 = torch.where(ok, torch.abs(Pt[0][0] * Pt[2][0]) / n, torch.ones_like(n))
    okf = ok.to(torch.float64)
    tot = 0.0
    for code, sig in ((0, s00), (2, s11), (1, s01)):
        for l in range(5):
            tot = tot + (okf * (P[code][l] - Pt[code][l]) ** 2 / sig).sum()
    return tot


@pytest.mark.timeout(900)
def test_c4_gradient_against_fp64_finite_differences(c1_catalogue):
    import torch
    from capow_amd import autograd as AG
    say("C4: loss gradient at 512^3, 16,777,216 particles, vs FP64 central differences")
    pos, vel = c1_catalogue
    post, velt = mocks.zeldovich(C1_NL, C1_L, 2)
    x64 = torch.tensor(mocks.xyzv(pos, vel, 0), device="cuda", dtype=torch.float64)
    xt = torch.tensor(mocks.xyzv(post, velt, 0), device="cuda", dtype=torch.float64)
    w64 = torch.ones(x64.shape[0], device="cuda", dtype=torch.float64)
    ps64 = api.PowerSpectrum(_c1_params(api.FP64))
    with torch.no_grad():
        Pt = AG.multipoles(ps64, xt, torch.ones_like(w64))
        _, nmt = AG.raw_bins(ps64, xt, torch.ones_like(w64))
    del xt

    def loss(ps, xx, ww):
        return _loss_parts(AG.multipoles(ps, xx, ww), Pt, nmt)

    grads = {}
    for prec, dt in ((api.FP64, torch.float64), (api.FP32, torch.float32)):
        ps = ps64 if prec == api.FP64 else api.PowerSpectrum(_c1_params(api.FP32))
        xr = x64.to(dt).clone().requires_grad_(True)
        wr = w64.to(dt).clone().requires_grad_(True)
        Lv = loss(ps, xr, wr)
        gx, gw = torch.autograd.grad(Lv, [xr, wr])
        grads[prec] = (gx.double(), gw.double())
        say("  %s loss %.6e" % ("fp64" if prec == api.FP64 else "fp32", Lv.item()))
        if ps is not ps64:
            ps.close()
        del xr, wr
    dx = C1_L / C1_N
    h = np.array([1e-4 * dx] * 3 + [1e-4 * 50.0, 1e-4])   # x y z [Mpc/h], v [km/s], w
    rng = np.random.default_rng(44)
    idx = rng.choice(x64.shape[0], size=1000, replace=False)
    u = rng.normal(size=(1000, 5))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    fd = np.zeros(1000)
    with torch.no_grad():
        for t, i in enumerate(idx):
            d = torch.tensor(u[t] * h, device="cuda", dtype=torch.float64)
            row, wi = x64[i].clone(), w64[i].clone()
            x64[i] = row + d[:4]
            w64[i] = wi + d[4]
            lp = loss(ps64, x64, w64).item()
            x64[i] = row - d[:4]
            w64[i] = wi - d[4]
            lm = loss(ps64, x64, w64).item()
            x64[i] = row
            w64[i] = wi
            fd[t] = (lp - lm) / 2
            if t % 200 == 0:
                say("  finite differences: %d of 1000" % t)
    ps64.close()
    for prec, tol in ((api.FP64, 1e-5), (api.FP32, 5e-3)):
        gx, gw = grads[prec]
        g = torch.cat([gx[idx], gw[idx, None]], dim=1).cpu().numpy()
        exp = (g * u * h).sum(axis=1)
        err = np.abs(exp - fd)
        scale = np.abs(fd).max()
        say("  %s backward vs FD: max |err| / max |fd| = %.2e (median %.2e)" %
            ("fp64" if prec == api.FP64 else "fp32", err.max() / scale, np.median(err) / scale))
        assert err.max() <= tol * scale


def test_c4_reference_gradient_64(oracle):
    """capow_backward on a 64^3 grid with 262,144 particles (64^3 Zel'dovich
    lattice, L = 250 Mpc/h, TSC + interlacing, P01 with its autos) against
    the FP64 reference gradient, for dL/draw of the C4 loss."""
    import torch
    from capow_amd import autograd as AG
    from oracle import adjoint_ref as AR
    say("C4: reference gradient at 64^3 with 262,144 particles")
    L, n, nk = 250.0, 64, 32
    kw = dict(nx=n, ny=n, nz=n, xmax=L, ymax=L, zmax=L, interp_order=3, do_interlacing=1,
              los=0, momentum=1, nk=nk, maxk=nk * 0.025, odd_multipoles=1)
    pos, vel = mocks.zeldovich(64, L, 5)
    post, velt = mocks.zeldovich(64, L, 6)
    pos = pos.astype(np.float32).astype(np.float64)
    vel = vel.astype(np.float32).astype(np.float64)
    ps = api.PowerSpectrum(api.make_params(precision=api.FP64, extra_spectra=3, **kw))
    x = torch.tensor(mocks.xyzv(pos, vel, 0), device="cuda", dtype=torch.float64)
    xt = torch.tensor(mocks.xyzv(post, velt, 0), device="cuda", dtype=torch.float64)
    one = torch.ones(x.shape[0], device="cuda", dtype=torch.float64)
    with torch.no_grad():
        Pt = AG.multipoles(ps, xt, one)
        _, nmt = AG.raw_bins(ps, xt, one)
    # dL/draw: the normalisation restated in torch, raw as the leaf
    res = ps.power()
    raw = torch.tensor(np.stack([res[c]["pk"] for c in res["codes"]]), requires_grad=True)
    N = float(x.shape[0])
    norm = N * N / L ** 3
    v = x[:, 3].cpu()
    shots = {0: N, 2: float((v ** 2).sum()), 1: float(v.sum())}
    nm = torch.tensor(res["nmodes"].astype(np.float64))
    filled = nm > 0
    last = int(torch.nonzero(filled).max())
    P = {}
    for s, code in enumerate(res["codes"]):
        nmc = torch.where(filled, nm, torch.ones_like(nm))
        rows = []
        for l in range(5):
            val = (raw[s, l] - nmc * shots[code]) / (nmc * norm) if l == 0 else \
                (2 * l + 1) * raw[s, l] / (nmc * norm)
            keep = torch.ones_like(val)
            keep[last] = 0
            rows.append(torch.where(filled, val, torch.zeros_like(val)) * keep)
        P[code] = torch.stack(rows)
    Lv = _loss_parts(P, Pt, nmt)
    (G,) = torch.autograd.grad(Lv, [raw])
    G = G.numpy()
    ref_x, ref_w = AR.gradient(oracle.params(**kw), pos, vel, None, G, 3)
    for prec, tol in ((api.FP64, 1e-9), (api.FP32, 2e-3)):
        pb = api.PowerSpectrum(api.make_params(precision=prec, extra_spectra=3, **kw))
        xx = mocks.xyzv(pos, vel, 0)
        pb.set_particles(xx if prec == api.FP64 else xx.astype(np.float32))
        gx, gw = pb.backward(G)
        pb.close()
        for c in range(4):
            e = np.abs(gx[:, c] - ref_x[:, c]).max() / np.abs(ref_x[:, c]).max()
            say("  %s component %d max rel %.2e" % ("fp64" if prec == api.FP64 else "fp32", c, e))
            assert e <= tol, (prec, c, e)
        e = np.abs(gw - ref_w).max() / np.abs(ref_w).max()
        assert e <= tol, (prec, "w", e)
    ps.close()


# ---------------------------------------------------------------- lattice

@pytest.mark.timeout(300)
@pytest.mark.parametrize("n", [512, 1024])
@pytest.mark.parametrize("prec", [api.FP32, api.FP64])
@pytest.mark.parametrize("kind", ["cic_int_reference_edges", "tsc_int_periodic_edges"])
def test_lattice_known_answer_full_size(n, prec, kind):
    _lattice_known_answer(n, prec, kind)


@pytest.mark.timeout(300)
@pytest.mark.parametrize("kind", ["cic_int_reference_edges", "tsc_int_periodic_edges"])
def test_lattice_known_answer_2048(kind):
    """The largest grid: 2048^3 with a 1024^3 lattice (1,073,741,824
    particles), FP32 form (its 2048-point x and y passes are the 64 x 32
    register kernels; FP64 takes the Stockham kernels at 2048 and two FP64
    grids alone would hold 138 GB).

    The box (L = 4096, so N/L = 1/2) and the lattice offsets (1/4, 1/8, 3/8
    of a cell) make every position and every FP32 scaled coordinate exact,
    so the FP32 form must give the known answer to rounding.  With
    L = 4000 (N/L not a float) the FP32 form's reciprocal stretches the
    scaled coordinates by up to 2^-24 N cells across the box, and the FFT's
    rounding on the perfectly periodic lattice adds up coherently over its
    million identical rows: P_4 of the fundamental bin then carries 1.5e-4
    V/N (2.7e-6 at 1024^3); the FP64 form is exact either way."""
    _lattice_known_answer(2048, api.FP32, kind, L=4096.0, dk=0.0025, offs=(0.25, 0.125, 0.375),
                          tol=1e-6)


def _lattice_known_answer(n, prec, kind, L=None, dk=None, offs=(0.3, 0.1, 0.45), tol=None):
    import torch
    say("lattice known answer", n, "fp64" if prec == api.FP64 else "fp32", kind)
    L = 2000.0 * n / 1024 if L is None else L
    dk = 0.005 * 1024 / n if dk is None else dk
    nk = int(np.floor(np.pi * n / L / dk))
    order, edge = (2, api.EDGE_REFERENCE) if kind.startswith("cic") else (3, api.EDGE_PERIODIC)
    p = api.make_params(nx=n, ny=n, nz=n, xmax=L, ymax=L, zmax=L, interp_order=order,
                        do_interlacing=1, los=0, momentum=0, nk=nk, maxk=nk * dk,
                        precision=prec, edge_mode=edge)
    dt = torch.float64 if prec == api.FP64 else torch.float32
    dx = L / n
    m = n // 2
    q = torch.arange(m, device="cuda", dtype=torch.float64) * 2 * dx
    # rows written column by column in the working precision (a 2048^3 grid
    # takes 1024^3 particles: no float64 copies of the whole catalogue)
    xyzv = torch.zeros((m ** 3, 4), device="cuda", dtype=dt)
    for a, off in enumerate(offs):
        shape = [1, 1, 1]
        shape[a] = m
        col = (q + off * dx).to(dt).reshape(shape).expand(m, m, m)
        xyzv[:, a] = col.reshape(-1)
        del col
    ps = api.PowerSpectrum(p)
    ps.set_particles(xyzv.contiguous())
    res = ps.power()
    shot, norm = api.scalars(p, ps.stats, 0)
    ps.close()
    del xyzv
    torch.cuda.empty_cache()
    pk, nm = _fin(p, res[0]["pk"], res["nmodes"], shot, norm)
    vn = L ** 3 / m ** 3
    assert shot / norm == pytest.approx(vn, rel=1e-14)
    sel = np.asarray(nm) > 0
    assert sel.sum() >= nk - 2
    if tol is None:
        tol = 1e-10 if prec == api.FP64 else 1e-4
    d0 = np.abs(pk[0][sel] + vn).max() / vn
    d24 = max(np.abs(pk[2][sel]).max(), np.abs(pk[4][sel]).max()) / vn
    say("  max |P0 + V/N| / (V/N) %.2e, max |P2|,|P4| / (V/N) %.2e" % (d0, d24))
    assert d0 <= tol and d24 <= tol
