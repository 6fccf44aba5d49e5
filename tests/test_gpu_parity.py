"""GPU parity: the HIP path through the C ABI against the FP64 CPU oracle.

Bars (stated per test):
  * deposit grids (FP64 form): |gpu - oracle| <= 1e-10 * max|field value|
    per cell (each tap is rounded to an int64 quantum of about 2^-63 x
    copies in the tile x the largest |field value| before exact
    accumulation);
  * FFT (FP64): <= 1e-12 * max|X|; (FP32): <= 2e-5 * max|X|;
  * binning on identical k-space grids: mode counts bit-exact, sums within
    1e-12 of the per-bin Cauchy-Schwarz scale;
  * end to end (FP64): mode counts bit-exact, |dP_l(k)| <= 1e-10 * (|P_0(k)|
    + shot/norm) (cross spectra: geometric mean of the two autos' scales);
  * FP32 form vs FP64 form: <= 1e-4 of the same scale for k < k_Nyq/2.
"""
import sys

import numpy as np
import pytest

from capow_amd import api, mocks

pytestmark = pytest.mark.gpu

L = 1000.0


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def oracle_params(O, p):
    return O.params(nx=p.nx, ny=p.ny, nz=p.nz, xmin=p.xmin, xmax=p.xmax, ymin=p.ymin,
                    ymax=p.ymax, zmin=p.zmin, zmax=p.zmax, interp_order=p.interp_order,
                    do_interlacing=p.do_interlacing, los=p.los, momentum=p.momentum, nk=p.nk,
                    nmu=p.nmu, output_2d=p.output_2d, output_log=p.output_log,
                    odd_multipoles=p.odd_multipoles, mink=p.mink, maxk=p.maxk,
                    omega_m=p.omega_m, redshift=p.redshift)


def edge_particles(n, L, los, rng):
    """Particles placed on the reference's edge rules: near 0, near L, in
    [dx, 1.5 dx) and [0.5 dx, 1.5 dx) bands, plus exact lattice points."""
    pts = []
    for frac in (0.0, 1e-9, 0.3, 0.5, 0.9999, 1.0, 1.2, 1.49, 1.5, 2.0):
        for axis in range(3):
            q = rng.random(3) * L
            q[axis] = frac * (L / n)
            pts.append(q)
            q2 = rng.random(3) * L
            q2[axis] = L - frac * (L / n) - 1e-9 if frac > 0 else L * (1 - 1e-15)
            pts.append(q2)
    pts = np.array(pts)
    pts = np.clip(pts, 0, np.nextafter(L, 0))
    return pts


def make_catalogue(n_grid, npart, los, seed, momentum):
    rng = np.random.default_rng(seed)
    pos, vel = mocks.uniform(npart, L, seed)
    vel = rng.normal(0, 300.0, size=(npart, 3))
    e = edge_particles(n_grid, L, los, rng)
    pos = np.concatenate([pos, e])
    vel = np.concatenate([vel, rng.normal(0, 300.0, size=(len(e), 3))])
    if los == 0:
        pos = np.clip(pos, 0, L)
    return pos, vel


CASES = [
    # name, grid, order, interlace, los, momentum
    ("cic_los0_m0", 32, 2, 0, 0, 0),
    ("tsc_int_los0_m1", 32, 3, 1, 0, 1),
    ("ngp_int_los0_m2", 24, 1, 1, 0, 2),
    ("tsc_int_los3_m1", 32, 3, 1, 3, 1),
    ("cic_int_los1_m3", 20, 2, 1, 1, 3),
    ("ngp_los2_m4", 16, 1, 0, 2, 4),
    ("tsc_los0_m8", 40, 3, 0, 0, 8),
    ("cic_int_los0_m6", 48, 2, 1, 0, 6),
]


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_deposit_matches_oracle(case, oracle):
    name, n, order, inter, los, mom = case
    say("deposit case", name)
    p = api.make_params(nx=n, ny=n, nz=n, interp_order=order, do_interlacing=inter, los=los,
                        momentum=mom, precision=api.FP64, nk=10, maxk=0.3)
    pos, vel = make_catalogue(n, 20000, los, 7 + n, mom)
    op = oracle_params(oracle, p)
    rc, og, st = oracle.assign(op, pos, vel)
    assert rc == 0
    alpha, shot, norm = oracle.scalars(op, st)
    oracle.mean_subtract(op, alpha, og)
    ps = api.PowerSpectrum(p)
    ps.set_particles(mocks.xyzv(pos, vel, los))
    gst = ps.assign()
    assert gst.ngrid == st.ngrid
    assert gst.nread == st.nread and gst.nout == st.nout
    if mom:
        for q in ("nsq", "vr", "vr2", "vr3", "vr4"):
            ref = getattr(st, q)
            scale = np.abs(mocks.xyzv(pos, vel, los)[:, 3] ** int(q[-1] if q[-1].isdigit() else 1)).sum()
            assert abs(getattr(gst, q) - ref) <= 1e-13 * max(1.0, scale), q
    assert gst.alpha == pytest.approx(alpha, rel=1e-15)
    vlos = mocks.xyzv(pos, vel, los)[:, 3]
    fa = {0: 0, 1: 0, 2: 1, 3: 0, 4: 1, 5: 2, 6: 0, 7: 1, 8: 0}[mom]
    fb = {1: 1, 3: 2, 4: 2, 6: 3, 7: 3, 8: 4}.get(mom)
    for slot in range(4):
        if og[slot] is None:
            continue
        e = fa if slot < 2 else fb
        gmax = max(1e-300, np.abs(vlos ** e).max())
        ref = oracle.real_view(op, og[slot])[:n, :, :n]
        got = ps.get_grid(slot).reshape(n, n, -1)[:, :, :n]
        err = np.abs(got - ref).max()
        say("  slot", slot, "max err", err, "scale", gmax)
        assert err <= 1e-10 * gmax, (name, slot, err)
    ps.close()


HALO_CASES = [
    # name, (nx, ny, nz), order, interlace, los, momentum, weights, clustered copies
    ("tsc_int_los0_m1", (64, 64, 64), 3, 1, 0, 1, False, 0),
    ("cic_int_los3_m3", (96, 64, 128), 2, 1, 3, 3, False, 0),
    ("ngp_los0_m2", (64, 48, 64), 1, 0, 0, 2, False, 0),
    ("tsc_los1_m1_w", (64, 32, 96), 3, 0, 1, 1, True, 0),
    ("tsc_int_los2_m6", (48, 64, 64), 3, 1, 2, 6, False, 0),
    ("cic_los0_m0_w", (32, 64, 64), 2, 0, 0, 0, True, 0),
    # one tile with 900 copies of one bank key (beyond the 512 ranks of the
    # reorder table: that tile keeps the sorted order) ...
    ("tsc_int_cluster900", (64, 64, 64), 3, 1, 0, 1, False, 900),
    # ... and one with more copies than the registers hold (reloaded per phase)
    ("tsc_int_cluster3000_w", (64, 64, 64), 3, 1, 0, 3, True, 3000),
    # density alone (no B field: one int32 per halo-row cell), interlaced CIC
    # as in the batched mocks, clustered beyond the registers, and weighted
    ("cic_int_m0_cluster3000", (64, 64, 64), 2, 1, 0, 0, False, 3000),
    ("cic_int_m0_w", (64, 64, 96), 2, 1, 0, 0, True, 0),
]


@pytest.mark.parametrize("case", HALO_CASES, ids=[c[0] for c in HALO_CASES])
def test_fp32_halo_deposit_matches_plain_tiles(case, monkeypatch):
    """FP32 tiles of 16x16x32 with two or more tiles along z take the
    halo-row deposit (and, for an unweighted field w v^0, the packing
    without sign borrow).  Both round the same float products to the same
    integer quanta as the plain tile deposit, whose FP64 twin matches the
    oracle, and integer sums do not depend on the order: the grids must be
    bitwise equal.  Particles sit on every edge rule of every axis, and
    weights (when used) take both signs."""
    name, (nx, ny, nz), order, inter, los, mom, weighted, ncl = case
    say("halo deposit case", name)
    p = api.make_params(nx=nx, ny=ny, nz=nz, xmax=L, ymax=L, zmax=L, interp_order=order,
                        do_interlacing=inter, los=los, momentum=mom, precision=api.FP32,
                        nk=10, maxk=0.3)
    rng = np.random.default_rng(nx + 3 * ny + 7 * nz)
    pos, vel = make_catalogue(nx, 30000, los, nz, mom)
    extra = np.concatenate([edge_particles(ny, L, los, rng), edge_particles(nz, L, los, rng)])
    pos = np.concatenate([pos, extra])
    if ncl:
        # inside one cell of an interior tile: every copy has the same key
        c0 = np.array([20.3, 21.6, 40.2]) * (L / np.array([nx, ny, nz]))
        pos = np.concatenate([pos, c0 + rng.uniform(0, 0.2 * L / nx, size=(ncl, 3))])
    vel = np.concatenate([vel, rng.normal(0, 300.0, size=(len(pos) - len(vel), 3))])
    xyzv = mocks.xyzv(pos, vel, los)
    w = rng.uniform(-0.5, 2.0, size=len(pos)) if weighted else None

    def grids():
        ps = api.PowerSpectrum(p)
        ps.set_particles(xyzv, w)
        ps.assign()
        out = []
        for s in range(4):
            try:
                # real cells only: the deposit does not write the row padding
                out.append(ps.get_grid(s).reshape(nx, ny, -1)[:, :, :nz].copy())
            except RuntimeError:
                out.append(None)
        ps.close()
        return out

    halo = grids()   # bank-balanced copy order, masked rows
    monkeypatch.setenv("CAPOW_DEPOSIT_BANK", "0")
    halo_sorted = grids()
    monkeypatch.delenv("CAPOW_DEPOSIT_BANK")
    monkeypatch.setenv("CAPOW_DEPOSIT_HALO", "0")
    plain = grids()
    monkeypatch.delenv("CAPOW_DEPOSIT_HALO")
    assert any(g is not None for g in halo)
    for other in (halo_sorted, plain):
        for s in range(4):
            assert (halo[s] is None) == (other[s] is None), s
            if halo[s] is not None:
                assert np.array_equal(halo[s].view(np.uint32), other[s].view(np.uint32)), (name, s)


@pytest.mark.parametrize("prec", [api.FP64, api.FP32])
@pytest.mark.parametrize("shape", [(16, 16, 16), (24, 20, 30), (64, 48, 32), (128, 128, 128),
                                   (30, 40, 50)])
def test_fft_matches_oracle(prec, shape, oracle):
    nx, ny, nz = shape
    say("fft case", prec, shape)
    p = api.make_params(nx=nx, ny=ny, nz=nz, interp_order=1, precision=prec)
    ps = api.PowerSpectrum(p)
    rng = np.random.default_rng(nx * 7 + nz)
    rowlen = 2 * (nz // 2 + 1)
    g = np.zeros((nx, ny, rowlen))
    g[:, :, :nz] = rng.standard_normal((nx, ny, nz))
    ps.set_grid(0, g)
    ps.mark_assigned()
    ps.fft()
    got = ps.get_grid(0).astype(np.float64).reshape(nx, ny, rowlen)
    ref = g.ravel().copy()
    oracle.fft_r2c_3d(ref, nx, ny, nz)
    ref = ref.reshape(nx, ny, rowlen)
    scale = np.abs(ref).max()
    err = np.abs(got - ref).max() / scale
    say("  rel err", err)
    assert err <= (1e-12 if prec == api.FP64 else 2e-5)
    ps.close()


def _cs_scale(pk0_auto_a, pk0_auto_b):
    return np.sqrt(np.abs(pk0_auto_a) * np.abs(pk0_auto_b))


BIN_CASES = [
    ("m0_cic", dict(interp_order=2, do_interlacing=0, momentum=0)),
    ("m1_tsc_int_odd_2d", dict(interp_order=3, do_interlacing=1, momentum=1, odd_multipoles=1,
                               output_2d=1, nmu=7, extra_spectra=3)),
    ("m3_los1_log", dict(interp_order=2, do_interlacing=1, momentum=3, los=1, output_log=1,
                         mink=-2.3, maxk=-0.5, nk=17)),
    ("m2_los2_mink", dict(interp_order=1, do_interlacing=0, momentum=2, los=2, mink=0.013,
                          maxk=0.31, nk=23)),
]


@pytest.mark.parametrize("case", BIN_CASES, ids=[c[0] for c in BIN_CASES])
def test_binning_matches_oracle(case, oracle):
    name, kw = case
    say("binning case", name)
    n = 32
    base = dict(nx=n, ny=n, nz=n, nk=30, maxk=0.3, precision=api.FP64)
    base.update(kw)
    p = api.make_params(**base)
    op = oracle_params(oracle, p)
    pos, vel = make_catalogue(n, 30000, p.los, 3, p.momentum)
    rc, og, st = oracle.assign(op, pos, vel)
    assert rc == 0
    alpha, shot, norm = oracle.scalars(op, st)
    oracle.mean_subtract(op, alpha, og)
    for g in og:
        if g is not None:
            oracle.fft_r2c_3d(g, n, n, n)
    ob = oracle.compute_power(op, og)
    ps = api.PowerSpectrum(p)
    for slot in range(4):
        if og[slot] is not None:
            ps.set_grid(slot, oracle.real_view(op, og[slot])[:n])
    ps.mark_transformed()
    res = ps.bin()
    assert np.array_equal(res["nmodes"], ob.nmodes), name
    prim = res[p.momentum]["pk"]
    autos = {}
    # scale: the auto spectra when available, else |pk0|
    for code in res["codes"]:
        autos[code] = res[code]["pk"][0]
    scale = np.abs(prim[0])
    if len(res["codes"]) == 3:
        scale = _cs_scale(autos[res["codes"][1]], autos[res["codes"][2]])
    for l in ((0, 1, 2, 3, 4) if p.odd_multipoles else (0, 2, 4)):
        d = np.abs(prim[l] - ob.pk[l])
        assert np.all(d <= 1e-12 * scale + 1e-300), (name, l, (d / np.maximum(scale, 1e-300)).max())
    if p.output_2d:
        assert np.array_equal(res["nmodes2d"].ravel(), ob.nmodes2d)
        d = np.abs(res[p.momentum]["pk2d"].ravel() - ob.pk2d)
        assert d.max() <= 1e-12 * np.abs(ob.pk2d).max() * 10
    ps.close()


def _finalized(p, res, stats, code):
    shot, norm = api.scalars(p, stats, code)
    pk5, nm, _, _, last = api.finalize(p, shot, norm, res[code]["pk"].ravel(), res["nmodes"])
    return pk5.reshape(5, -1), nm, shot / norm


E2E = [
    ("c0_like_cic_m0", 64, dict(interp_order=2, do_interlacing=0, momentum=0), 100000, "uniform"),
    ("c1_like_tsc_int_m1", 64, dict(interp_order=3, do_interlacing=1, momentum=1,
                                    odd_multipoles=1, extra_spectra=3), 32, "zeldovich"),
    ("tsc_int_los3_m1", 48, dict(interp_order=3, do_interlacing=1, momentum=1, los=3,
                                 extra_spectra=3), 24, "zeldovich"),
    ("c1_like_128_tsc_int_m1", 128, dict(interp_order=3, do_interlacing=1, momentum=1,
                                         odd_multipoles=1, extra_spectra=3), 64, "zeldovich"),
]


def _catalogue(kind, size, seed):
    if kind == "uniform":
        return mocks.uniform(size, L, seed)
    return mocks.zeldovich(size, L, seed)


@pytest.mark.parametrize("case", E2E, ids=[c[0] for c in E2E])
def test_end_to_end_fp64(case, oracle):
    name, n, kw, size, kind = case
    say("e2e case", name)
    base = dict(nx=n, ny=n, nz=n, nk=int(np.pi * n / L / 0.01), maxk=0.01 * int(np.pi * n / L / 0.01),
                precision=api.FP64)
    base.update(kw)
    p = api.make_params(**base)
    pos, vel = _catalogue(kind, size, 11)
    ps = api.PowerSpectrum(p)
    ps.set_particles(mocks.xyzv(pos, vel, p.los))
    res = ps.power()
    scales = {}
    for code in res["codes"]:
        op = oracle_params(oracle, p)
        op.momentum = code
        rc, ob, ost, oa, oshot, onorm = oracle.run(op, pos, vel)
        assert rc == 0
        assert np.array_equal(res["nmodes"], ob.nmodes), (name, code)
        gpk, gnm, gsn = _finalized(p, res, ps.stats, code)
        o5 = np.stack([ob.pk[l] for l in range(5)])
        opk, onm, _, _, _ = api.finalize(p, oshot, onorm, o5.ravel(), ob.nmodes)
        opk = opk.reshape(5, -1)
        assert np.array_equal(gnm, onm)
        if code in (0, 2, 5):
            scales[code] = np.abs(opk[0]) + oshot / onorm
            scale = scales[code]
        else:
            scale = None
        if scale is None:
            # cross: geometric mean of the autos (computed by the oracle)
            sa = []
            for auto in (0, 2):
                op2 = oracle_params(oracle, p)
                op2.momentum = auto
                rc2, ob2, _, _, s2, n2 = oracle.run(op2, pos, vel)
                a5 = np.stack([ob2.pk[l] for l in range(5)])
                apk = api.finalize(p, s2, n2, a5.ravel(), ob2.nmodes)[0].reshape(5, -1)
                sa.append(np.abs(apk[0]) + s2 / n2)
            scale = np.sqrt(sa[0] * sa[1])
        ls = (0, 1, 2, 3, 4) if p.odd_multipoles else (0, 2, 4)
        for l in ls:
            rel = np.abs(gpk[l] - opk[l]) / np.maximum(scale, 1e-300)
            say("  code", code, "l", l, "max rel", rel.max())
            assert rel.max() <= 1e-10, (name, code, l, rel.max())
    ps.close()


@pytest.mark.parametrize("n", [64, 128])
def test_fp32_form_error_table(n):
    """FP32 form vs FP64 form on the same (float32-representable) catalogue:
    every l and k bin below k_Nyq/2 within 1e-4 of the FP64 scale."""
    say("fp32 vs fp64 error table")
    kw = dict(nx=n, ny=n, nz=n, interp_order=3, do_interlacing=1, momentum=1, extra_spectra=3,
              nk=20, maxk=0.2)
    pos, vel = mocks.zeldovich(64, L, 5)
    pos = pos.astype(np.float32).astype(np.float64)
    vel = vel.astype(np.float32).astype(np.float64)
    out = {}
    for prec in (api.FP64, api.FP32):
        p = api.make_params(precision=prec, **kw)
        ps = api.PowerSpectrum(p)
        ps.set_particles(mocks.xyzv(pos, vel, 0))
        res = ps.power()
        out[prec] = {c: _finalized(p, res, ps.stats, c) for c in res["codes"]}
        ps.close()
    kc = (np.arange(20) + 0.5) * 0.01
    knyq = np.pi * n / L
    sel = kc < knyq / 2
    s00 = np.abs(out[api.FP64][0][0][0]) + out[api.FP64][0][2]
    s11 = np.abs(out[api.FP64][2][0][0]) + out[api.FP64][2][2]
    for code in (0, 1, 2):
        scale = s00 if code == 0 else (s11 if code == 2 else np.sqrt(s00 * s11))
        for l in (0, 2, 4):
            a = out[api.FP32][code][0][l]
            b = out[api.FP64][code][0][l]
            rel = np.abs(a - b) / scale
            say("  code", code, "l", l, "max rel (k<kNyq/2)", rel[sel].max())
            assert rel[sel].max() <= 1e-4


MIRROR_CASES = [
    # FP32 rows of 129 modes: two modes per lane, one chunk and a 1-mode tail
    ("fp32_two_per_lane", api.FP32, (64, 48, 256),
     dict(interp_order=3, do_interlacing=1, momentum=1, extra_spectra=3)),
    # 101 modes per row: one 64-mode chunk and a partial chunk of 37
    ("fp32_long_tail", api.FP32, (32, 40, 200), dict(interp_order=2, do_interlacing=1, momentum=0)),
    ("fp64_chunk_and_tail", api.FP64, (48, 64, 160),
     dict(interp_order=3, do_interlacing=1, momentum=2)),
    # odd multipoles with the line of sight along z: mu is the same on all four rows
    ("fp64_odd_los_z", api.FP64, (32, 32, 128),
     dict(interp_order=3, do_interlacing=1, momentum=1, odd_multipoles=1, los=3, extra_spectra=3)),
    # odd multipoles along x: mu changes sign between mirror rows -> row form
    ("fp32_odd_los_x", api.FP32, (32, 32, 64),
     dict(interp_order=2, do_interlacing=1, momentum=1, odd_multipoles=1, los=1)),
    # no interlacing, momentum cross with its autos, odd NX (row form)
    ("fp32_odd_nx", api.FP32, (45, 32, 64), dict(interp_order=1, momentum=3, extra_spectra=3)),
    # log bins (OutputLog): one exact log10 bin per group
    ("fp64_log_bins", api.FP64, (64, 48, 64),
     dict(interp_order=2, do_interlacing=1, momentum=1, extra_spectra=3, output_log=1,
          mink=-2.5, maxk=-0.6, nk=25)),
    ("fp32_log_two_per_lane", api.FP32, (32, 32, 256),
     dict(interp_order=3, do_interlacing=1, momentum=0, output_log=1, mink=-2.5, maxk=-0.9,
          nk=20)),
    # (k, mu) bins with the line of sight along z: per-k bounds, int64 quanta
    ("fp64_2d_los_z", api.FP64, (64, 48, 64),
     dict(interp_order=3, do_interlacing=1, momentum=1, extra_spectra=3, output_2d=1, nmu=7,
          odd_multipoles=1)),
    ("fp32_2d_two_per_lane", api.FP32, (32, 32, 256),
     dict(interp_order=2, do_interlacing=1, momentum=0, output_2d=1, nmu=10)),
    # (k, mu) bins over log k bins, momentum cross spectrum with its autos
    ("fp64_2d_log_bins", api.FP64, (48, 64, 96),
     dict(interp_order=3, do_interlacing=1, momentum=3, extra_spectra=3, output_2d=1, nmu=5,
          output_log=1, mink=-2.4, maxk=-0.7, nk=17, los=3)),
]


@pytest.mark.parametrize("case", MIRROR_CASES, ids=[c[0] for c in MIRROR_CASES])
def test_mirror_group_binning(case, oracle, monkeypatch):
    """Binning of the mirror rows (i, j), (NX-i, j), (i, NY-j), (NX-i, NY-j)
    together, against the oracle's compute_power and against the row form
    (CAPOW_BIN_MIRROR=0) on the same k-space grids: mode counts exact, sums
    within 1e-12 of the bin's scale."""
    name, prec, (nx, ny, nz), kw = case
    say("mirror binning case", name)
    base = dict(nx=nx, ny=ny, nz=nz, nk=40, maxk=0.2, precision=prec)
    base.update(kw)
    p = api.make_params(**base)
    op = oracle_params(oracle, p)
    rng = np.random.default_rng(nx + 3 * ny + 7 * nz)
    has_b = p.momentum not in (0, 2, 5)
    il = bool(p.do_interlacing)
    og = []
    for s, used in enumerate((True, il, has_b, has_b and il)):
        if not used:
            og.append(None)
            continue
        g = oracle.new_grid(op)
        v = oracle.real_view(op, g)
        # float32-representable, so that both precisions see the same values
        v[:nx] = rng.standard_normal(v[:nx].shape).astype(np.float32)
        og.append(g)
    ob = oracle.compute_power(op, og)
    res = {}
    for form in ("mirror", "rows"):
        if form == "rows":
            monkeypatch.setenv("CAPOW_BIN_MIRROR", "0")
        ps = api.PowerSpectrum(p)
        for s in range(4):
            if og[s] is not None:
                ps.set_grid(s, oracle.real_view(op, og[s])[:nx])
        ps.mark_transformed()
        res[form] = ps.bin()
        ps.close()
    monkeypatch.delenv("CAPOW_BIN_MIRROR")
    ls = (0, 1, 2, 3, 4) if p.odd_multipoles else (0, 2, 4)
    for form, r in res.items():
        assert np.array_equal(r["nmodes"], ob.nmodes), (name, form)
        prim = r[p.momentum]["pk"]
        scale = np.abs(prim[0])
        if len(r["codes"]) == 3:
            scale = _cs_scale(r[r["codes"][1]]["pk"][0], r[r["codes"][2]]["pk"][0])
        for l in ls:
            d = np.abs(prim[l] - ob.pk[l])
            assert np.all(d <= 1e-12 * scale + 1e-300), (name, form, l)
    a, b = res["mirror"], res["rows"]
    for code in a["codes"]:
        scale = np.abs(b[code]["pk"][0])
        if code != p.momentum and len(a["codes"]) == 3:
            scale = _cs_scale(b[a["codes"][1]]["pk"][0], b[a["codes"][2]]["pk"][0])
        for l in ls:
            d = np.abs(a[code]["pk"][l] - b[code]["pk"][l])
            assert np.all(d <= 1e-12 * (scale + np.abs(b[code]["pk"][0])) + 1e-300), (name, code, l)
    if p.output_2d:
        for form, r in res.items():
            assert np.array_equal(r["nmodes2d"].ravel(), ob.nmodes2d), (name, form)
            d = np.abs(r[p.momentum]["pk2d"].ravel() - ob.pk2d)
            assert d.max() <= 1e-12 * np.abs(ob.pk2d).max() * 10, (name, form)
        for code in a["codes"]:
            ref2 = b[code]["pk2d"]
            d = np.abs(a[code]["pk2d"] - ref2)
            assert d.max() <= 1e-12 * np.abs(ref2).max() * 10, (name, code)


@pytest.mark.parametrize("prec", [api.FP64, api.FP32])
@pytest.mark.parametrize("order,inter", [(1, 1), (2, 0), (2, 1), (3, 0)])
def test_lattice_power_is_minus_shot(prec, order, inter):
    """The known answer of tests/test_oracle_power.py::
    test_lattice_power_is_minus_shot through the whole GPU path (capow_power
    and the host finalisation): a 2-cell-spaced lattice of particles that no
    edge rule touches has no power below k_L = pi/dx, so P_0 = -V/N there
    (FP64 within 1e-10, FP32 form within 1e-4 of V/N) and P_2 = P_4 = 0;
    the mode counts are bit-exact with the oracle's (checked elsewhere)."""
    say("lattice case", prec, order, inter)
    n, nk, maxk = 32, 20, 0.2
    dx = L / n
    m = n // 2
    q = np.arange(m) * 2 * dx
    X, Y, Z = np.meshgrid(q + 0.3 * dx, q + 0.1 * dx, q + 0.45 * dx, indexing="ij")
    pos = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
    vel = np.zeros_like(pos)
    p = api.make_params(nx=n, ny=n, nz=n, xmax=L, ymax=L, zmax=L, interp_order=order,
                        do_interlacing=inter, momentum=0, nk=nk, maxk=maxk, precision=prec)
    ps = api.PowerSpectrum(p)
    ps.set_particles(mocks.xyzv(pos, vel, 0))
    res = ps.power()
    pk, nm, pshot = _finalized(p, res, ps.stats, 0)
    ps.close()
    assert pshot == pytest.approx(L ** 3 / m ** 3, rel=1e-14)
    edges_hi = (np.arange(nk) + 1) * (maxk / nk)
    sel = (edges_hi <= np.pi / dx) & (np.asarray(nm) > 0)
    assert sel.sum() >= 8
    tol = 1e-10 if prec == api.FP64 else 1e-4
    np.testing.assert_allclose(pk[0][sel], -pshot, rtol=tol)
    assert np.abs(pk[2][sel]).max() <= tol * pshot
    assert np.abs(pk[4][sel]).max() <= tol * pshot


def test_determinism():
    """Same input twice: grids and bins bitwise identical (FP32 form)."""
    say("determinism")
    n = 64
    p = api.make_params(nx=n, ny=n, nz=n, interp_order=3, do_interlacing=1, momentum=1,
                        extra_spectra=3, precision=api.FP32, nk=20, maxk=0.2)
    pos, vel = mocks.zeldovich(48, L, 9)
    outs = []
    for rep in range(2):
        ps = api.PowerSpectrum(p)
        ps.set_particles(mocks.xyzv(pos, vel, 0))
        ps.assign()
        grids = [ps.get_grid(s).copy() for s in range(4)]
        ps.fft()
        res = ps.bin()
        outs.append((grids, res))
        ps.close()
    for s in range(4):
        # real cells only: the row padding is not part of the real-space grid
        a = outs[0][0][s].reshape(n, n, -1)[:, :, :n]
        b = outs[1][0][s].reshape(n, n, -1)[:, :, :n]
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    for code in outs[0][1]["codes"]:
        assert np.array_equal(outs[0][1][code]["pk"].view(np.uint64),
                              outs[1][1][code]["pk"].view(np.uint64))


def _grids_and_bins(p, xyzv):
    ps = api.PowerSpectrum(p)
    ps.set_particles(xyzv)
    ps.assign()
    n = p.nz
    grids = [ps.get_grid(s).reshape(p.nx, p.ny, -1)[:, :, :n].copy() for s in range(4)]
    ps.fft()
    res = ps.bin()
    ps.close()
    return grids, res


@pytest.mark.parametrize("prec", [api.FP32, api.FP64])
def test_sort_order_and_table_overflow_do_not_change_grids(prec, monkeypatch):
    """The deposit sums exact integers, so the grids and bins do not depend
    on the catalogue order or on which path the counting sort takes: the
    LDS-aggregated path, a shuffled catalogue (no runs of equal keys) and
    the global-counter path (CAPOW_SORT_PROBES=0, as when the table is full)
    give bitwise identical results.  Grid 96^3 spans several tiles per axis
    (ghost copies on every face)."""
    say("sort paths", prec)
    n = 96
    p = api.make_params(nx=n, ny=n, nz=n, interp_order=3, do_interlacing=1, momentum=1,
                        extra_spectra=3, precision=prec, nk=20, maxk=0.2)
    pos, vel = mocks.zeldovich(64, L, 11)
    xyzv = mocks.xyzv(pos, vel, 0)
    perm = np.random.default_rng(3).permutation(len(xyzv))
    base = _grids_and_bins(p, xyzv)
    shuffled = _grids_and_bins(p, xyzv[perm])
    monkeypatch.setenv("CAPOW_SORT_PROBES", "0")
    forced = _grids_and_bins(p, xyzv)
    monkeypatch.delenv("CAPOW_SORT_PROBES")
    u = np.uint32 if prec == api.FP32 else np.uint64
    for other in (shuffled, forced):
        for s in range(4):
            assert np.array_equal(base[0][s].view(u), other[0][s].view(u)), s
        for code in base[1]["codes"]:
            assert np.array_equal(base[1][code]["pk"], other[1][code]["pk"]), code


def test_batch_matches_single_contexts():
    """capow_batch_run (several streams, two grid configs, no host
    synchronisation between mocks) gives bitwise the bins and sums of
    running each catalogue alone through capow_power, including one with
    every particle next to a tile corner (its taps reach four tiles, so the
    corner, edge and face class ranges of the sort carry all of it)."""
    import torch
    say("batch vs single contexts")
    Lb = 500.0
    cfgs = [dict(nx=32, ny=32, nz=32, xmax=Lb, ymax=Lb, zmax=Lb, interp_order=2, do_interlacing=1,
                 momentum=0, nk=20, maxk=0.2, precision=api.FP32),
            dict(nx=48, ny=48, nz=48, xmax=Lb, ymax=Lb, zmax=Lb, interp_order=2, do_interlacing=1,
                 momentum=0, nk=30, maxk=0.3, precision=api.FP32)]
    counts = [3000, 5000, 4000, 6000, 2500, 4500]
    cats = [mocks.mixed_mock_torch(n, Lb, seed=100 + i) for i, n in enumerate(counts)]
    # 6000 particles in the last cell of a tile in x and y (cell 15.5 of 32^3,
    # tiles of 16 cells; z is one tile): deposited by 4 tiles each
    corner = torch.full((6000, 4), 15.5 * Lb / 32, device="cuda", dtype=torch.float32)
    corner[:, 3] = 0
    cats.append(corner)
    which = [0, 1, 0, 1, 0, 1, 0]
    b = api.Batch(cfgs, max_particles=6000, nstreams=3)
    res, stats, ms = b.run(cats, which)
    b.close()
    assert len(res) == len(cats) and np.all(ms > 0)
    for i, (cat, q) in enumerate(zip(cats, which)):
        ps = api.PowerSpectrum(api.make_params(**cfgs[q]))
        ps.set_particles(cat)
        ref = ps.power()
        st = ps.stats
        ps.close()
        assert stats[i].ngrid == st.ngrid and stats[i].nread == st.nread, i
        for code in ref["codes"]:
            assert np.array_equal(res[i][code]["pk"], ref[code]["pk"]), (i, code)
        assert np.array_equal(res[i]["nmodes"], ref["nmodes"]), i


def test_fp32_quanta_follow_each_tiles_own_values():
    """One particle with a line-of-sight velocity 10^5 times the others'
    coarsens the FP32 rounding only in the tiles that deposit it: each tile
    takes its quantum from its own largest |w v|.  Away from the outlier the
    FP32 momentum grid agrees with the FP64 one within 1e-4 of the largest
    |w v| there (the FP32 tap weights of float32 positions alone leave about
    1e-5; a quantum from the global maximum would leave errors of order
    1e-3)."""
    say("per-tile quanta with an outlier")
    n = 128
    rng = np.random.default_rng(17)
    pos = rng.uniform(0, L, size=(60000, 3)).astype(np.float32).astype(np.float64)
    vel = rng.normal(0, 300.0, size=(60000, 3)).astype(np.float32).astype(np.float64)
    pos[0] = (0.1 * L, 0.1 * L, 0.1 * L)
    vel[0, 2] = 3.0e7
    grids = {}
    for prec in (api.FP32, api.FP64):
        p = api.make_params(nx=n, ny=n, nz=n, interp_order=3, do_interlacing=1, momentum=1,
                            precision=prec, nk=10, maxk=0.2)
        ps = api.PowerSpectrum(p)
        xyzv = mocks.xyzv(pos, vel, 0)
        ps.set_particles(xyzv.astype(np.float32) if prec == api.FP32 else xyzv)
        ps.assign()
        grids[prec] = ps.get_grid(2).astype(np.float64).reshape(n, n, -1)[:, :, :n]
        ps.close()
    # the outlier sits in cell 12 of the first 16-cell tile along x; its taps
    # stay in that tile, whose neighbours are x tiles 1 and 7: compare x
    # planes 32 .. 111, whose tiles never see it
    far = np.zeros((n, n, n), bool)
    far[32:112] = True
    b64 = grids[api.FP64]
    scale = np.abs(b64[far]).max()
    err = np.abs(grids[api.FP32][far] - b64[far]).max() / scale
    say("  max |B32 - B64| / max |B| away from the outlier: %.2e" % err)
    assert scale < 1e5 and err <= 1e-4
