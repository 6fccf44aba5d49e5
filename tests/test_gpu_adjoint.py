"""GPU adjoint (capow_backward through the C ABI) against the FP64 reference
gradient (oracle/adjoint_ref.py, itself pinned by finite differences of the
C oracle) and against finite differences of the GPU's own FP64 forward.

Bars: FP64 form within 1e-9 of max|gradient| per component; FP32 form
within 2e-3 of max|gradient| (FP32 grids and FFT); finite differences of the
FP64 forward within 1e-6.
"""
import sys

import numpy as np
import pytest

from capow_amd import api, mocks
from oracle import adjoint_ref as AR

pytestmark = pytest.mark.gpu

L = 1000.0


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _catalogue(n, seed):
    rng = np.random.default_rng(seed)
    pos = rng.uniform(0, L, size=(n, 3))
    vel = rng.normal(0, 100.0, size=(n, 3))
    w = rng.uniform(0.5, 1.5, size=n)
    return pos, vel, w


CASES = [
    # name, order, interlace, momentum, extra, odd
    ("tsc_int_m1_x3_odd", 3, 1, 1, 3, 1),
    ("cic_int_m0", 2, 1, 0, 0, 0),
    ("tsc_m3_x1", 3, 0, 3, 1, 0),
    ("ngp_int_m2", 1, 1, 2, 0, 0),
]


@pytest.mark.parametrize("prec", [api.FP64, api.FP32])
@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_backward_matches_reference_gradient(case, prec, oracle):
    name, order, inter, mom, extra, odd = case
    say("adjoint", name, "fp64" if prec == api.FP64 else "fp32")
    n = 32
    pos, vel, w = _catalogue(3000, 11)
    if prec == api.FP32:
        pos = pos.astype(np.float32).astype(np.float64)
        vel = vel.astype(np.float32).astype(np.float64)
        w = w.astype(np.float32).astype(np.float64)
    kw = dict(nx=n, ny=n, nz=n, xmax=L, ymax=L, zmax=L, interp_order=order, do_interlacing=inter,
              momentum=mom, nk=12, maxk=0.12, odd_multipoles=odd)
    op = oracle.params(**kw)
    ns = len(AR.spectra(mom, extra))
    G = np.random.default_rng(5).normal(size=(ns, 5, 12))
    ref_x, ref_w = AR.gradient(op, pos, vel, w, G, extra)
    ps = api.PowerSpectrum(api.make_params(precision=prec, extra_spectra=extra, **kw))
    ps.set_particles(mocks.xyzv(pos, vel, 0), w)
    gx, gw = ps.backward(G)
    ps.close()
    tol = 1e-9 if prec == api.FP64 else 2e-3
    for c in range(4):
        s = np.abs(ref_x[:, c]).max()
        if s == 0:
            assert np.abs(gx[:, c]).max() == 0, c
            continue
        err = np.abs(gx[:, c] - ref_x[:, c]).max() / s
        say("  component", c, "max rel err", err)
        assert err <= tol, (c, err)
    err = np.abs(gw - ref_w).max() / np.abs(ref_w).max()
    say("  weights max rel err", err)
    assert err <= tol


def test_backward_matches_gpu_finite_differences():
    """FP64 GPU forward, central differences on 24 random particles of a
    Zel'dovich catalogue (64^3 grid, TSC + interlacing, density and
    momentum): dL/dx_i for L = sum G raw."""
    say("adjoint vs GPU finite differences")
    n = 64
    pos, vel = mocks.zeldovich(32, L, 21)
    xyzv = mocks.xyzv(pos, vel, 0)
    kw = dict(nx=n, ny=n, nz=n, xmax=L, ymax=L, zmax=L, interp_order=3, do_interlacing=1,
              momentum=1, nk=20, maxk=0.2, precision=api.FP64, extra_spectra=3)
    ps = api.PowerSpectrum(api.make_params(**kw))
    G = np.random.default_rng(8).normal(size=(3, 5, 20))

    def loss(x):
        ps.set_particles(x)
        r = ps.power()
        return float(sum((G[s] * r[c]["pk"]).sum() for s, c in enumerate(r["codes"])))

    ps.set_particles(xyzv)
    gx, _ = ps.backward(G)
    rng = np.random.default_rng(4)
    idx = rng.choice(len(xyzv), size=24, replace=False)
    scale = np.abs(gx[:, :3]).max()
    h = 1e-4 * L / n
    for i in idx:
        for a in range(3):
            xp, xm = xyzv.copy(), xyzv.copy()
            xp[i, a] += h
            xm[i, a] -= h
            fd = (loss(xp) - loss(xm)) / (2 * h)
            assert abs(fd - gx[i, a]) <= 1e-6 * scale, (i, a, fd, gx[i, a])
    ps.close()


def test_autograd_loss_gradient():
    """torch.autograd over the C ABI with the host normalisation in torch:
    L = sum_{l,k} (P_l - P_l^target)^2 / sigma^2, sigma^2 = 2 P_0^2 / N_modes
    of the target; dL/dx and dL/dw against central differences of L (FP64)."""
    import torch
    from capow_amd import autograd as AG
    say("autograd loss")
    n = 32
    kw = dict(nx=n, ny=n, nz=n, xmax=L, ymax=L, zmax=L, interp_order=3, do_interlacing=1,
              momentum=0, nk=12, maxk=0.12, precision=api.FP64)
    ps = api.PowerSpectrum(api.make_params(**kw))
    pos, vel = mocks.zeldovich(24, L, 31)
    post, velt = mocks.zeldovich(24, L, 32)
    x = torch.tensor(mocks.xyzv(pos, vel, 0), device="cuda", dtype=torch.float64)
    xt = torch.tensor(mocks.xyzv(post, velt, 0), device="cuda", dtype=torch.float64)
    w0 = torch.ones(x.shape[0], device="cuda", dtype=torch.float64)
    with torch.no_grad():
        Pt = AG.multipoles(ps, xt, torch.ones_like(w0))[0]
        _, nmt = AG.raw_bins(ps, xt, torch.ones_like(w0))
    sig2 = torch.where(nmt > 0, 2 * Pt[0] ** 2 / torch.clamp(nmt, min=1), torch.ones_like(nmt))
    sig2 = torch.where(sig2 > 0, sig2, torch.ones_like(sig2))

    def loss_of(xx, ww):
        P = AG.multipoles(ps, xx, ww)[0]
        return (((P - Pt)[[0, 2, 4]] ** 2) / sig2).sum()

    xr = x.clone().requires_grad_(True)
    wr = w0.clone().requires_grad_(True)
    Lv = loss_of(xr, wr)
    gx, gw = torch.autograd.grad(Lv, [xr, wr])
    rng = np.random.default_rng(2)
    h = 1e-4 * L / n
    sx = gx[:, :3].abs().max().item()
    sw = gw.abs().max().item()
    for i in rng.choice(x.shape[0], size=6, replace=False):
        for a in range(3):
            xp, xm = x.clone(), x.clone()
            xp[i, a] += h
            xm[i, a] -= h
            fd = (loss_of(xp, w0) - loss_of(xm, w0)).item() / (2 * h)
            assert abs(fd - gx[i, a].item()) <= 1e-5 * sx, (i, a, fd, gx[i, a].item())
        hw = 1e-6
        wp, wm = w0.clone(), w0.clone()
        wp[i] += hw
        wm[i] -= hw
        fd = (loss_of(x, wp) - loss_of(x, wm)).item() / (2 * hw)
        assert abs(fd - gw[i].item()) <= 1e-5 * sw, (i, fd, gw[i].item())
    ps.close()


TILE_CASES = [
    # name, (nx, ny, nz), order, interlace, los, momentum, extra, weights, precision
    # (register FFT lengths only: NX, NY, NZ/2 in 16, 32, 64, 96, ...)
    ("tsc_int_los0_m1_fp64", (64, 96, 128), 3, 1, 0, 1, 3, False, api.FP64),
    ("tsc_int_los0_m1_fp32", (64, 64, 64), 3, 1, 0, 1, 3, False, api.FP32),
    ("cic_int_los3_m3_w_fp32", (96, 64, 128), 2, 1, 3, 3, 1, True, api.FP32),
    ("ngp_los1_m0_w_fp64", (32, 64, 64), 1, 0, 1, 0, 0, True, api.FP64),
    ("tsc_los2_m2_fp32", (32, 64, 192), 3, 0, 2, 2, 0, False, api.FP32),
]


@pytest.mark.parametrize("case", TILE_CASES, ids=[c[0] for c in TILE_CASES])
def test_tile_gather_matches_particle_gather(case, monkeypatch):
    """The tile gather (the forward's sort tiles, gradient regions staged in
    LDS, rows written back at the input index) against the per-particle
    gather (CAPOW_GATHER_TILES=0) on the same catalogue: the same sums in
    another order, so within 1e-12 (FP64) or 2e-5 (FP32) of max|gradient|
    per component.  Shuffled input order, particles on every edge rule and
    outside the box (line of sight != 0: dropped by the sort, zero gradient
    in both), signed weights, and one tile with 3000 particles (more than
    one pass over its region)."""
    from tests.test_gpu_parity import edge_particles
    name, (nx, ny, nz), order, inter, los, mom, extra, weighted, prec = case
    say("tile gather", name)
    Lb = 1000.0
    rng = np.random.default_rng(nx + ny + nz + order)
    pos = rng.uniform(-0.02 * Lb, 1.02 * Lb, size=(40000, 3)) if los else rng.uniform(0, Lb, size=(40000, 3))
    pos = np.concatenate([pos, edge_particles(nx, Lb, los, rng), edge_particles(nz, Lb, los, rng)])
    c0 = np.array([20.3, 21.6, 40.2]) * (Lb / np.array([nx, ny, nz]))
    pos = np.concatenate([pos, c0 + rng.uniform(0, 0.3 * Lb / nx, size=(3000, 3))])
    if los == 0:
        pos = np.clip(pos, 0, Lb)
    vel = rng.normal(0, 200.0, size=pos.shape)
    perm = rng.permutation(len(pos))
    pos, vel = pos[perm], vel[perm]
    w = rng.uniform(-0.5, 2.0, size=len(pos)) if weighted else None
    kw = dict(nx=nx, ny=ny, nz=nz, xmax=Lb, ymax=Lb, zmax=Lb, interp_order=order,
              do_interlacing=inter, los=los, momentum=mom, nk=16, maxk=0.16,
              extra_spectra=extra, precision=prec)
    p = api.make_params(**kw)
    ns = len(AR.spectra(mom, extra))
    G = np.random.default_rng(6).normal(size=(ns, 5, 16))
    xyzv = mocks.xyzv(pos, vel, los)

    def grads():
        ps = api.PowerSpectrum(p)
        ps.set_particles(xyzv, w)
        gx, gw = ps.backward(G)
        ps.close()
        return gx, gw

    tiles = grads()
    monkeypatch.setenv("CAPOW_GATHER_TILES", "0")
    plain = grads()
    monkeypatch.delenv("CAPOW_GATHER_TILES")
    tol = 1e-12 if prec == api.FP64 else 2e-5
    for c in range(4):
        s = np.abs(plain[0][:, c]).max()
        assert np.isfinite(tiles[0][:, c]).all()
        if s == 0:
            assert np.abs(tiles[0][:, c]).max() == 0, c
            continue
        err = np.abs(tiles[0][:, c] - plain[0][:, c]).max() / s
        say("  component", c, "max rel diff", err)
        assert err <= tol, (c, err)
    if weighted or mom:
        s = np.abs(plain[1]).max()
        err = np.abs(tiles[1] - plain[1]).max() / s
        say("  weights max rel diff", err)
        assert err <= tol
    # particles the sort drops have zero gradients in both
    dead = np.all(plain[0] == 0, axis=1)
    assert np.all(tiles[0][dead] == 0)


@pytest.mark.parametrize("case", [TILE_CASES[1], TILE_CASES[2], TILE_CASES[0]],
                         ids=[TILE_CASES[1][0], TILE_CASES[2][0], TILE_CASES[0][0]])
def test_kept_forward_matches_recompute(case):
    """capow_keep_forward: a backward right after the forward (capow_power)
    reuses its sort and transform; the gradients must be bitwise those of
    the backward that recomputes them (the same deposit sums, transform and
    per-particle tap sums), the forward's bins unchanged by the recorded
    indices, and anything that changes the particles or grids in between
    must make the backward recompute."""
    name, (nx, ny, nz), order, inter, los, mom, extra, weighted, prec = case
    say("kept forward", name)
    Lb = 1000.0
    rng = np.random.default_rng(11)
    pos = rng.uniform(0, Lb, size=(30000, 3))
    vel = rng.normal(0, 200.0, size=pos.shape)
    w = rng.uniform(-0.5, 2.0, size=len(pos)) if weighted else None
    p = api.make_params(nx=nx, ny=ny, nz=nz, xmax=Lb, ymax=Lb, zmax=Lb, interp_order=order,
                        do_interlacing=inter, los=los, momentum=mom, nk=16, maxk=0.16,
                        extra_spectra=extra, precision=prec)
    ns = len(AR.spectra(mom, extra))
    G = np.random.default_rng(6).normal(size=(ns, 5, 16))
    xyzv = mocks.xyzv(pos, vel, los)

    ref = api.PowerSpectrum(p)
    ref.set_particles(xyzv, w)
    fwd_ref = ref.power()
    ref.set_particles(xyzv, w)
    gref = ref.backward(G)
    assert not ref.backward_reused()

    ps = api.PowerSpectrum(p)
    ps.keep_forward(True)
    ps.set_particles(xyzv, w)
    fwd = ps.power()
    for code in fwd_ref["codes"]:
        assert np.array_equal(fwd[code]["pk"], fwd_ref[code]["pk"]), code
    g = ps.backward(G)
    assert ps.backward_reused()
    assert np.array_equal(g[0], gref[0])
    if gref[1] is not None:
        assert np.array_equal(g[1], gref[1])
    # the backward leaves gradients in the grids: the next one recomputes
    g2 = ps.backward(G)
    assert not ps.backward_reused()
    assert np.array_equal(g2[0], gref[0])
    # a forward, then new particles: recompute, on the new particles
    ps.power()
    pos2 = rng.uniform(0, Lb, size=(20000, 3))
    xyzv2 = mocks.xyzv(pos2, rng.normal(0, 200.0, size=pos2.shape), los)
    w2 = rng.uniform(-0.5, 2.0, size=len(pos2)) if weighted else None
    ps.set_particles(xyzv2, w2)
    g3 = ps.backward(G)
    assert not ps.backward_reused()
    ref.set_particles(xyzv2, w2)
    g3r = ref.backward(G)
    assert np.array_equal(g3[0], g3r[0])
    ref.close()
    ps.close()
