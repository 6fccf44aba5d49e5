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
