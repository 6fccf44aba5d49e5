"""Pins the FP64 reference gradient (oracle/adjoint_ref.py) -- the checker of
capow_backward -- against central finite differences of the C oracle's own
forward (orc_assign + FFT + orc_compute_power): for L = sum G raw_bins with
random G, dL/dx, dL/dy, dL/dz, dL/dv_los and dL/dw of sampled particles.
TSC and CIC weights are piecewise polynomials, so central differences with
a step of 1e-4 cell agree to O(h^2) away from the (measure-zero) kinks."""
import numpy as np
import pytest

from oracle import adjoint_ref as AR


def _case(O, order, inter, mom, n=16, npart=300, seed=3):
    rng = np.random.default_rng(seed)
    L = 1000.0
    p = O.params(nx=n, ny=n, nz=n, xmax=L, ymax=L, zmax=L, interp_order=order,
                 do_interlacing=inter, momentum=mom, nk=10, maxk=0.1, odd_multipoles=1)
    pos = rng.uniform(0, L, size=(npart, 3))
    vel = rng.normal(0, 100.0, size=(npart, 3))
    w = rng.uniform(0.5, 1.5, size=npart)
    return p, pos, vel, w, rng


@pytest.mark.parametrize("order,inter,mom,extra", [(3, 1, 1, 3), (2, 1, 0, 0), (3, 0, 3, 1)])
def test_reference_gradient_matches_finite_differences(oracle, order, inter, mom, extra):
    p, pos, vel, w, rng = _case(oracle, order, inter, mom)
    ns = len(AR.spectra(mom, extra))
    G = rng.normal(size=(ns, 5, p.nk))

    def loss(pos_, vel_, w_):
        return float((G * AR.raw_bins(p, pos_, vel_, w_, extra)).sum())

    gx, gw = AR.gradient(p, pos, vel, w, G, extra)
    d = 1000.0 / 16
    sample = rng.choice(len(pos), size=8, replace=False)
    errs = []
    scale_x = np.abs(gx[:, :3]).max()
    for i in sample:
        for a in range(3):
            h = 1e-4 * d
            pp, pm = pos.copy(), pos.copy()
            pp[i, a] += h
            pm[i, a] -= h
            fd = (loss(pp, vel, w) - loss(pm, vel, w)) / (2 * h)
            errs.append(abs(fd - gx[i, a]) / scale_x)
        # v_los = vz at LOS = 0
        h = 1e-3
        vp, vm = vel.copy(), vel.copy()
        vp[i, 2] += h
        vm[i, 2] -= h
        fd = (loss(pos, vp, w) - loss(pos, vm, w)) / (2 * h)
        sv = max(np.abs(gx[:, 3]).max(), 1e-300)
        if np.abs(gx[:, 3]).max() > 0:
            errs.append(abs(fd - gx[i, 3]) / sv)
        else:
            assert abs(fd) < 1e-6 * scale_x
        h = 1e-6
        wp, wm = w.copy(), w.copy()
        wp[i] += h
        wm[i] -= h
        fd = (loss(pos, vel, wp) - loss(pos, vel, wm)) / (2 * h)
        errs.append(abs(fd - gw[i]) / np.abs(gw).max())
    assert max(errs) < 1e-5, max(errs)


def test_reference_gradient_pins_a_planted_error(oracle):
    """A dropped factor in the reference gradient (here: no interlaced
    contribution) must fail the finite-difference check above."""
    p, pos, vel, w, rng = _case(oracle, 3, 1, 0)
    G = rng.normal(size=(1, 5, p.nk))
    gx, _ = AR.gradient(p, pos, vel, w, G, 0)
    p0 = oracle.params(**{f: getattr(p, f) for f, _ in p._fields_})
    p0.do_interlacing = 0
    gx0, _ = AR.gradient(p0, pos, vel, w, G, 0)
    i = 5
    h = 1e-4 * 1000.0 / 16
    pp, pm = pos.copy(), pos.copy()
    pp[i, 0] += h
    pm[i, 0] -= h
    fd = ((G * AR.raw_bins(p, pp, vel, w, 0)).sum() - (G * AR.raw_bins(p, pm, vel, w, 0)).sum()) / (2 * h)
    assert abs(fd - gx[i, 0]) < 1e-5 * np.abs(gx[:, :3]).max()
    assert abs(fd - gx0[i, 0]) > 1e-3 * np.abs(gx[:, :3]).max()
