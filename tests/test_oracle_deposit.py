"""Pins the oracle's add_to_grid restatement (reference
src/read_data.c:514-745 at one MPI rank) with known answers.

Grid: 16^3 cells of size 1 (box [0, 16)), so scaled coordinates equal
positions.  Expected values are worked out by hand from the textbook
weights (NGP 1; CIC 1-f, f; TSC 0.5(0.5-f)^2, 0.75-f^2, 0.5(0.5+f)^2) and
the reference's edge rules:
  * TSC drops a low-neighbour tap whose index is 0 (guards "> 0",
    src/read_data.c:692-741);
  * LOS = 0 wraps (buffer slices onto 0.., src/read_data.c:152-175;
    y/z wrap, src/read_data.c:656-671); LOS != 0 drops taps outside;
  * interlaced (x + d/2) CIC/NGP reject shifted coordinates >= MAX.
"""
import numpy as np
import pytest

N = 16


def grid_after_halo(O, p, g):
    """Real cells after the NTask = 1 halo add (orc_assign does it; for
    single orc_add_to_grid calls we apply the same wrap here for LOS = 0)."""
    v = O.real_view(p, g).copy()
    if p.los == 0:
        for s in range(p.interp_order):
            v[s] += v[N + s]
    return v[:N, :, :N]


def deposit(O, order, x, y, z, los=0, w=1.0, vr=1.0):
    p = O.params(nx=N, ny=N, nz=N, xmax=16.0, ymax=16.0, zmax=16.0, interp_order=order, los=los)
    g = O.new_grid(p)
    acc = O.add_to_grid(p, x, y, z, w, vr, g)
    return acc, grid_after_halo(O, p, g)


def test_tsc_interior_weights(oracle):
    acc, g = deposit(oracle, 3, 5.3, 7.6, 9.5)
    assert acc == 1.0
    # x: c=5 f=0.3 -> 4:0.02 5:0.66 6:0.32 ; y: c=8 f=-0.4 -> 7:0.405 8:0.59 9:0.005
    # z: s=9.5 -> c=10 f=-0.5 -> 9:0.5 10:0.5 11:0
    wx = {4: 0.02, 5: 0.66, 6: 0.32}
    wy = {7: 0.405, 8: 0.59, 9: 0.005}
    wz = {9: 0.5, 10: 0.5}
    exp = np.zeros_like(g)
    for i, a in wx.items():
        for j, b in wy.items():
            for k, c in wz.items():
                exp[i, j, k] = a * b * c
    assert np.abs(g - exp).max() < 1e-15
    assert g.sum() == pytest.approx(1.0, abs=1e-14)


def test_cic_and_ngp_interior(oracle):
    acc, g = deposit(oracle, 2, 3.25, 4.5, 10.0)
    exp = np.zeros_like(g)
    for i, a in {3: 0.75, 4: 0.25}.items():
        for j, b in {4: 0.5, 5: 0.5}.items():
            exp[i, j, 10] = a * b
    assert np.abs(g - exp).max() < 1e-15
    acc, g = deposit(oracle, 1, 3.99, 4.0, 15.9999)
    assert g[3, 4, 15] == 1.0 and g.sum() == 1.0


def test_tsc_drops_low_tap_at_index_zero_x(oracle):
    # x = 1.2: centre 1, low tap would be cell 0 -> dropped (weight 0.045)
    acc, g = deposit(oracle, 3, 1.2, 8.0, 8.0)
    assert acc == 1.0
    assert g[0].sum() == 0.0
    assert g.sum() == pytest.approx(1.0 - 0.5 * 0.3 ** 2, abs=1e-14)


def test_tsc_drops_low_tap_at_index_zero_y(oracle):
    # y = 0.7: centre 1, f = -0.3, low tap (cell 0, weight 0.32) dropped
    acc, g = deposit(oracle, 3, 8.0, 0.7, 8.0)
    assert g[:, 0, :].sum() == 0.0
    assert g.sum() == pytest.approx(0.68, abs=1e-14)


def test_tsc_periodic_wrap_low_y_keeps_low_tap(oracle):
    # y = 0.3: centre 0, the low tap (-1) wraps to cell 15 and is kept
    # (iyneighlow < 0 -> NY-1, then the "> 0" guard passes)
    acc, g = deposit(oracle, 3, 8.0, 0.3, 8.0)
    np.testing.assert_allclose([g[:, 15].sum(), g[:, 0].sum(), g[:, 1].sum()],
                               [0.02, 0.66, 0.32], atol=1e-14)
    acc, g = deposit(oracle, 3, 8.0, 8.0, 0.3)
    np.testing.assert_allclose([g[:, :, 15].sum(), g[:, :, 0].sum(), g[:, :, 1].sum()],
                               [0.02, 0.66, 0.32], atol=1e-14)
    # the same position without the wrap (LOS != 0) drops it
    acc, g = deposit(oracle, 3, 8.0, 0.3, 8.0, los=2)
    assert g.sum() == pytest.approx(0.98, abs=1e-14)


def test_tsc_periodic_wrap_low_x(oracle):
    # x = 0.3 < dx: shifted by L to 16.3 -> centre 16 == cell 0, low 15, high 1
    acc, g = deposit(oracle, 3, 0.3, 8.0, 8.0)
    assert acc == 1.0
    assert g.sum() == pytest.approx(1.0, abs=1e-14)
    np.testing.assert_allclose([g[15].sum(), g[0].sum(), g[1].sum()], [0.02, 0.66, 0.32],
                               atol=1e-14)


def test_tsc_periodic_wrap_high_y(oracle):
    # y = 15.8 -> centre 16 -> cells 15, 0, 1 with f = -0.2
    acc, g = deposit(oracle, 3, 8.0, 15.8, 8.0)
    np.testing.assert_allclose([g[:, 15].sum(), g[:, 0].sum(), g[:, 1].sum()],
                               [0.245, 0.71, 0.045], atol=1e-14)


def test_tsc_nonperiodic_los3_drops_outside(oracle):
    # LOS != 0: y = 15.8 keeps only the low tap (cell 15): 0.245
    acc, g = deposit(oracle, 3, 8.0, 15.8, 8.0, los=3)
    assert acc == 1.0
    assert g.sum() == pytest.approx(0.245, abs=1e-14)
    assert g[:, 15].sum() == pytest.approx(0.245, abs=1e-14)


def test_cic_wrap_and_interlace_reject(oracle):
    acc, g = deposit(oracle, 2, 15.6, 8.0, 8.0)
    np.testing.assert_allclose([g[15].sum(), g[0].sum()], [0.4, 0.6], atol=1e-14)
    # the interlaced call passes x + dx/2 = 16.1 >= XMAX: rejected
    acc, g = deposit(oracle, 2, 16.1, 8.5, 8.5)
    assert acc == 0.0 and g.sum() == 0.0


def test_tsc_accepts_x_equal_max_los0(oracle):
    # x = XMAX is inside [xmin+dx, xmax+dx) -> centre 16 == cell 0
    acc, g = deposit(oracle, 3, 16.0, 8.0, 8.0)
    assert acc == 1.0 and g.sum() == pytest.approx(1.0)


def test_mass_conservation_and_moments(oracle):
    """Away from the edges every order conserves sum(w*vr); the read loop's
    moments equal numpy sums."""
    rng = np.random.default_rng(4)
    n = 5000
    pos = rng.uniform(3.0, 12.0, size=(n, 3))
    vel = rng.normal(0, 10, size=(n, 3))
    w = rng.uniform(0.5, 2.0, size=n)
    for order in (1, 2, 3):
        p = oracle.params(nx=N, ny=N, nz=N, xmax=16.0, ymax=16.0, zmax=16.0,
                          interp_order=order, momentum=1, do_interlacing=1)
        rc, grids, st = oracle.assign(p, pos, vel, w)
        assert rc == 0
        assert grids[0].sum() == pytest.approx(w.sum(), rel=1e-12)
        assert grids[1].sum() == pytest.approx(w.sum(), rel=1e-12)
        assert grids[2].sum() == pytest.approx((w * vel[:, 2]).sum(), rel=1e-10)
        assert st.ngrid == pytest.approx(w.sum(), rel=1e-14)
        assert st.nsq == pytest.approx(w.sum(), rel=1e-14)
        for k, name in enumerate(("vr", "vr2", "vr3", "vr4"), start=1):
            assert getattr(st, name) == pytest.approx((w * vel[:, 2] ** k).sum(), rel=1e-12)


def test_rsd_shift_and_nout(oracle):
    """LOS = 3 shifts z by v (1+z) / (100 E(z)); particles pushed outside are
    counted in NOUT (src/read_data.c:67-88)."""
    om, zr = 0.3, 0.5
    E = np.sqrt(om * (1 + zr) ** 3 + 1 - om)
    pos = np.array([[8.0, 8.0, 15.0], [8.0, 8.0, 1.0], [8.0, 8.0, 8.0]])
    vel = np.array([[0.0, 0.0, 300.0], [0.0, 0.0, -300.0], [0.0, 0.0, 10.0]])
    shifts = vel[:, 2] * (1 + zr) / (100 * E)
    p = oracle.params(nx=N, ny=N, nz=N, xmax=16.0, ymax=16.0, zmax=16.0, interp_order=1,
                      los=3, omega_m=om, redshift=zr)
    rc, grids, st = oracle.assign(p, pos, vel)
    assert rc == 0
    znew = pos[:, 2] + shifts
    assert st.nout == np.sum((znew < 0) | (znew > 16))
    assert st.ngrid == np.sum((znew >= 0) & (znew < 16))
    g = oracle.real_view(p, grids[0])
    assert g[8, 8, int(znew[2])] == 1.0


def test_los0_out_of_box_is_fatal(oracle):
    p = oracle.params(nx=N, ny=N, nz=N, xmax=16.0, ymax=16.0, zmax=16.0, interp_order=2)
    rc, _, _ = oracle.assign(p, np.array([[8.0, 16.5, 8.0]]), np.zeros((1, 3)))
    assert rc == 110
