"""Pins the oracle's compute_periodic_power restatement (reference
src/main.c:302-468) and the output normalisation (src/main.c:1014-1056).

* Mode counts: the half-space NM weights must reproduce a brute-force count
  over the FULL k-space lattice plus exactly the reference's documented
  excess (DC, the kz = NZ/2 plane and the (NX/2, j, 0) line carry NM = 2
  although each is a single full-space mode).
* Plane waves: a single Fourier mode lands in one bin with the closed-form
  power 2 (A N^3/2)^2 / sinc^(2p), mu = 1 along the line of sight gives
  L2 = L4 = 1, mu = 0 gives L2 = -1/2, L4 = 3/8.
* Interlacing: a half-cell-shifted copy of the same wave combined with the
  e^{i k.h} phase reproduces the un-interlaced power.
"""
import numpy as np
import pytest


def ftab(n, L):
    """Frequencies the way the reference forms them: float casts
    (src/main.c:329-352)."""
    d = L / n
    s = 1.0 / (np.float64(np.float32(n)) * d)
    i = np.arange(n)
    f = np.where(i <= n // 2, i.astype(np.float32).astype(np.float64) * s,
                 (i.astype(np.float32) - np.float32(n)).astype(np.float64) * s)
    return f


def full_space_counts(n, L, nk, mink, maxk):
    f = ftab(n, L)
    fx, fy, fz = np.meshgrid(f, f, f, indexing="ij")
    fk = 2.0 * np.pi * np.sqrt(fx * fx + fy * fy + fz * fz)
    bw = (maxk - mink) / nk
    kb = np.trunc((fk - mink).astype(np.float32).astype(np.float64) / bw).astype(np.int64)
    ok = (kb >= 0) & (kb < nk) & (fk <= np.pi / (L / n))
    full = np.bincount(kb[ok], minlength=nk)[:nk]
    # excess of the reference's NM rule over one count per full-space mode
    kz = np.arange(n)
    ex = np.zeros_like(ok)
    ex[:, :, n // 2] = True                 # kz = NZ/2 plane: NM = 2, one mode
    ex[n // 2, :, 0] = True                 # (NX/2, j, 0): NM = 2
    ex[0, 0, 0] = True                      # DC: NM = 2
    extra = np.bincount(kb[ok & ex], minlength=nk)[:nk]
    del kz
    return full + extra


@pytest.mark.parametrize("n,nk,mink,maxk", [(16, 12, 0.0, 0.06), (24, 30, 0.0, 0.3),
                                            (20, 7, 0.013, 0.11), (32, 50, 0.0, 0.11)])
def test_mode_counts_full_space(n, nk, mink, maxk, oracle):
    L = 1000.0 * n / 32
    p = oracle.params(nx=n, ny=n, nz=n, xmax=L, ymax=L, zmax=L, nk=nk, mink=mink, maxk=maxk)
    g = oracle.new_grid(p)
    b = oracle.compute_power(p, [g, None, None, None])
    exp = full_space_counts(n, L, nk, mink, maxk)
    assert np.array_equal(b.nmodes, exp)


def _wave_grid(p, n, m, shift=0.0, amp=1.0):
    """cos(2 pi m.(x - shift d)/n) on the padded grid of p."""
    import itertools
    g = np.zeros((n + p.interp_order, n, 2 * (n // 2 + 1)))
    i = np.arange(n)
    X, Y, Z = np.meshgrid(i, i, i, indexing="ij")
    ph = 2 * np.pi * (m[0] * (X - shift) + m[1] * (Y - shift) + m[2] * (Z - shift)) / n
    g[:n, :, :n] = amp * np.cos(ph)
    del itertools
    return g.ravel()


@pytest.mark.parametrize("m,mu_exp", [((0, 0, 3), 1.0), ((3, 0, 0), 0.0), ((0, 3, 0), 0.0),
                                      ((2, 2, 1), 1.0 / 3.0)])
@pytest.mark.parametrize("order", [1, 2, 3])
def test_plane_wave_power(m, mu_exp, order, oracle):
    n, L, A = 16, 16.0, 0.7
    p = oracle.params(nx=n, ny=n, nz=n, xmax=L, ymax=L, zmax=L, nk=15, maxk=np.pi * 1.0,
                      interp_order=order)
    g = _wave_grid(p, n, m, amp=A)
    oracle.fft_r2c_3d(g, n, n, n)
    b = oracle.compute_power(p, [g, None, None, None])
    kmag = 2 * np.pi * np.sqrt(sum(c * c for c in m)) / L
    kb = int(kmag / (np.pi / 15))   # bins chosen so no k lies on an edge
    sinc = 1.0
    for c in m:
        a = np.pi * c / n
        if c:
            sinc *= np.sin(a) / a
    expected = 2 * (A * n ** 3 / 2) ** 2 / sinc ** (2 * order)
    assert b.pk[0][kb] == pytest.approx(expected, rel=1e-12)
    others = np.delete(b.pk[0], kb)
    assert np.abs(others).max() < 1e-12 * expected
    L2 = 1.5 * mu_exp ** 2 - 0.5
    L4 = (35 * mu_exp ** 4 - 30 * mu_exp ** 2 + 3) / 8
    assert b.pk[2][kb] == pytest.approx(L2 * expected, rel=1e-12, abs=1e-9 * expected)
    assert b.pk[4][kb] == pytest.approx(L4 * expected, rel=1e-12, abs=1e-9 * expected)


@pytest.mark.parametrize("m", [(1, 2, 3), (5, 0, 2), (0, 3, 4), (3, 0, 0)])
def test_interlace_phase_combines_shifted_copy(m, oracle):
    """The interlaced grid samples the field half a cell lower (a particle at
    x deposits at x + d/2); combining it with e^{i k.h} must double the
    amplitude, which grid_cor *= 0.25 undoes."""
    n, L = 16, 16.0
    kw = dict(nx=n, ny=n, nz=n, xmax=L, ymax=L, zmax=L, nk=15, maxk=np.pi, interp_order=2)
    p0 = oracle.params(do_interlacing=0, **kw)
    p1 = oracle.params(do_interlacing=1, **kw)
    g = _wave_grid(p0, n, m)
    gi = _wave_grid(p1, n, m, shift=0.5)
    g1 = g.copy()
    oracle.fft_r2c_3d(g, n, n, n)
    oracle.fft_r2c_3d(g1, n, n, n)
    oracle.fft_r2c_3d(gi, n, n, n)
    b0 = oracle.compute_power(p0, [g, None, None, None])
    b1 = oracle.compute_power(p1, [g1, gi, None, None])
    assert np.abs(b0.pk[0]).max() > 0
    np.testing.assert_allclose(b1.pk[0], b0.pk[0], rtol=1e-12, atol=1e-9 * np.abs(b0.pk[0]).max())


def test_finalize_known_answers(oracle):
    p = oracle.params(nk=4, odd_multipoles=1, maxk=0.4)
    b = oracle.Bins(p)
    b.nmodes[:] = [2, 4, 6, 0]
    b.pk[0][:] = [10.0, 20.0, 30.0, 0.0]
    b.pk[1][:] = [1.0, 2.0, 3.0, 0.0]
    b.pk[2][:] = [4.0, 8.0, 12.0, 0.0]
    b.pk[3][:] = [1.0, 1.0, 1.0, 0.0]
    b.pk[4][:] = [2.0, 2.0, 2.0, 0.0]
    shot, norm = 3.0, 0.5
    zb = oracle.finalize(p, shot, norm, b)
    # highest filled bin (index 2) is zeroed
    assert zb == 2
    np.testing.assert_allclose(b.pk[0], [(10 - 6) / 1.0, (20 - 12) / 2.0, 0, 0])
    np.testing.assert_allclose(b.pk[2], [4 * 5 / 1.0, 8 * 5 / 2.0, 0, 0])
    np.testing.assert_allclose(b.pk[4], [2 * 9 / 1.0, 2 * 9 / 2.0, 0, 0])
    np.testing.assert_allclose(b.pk[1], [1 * 3 / 1.0, 2 * 3 / 2.0, 0, 0])
    np.testing.assert_allclose(b.pk[3], [1 * 7 / 1.0, 1 * 7 / 2.0, 0, 0])
    assert list(b.nmodes) == [2, 4, 0, 0]


def test_output_text_known_answer(oracle, tmp_path):
    p = oracle.params(nk=3, maxk=0.3)
    b = oracle.Bins(p)
    b.nmodes[:] = [2, 5, 0]
    b.pk[0][:] = [1.5, 2e-7, 0.0]
    b.pk[2][:] = [-0.25, 1234567.0, 0.0]
    b.pk[4][:] = [0.0, 3.0, 0.0]
    path = str(tmp_path / "pk.txt")
    assert oracle.write_output(p, 2.0, 4.0, b, path) == 0
    text = open(path).read()
    assert text == ("# shot = 0.5\n# norm = 4\n# k  pk0  pk2  pk4  Nmodes\n"
                    "0.05 1.5 -0.25 0 2\n0.15 2e-07 1.23457e+06 3 5\n0.25 0 0 0 0\n")


# A periodic lattice of particles, spacing 2 cells, offsets chosen so that no
# edge rule of the reference drops a tap or a particle (the TSC low tap at
# index 0, the interlaced >= MAX rejection): its density is 2-cell periodic,
# so every Fourier mode with all |k_i| below the lattice harmonic
# k_L = pi/dx vanishes, and after the shot-noise subtraction
# (src/main.c:1014-1056: P_0 = sum/(norm Nmodes) - shot/norm, shot = N,
# norm = N^2/V) P_0 = -V/N exactly in every bin below k_L while
# P_2 = P_4 = 0.  Interlaced TSC always meets one of the edge rules and is
# left out.
@pytest.mark.parametrize("order,inter", [(1, 0), (1, 1), (2, 0), (2, 1), (3, 0)])
def test_lattice_power_is_minus_shot(order, inter, oracle):
    n, L, nk, maxk = 32, 1000.0, 20, 0.2
    dx = L / n
    m = n // 2
    q = np.arange(m) * 2 * dx
    X, Y, Z = np.meshgrid(q + 0.3 * dx, q + 0.1 * dx, q + 0.45 * dx, indexing="ij")
    pos = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
    vel = np.zeros_like(pos)
    p = oracle.params(nx=n, ny=n, nz=n, interp_order=order, do_interlacing=inter, nk=nk, maxk=maxk)
    rc, b, st, alpha, shot, norm = oracle.run(p, pos, vel)
    assert rc == 0 and st.ngrid == m ** 3
    raw0 = b.pk[0].copy()
    oracle.finalize(p, shot, norm, b)
    kl = np.pi / dx
    edges_hi = (np.arange(nk) + 1) * (maxk / nk)
    sel = (edges_hi <= kl) & (b.nmodes > 0)
    assert sel.sum() >= 8
    # the raw sums of these bins are FFT round-off of a grid of O(1) values
    assert np.abs(raw0[sel]).max() <= 1e-12 * b.nmodes[sel].max() * shot
    pshot = L ** 3 / m ** 3
    assert shot / norm == pytest.approx(pshot, rel=1e-14)
    np.testing.assert_allclose(b.pk[0][sel], -pshot, rtol=1e-12)
    assert np.abs(b.pk[2][sel]).max() <= 1e-12 * pshot
    assert np.abs(b.pk[4][sel]).max() <= 1e-12 * pshot
    # control: one particle moved by one cell breaks the 2-cell periodicity
    # (for every assignment order) and puts power into every one of these bins
    pos2 = pos.copy()
    pos2[m * m * 3 + 5, 1] += dx
    rc, b2, _, _, shot2, norm2 = oracle.run(p, pos2, vel)
    assert rc == 0
    assert np.all(np.abs(b2.pk[0][sel]) > 1e6 * np.abs(raw0[sel]).max() + 1e-300)
