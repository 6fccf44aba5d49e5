"""Pins the oracle's survey (Yamamoto) estimator, a restatement of reference
src/main.c:472-967, against closed forms that follow from its definition and
not from its code.

For point sources the r^-weighted fields are known exactly: with unit
sources at cells a and b, F0(k) = e^{-i phi_a} + e^{-i phi_b} and the field
weighted by a product P of unit-vector components is P(n_a) e^{-i phi_a} +
P(n_b) e^{-i phi_b}.  Summing the reference's index combinations with their
multiplicities rebuilds (k^.n)^l, so with theta = phi_b - phi_a, w = NM cor:
  even l:  Pk_l = sum w (1 + cos theta) [L_l(k^.n_a) + L_l(k^.n_b)]
  odd l:   Pk_l = sum w sin theta [L_l(k^.n_b) - L_l(k^.n_a)]
(the reference's Pk_l carry no 2l+1 factor).  A wrong multiplicity, a
missing or repeated index combination, a swapped sign or a transposed
product changes these sums at generic directions.  The per-mode weights w
are restated here and checked against the periodic oracle's P0 of a single
unit source (|F|^2 = 1).  No GPU needed.
"""
import itertools

import numpy as np
import pytest

from oracle import pyoracle as orc

L = 1000.0


def _legendre(l, mu):
    return {0: np.ones_like(mu), 1: mu, 2: 1.5 * mu**2 - 0.5, 3: 2.5 * mu**3 - 1.5 * mu,
            4: 4.375 * mu**4 - 3.75 * mu**2 + 0.375}[l]


def _modes(p):
    """Binned modes of the half-space grid: (kbin, NM, cor, kvec, integer
    indices), with the reference's float casts (src/main.c:796-830)."""
    NX, NY, NZ = p.nx, p.ny, p.nz
    dx, dy, dz = (p.xmax - p.xmin) / NX, (p.ymax - p.ymin) / NY, (p.zmax - p.zmin) / NZ
    f32 = np.float32
    sx = 1.0 / (float(f32(NX)) * dx)
    sy = 1.0 / (float(f32(NY)) * dy)
    sz = 1.0 / (float(f32(NZ)) * dz)
    i, j, k = np.meshgrid(np.arange(NX), np.arange(NY), np.arange(NZ // 2 + 1), indexing="ij")
    fx = np.where(i <= NX // 2, i.astype(f32).astype(float), (i.astype(f32) - f32(NX)).astype(float)) * sx
    fy = np.where(j <= NY // 2, j.astype(f32).astype(float), (j.astype(f32) - f32(NY)).astype(float)) * sy
    fz = k.astype(f32).astype(float) * sz
    fk = 2.0 * np.pi * np.sqrt(fx * fx + fy * fy + fz * fz)
    bw = (p.maxk - p.mink) / p.nk
    kbin = np.trunc((fk - p.mink).astype(f32).astype(float) / bw).astype(int)
    nyq = min(np.pi / dx, np.pi / dy, np.pi / dz)
    ok = (kbin >= 0) & (kbin < p.nk) & (fk <= nyq)
    ax, ay, az = np.pi * fx * dx, np.pi * fy * dy, np.pi * fz * dz
    with np.errstate(invalid="ignore", divide="ignore"):
        sxx = np.where(fx != 0, np.sin(ax) / ax, 1.0)
        syy = np.where(fy != 0, np.sin(ay) / ay, 1.0)
        szz = np.where(fz != 0, np.sin(az) / az, 1.0)
    cor = (1.0 / (sxx * syy * szz)) ** (2 * p.interp_order)
    NM = np.full(i.shape, 2.0)
    k0 = k == 0
    NM[k0 & (j == 0) & (i != 0) & (i != NX // 2)] = 1.0
    NM[k0 & (j != 0) & (i != NX // 2)] = 1.0
    kvec = 2.0 * np.pi * np.stack([fx, fy, fz], axis=-1)
    sel = ok
    return (kbin[sel], NM[sel], cor[sel], kvec[sel], fk[sel],
            np.stack([i[sel], j[sel], k[sel]], axis=-1))


def _unit_grid(p, cells):
    g = orc.new_grid(p)
    v = orc.real_view(p, g)
    for c in cells:
        v[c] += 1.0
    return g


def test_combos_cover_every_monomial_once():
    """The index sets of the reference loops: multisets {ii,jj}, {ii,jj,kk},
    {ii,ii,jj,kk} are distinct and cover all monomials of degree l."""
    for l, key in ((1, lambda c: (c[0],)), (2, lambda c: tuple(sorted(c[:2]))),
                   (3, lambda c: tuple(sorted(c))), (4, lambda c: tuple(sorted((c[0],) + c)))):
        combos = orc.survey_combos(l)
        sets = [key(c) for c in combos]
        want = set(itertools.combinations_with_replacement(range(3), l))
        assert len(sets) == len(want) == {1: 3, 2: 6, 3: 10, 4: 15}[l]
        assert set(sets) == want


def test_mode_weights_match_periodic_oracle():
    """sum NM cor per bin (the weights restated here) equals the periodic
    oracle's P0 of one unit source, whose |F(k)|^2 = 1."""
    p = orc.params(nx=24, ny=20, nz=16, xmax=L, ymax=L * 20 / 24, zmax=L * 16 / 24,
                   interp_order=3, nk=18, maxk=0.18)
    g = _unit_grid(p, [(5, 7, 3)])
    orc.fft_r2c_3d(g, p.nx, p.ny, p.nz)
    b = orc.compute_power(p, [g, None, None, None])
    kbin, NM, cor, _, _, _ = _modes(p)
    s0 = np.bincount(kbin, weights=NM * cor, minlength=p.nk)
    assert np.array_equal(np.bincount(kbin, weights=NM, minlength=p.nk).astype(np.int64), b.nmodes)
    assert np.allclose(b.pk[0], s0, rtol=1e-12, atol=0)


@pytest.mark.parametrize("order,odd", [(2, 1), (3, 0), (1, 1)])
def test_two_sources_closed_form(order, odd):
    p = orc.params(nx=24, ny=20, nz=16, xmax=L, ymax=L * 20 / 24, zmax=L * 16 / 24,
                   interp_order=order, nk=18, maxk=0.18, odd_multipoles=odd)
    d = np.array([L / 24] * 3)
    ca, cb = (3, 15, 11), (19, 4, 6)
    origin = np.array([-310.0, 1240.0, -455.0])   # generic directions, no zero components
    g = _unit_grid(p, [ca, cb])
    res = orc.survey_power(p, origin, [g, None, None, None])
    kbin, NM, cor, kvec, fk, idx = _modes(p)
    w = NM * cor
    na = np.array(ca) * d - origin
    nb = np.array(cb) * d - origin
    na /= np.linalg.norm(na)
    nb /= np.linalg.norm(nb)
    khat = kvec / np.where(fk > 0, fk, 1.0)[:, None]
    mua, mub = khat @ na, khat @ nb
    n = np.array([p.nx, p.ny, p.nz])
    theta = 2.0 * np.pi * (idx * (np.array(cb) - np.array(ca)) / n).sum(axis=1)
    nonzero = fk > 0
    assert np.array_equal(res.nmodes, np.bincount(kbin, weights=NM, minlength=p.nk).astype(np.int64))
    for l in (0, 1, 2, 3, 4):
        if l % 2 and not odd:
            continue
        if l % 2:
            term = np.sin(theta) * (_legendre(l, mub) - _legendre(l, mua)) * nonzero
        elif l == 0:
            term = (1 + np.cos(theta)) * 2.0
        else:
            # the k = 0 mode (fktot = 0) enters only the monopole pass:
            # -0.5 and 0.375 of its power for l = 2, 4
            term = np.where(nonzero, (1 + np.cos(theta)) * (_legendre(l, mua) + _legendre(l, mub)),
                            (1 + np.cos(theta)) * 2.0 * (-0.5 if l == 2 else 0.375))
        want = np.bincount(kbin, weights=w * term, minlength=p.nk)
        scale = np.bincount(kbin, weights=w * 4.0, minlength=p.nk)
        assert np.all(np.abs(res.pk[l] - want) <= 1e-11 * scale), l


def test_momentum_without_grid_is_rejected():
    p = orc.params(nx=8, ny=8, nz=8, momentum=2)
    g = orc.new_grid(p)
    with pytest.raises(ValueError):
        orc.survey_power(p, [0.0, 0.0, 0.0], [g, None, None, None])


@pytest.mark.parametrize("odd", [0, 1])
def test_momentum_cross_closed_form(odd):
    """Momentum with a ddg_mom grid (src/main.c:877-898): A = unit source at a
    (ddg), B = unit source at b (ddg_mom).  The monopole pass gives
    Im(A0* B0) = -sin theta, the even passes 0.5 sin theta (P(n_b) - P(n_a)),
    the odd ones 0.5 cos theta (P(n_a) + P(n_b))."""
    p = orc.params(nx=20, ny=24, nz=16, xmax=L * 20 / 24, ymax=L, zmax=L * 16 / 24,
                   interp_order=2, nk=15, maxk=0.15, momentum=3, odd_multipoles=odd)
    d = np.array([L / 24] * 3)
    ca, cb = (2, 17, 9), (13, 5, 3)
    origin = np.array([905.0, -372.0, 611.0])
    ga, gb = _unit_grid(p, [ca]), _unit_grid(p, [cb])
    res = orc.survey_power(p, origin, [ga, None, gb, None])
    kbin, NM, cor, kvec, fk, idx = _modes(p)
    w = NM * cor
    na = np.array(ca) * d - origin
    nb = np.array(cb) * d - origin
    na /= np.linalg.norm(na)
    nb /= np.linalg.norm(nb)
    khat = kvec / np.where(fk > 0, fk, 1.0)[:, None]
    mua, mub = khat @ na, khat @ nb
    theta = 2.0 * np.pi * (idx * (np.array(cb) - np.array(ca)) / np.array([p.nx, p.ny, p.nz])).sum(axis=1)
    s, c, nz = np.sin(theta), np.cos(theta), fk > 0
    terms = {
        0: -s,
        2: 0.5 * s + nz * 0.5 * s * 1.5 * (mub**2 - mua**2),
        4: -0.375 * s + nz * 0.5 * s * (-3.75 * (mub**2 - mua**2) + 4.375 * (mub**4 - mua**4)),
        1: nz * 0.5 * c * (mua + mub),
        3: nz * 0.5 * c * (_legendre(3, mua) + _legendre(3, mub)),
    }
    scale = np.bincount(kbin, weights=w * 4.0, minlength=p.nk)
    for l in ((0, 1, 2, 3, 4) if odd else (0, 2, 4)):
        want = np.bincount(kbin, weights=w * terms[l], minlength=p.nk)
        assert np.all(np.abs(res.pk[l] - want) <= 1e-11 * scale), l
