"""Pins the Momentum-code tables of the oracle (and, marked gpu, of the HIP
path) by closed forms that do not restate those tables.

The reference names its spectra in comments: Momentum 0..8 are P00, P01,
P11, P02, P12, P22, P03, P13, P04 (src/main.c:141-148, 213, 405).  P_ab is
the power of the fields w v^a and w v^b (v the line-of-sight velocity).
From the name alone, and the physics of those fields:

* field A = w v^a, field B = w v^b (no second field when a = b);
* the spectrum is |A|^2 when a = b; when a != b the cross spectrum is real
  for a + b even (Re A* B) and imaginary, odd in mu, for a + b odd
  (Im A* B);
* the mean is subtracted from a density field (a = 0) only;
* the Poisson shot noise of w v^a x w v^b is sum w^2 v^(a+b), i.e.
  N <v^(a+b)> for unit weights.

Three checks use these rules, never the oracle's own tables:

1. closed form: three NGP particles on grid points have exactly the Fourier
   amplitudes e^{-i k.x_j}.  Every stored mode of a 16^3 grid is
   enumerated in numpy (frequencies with the reference's float casts, the
   bin formula, mu, the Hermitian weight NM and the window deconvolution,
   src/main.c:329-429) and the raw sums of every bin and multipole, and
   the normalised P_l(k) with their shot noise, must match the oracle's
   for all nine codes.  Three distinct velocities separate (a, b) from
   (b, a) and from other splits of a + b, and fix the sign of Im;
2. constant velocity v0 on a clustered catalogue: for k != 0, B = v0^b A,
   so Re and auto codes are v0^(a+b) times the P00 sums and Im codes vanish
   (TSC + interlacing, every bin but the DC bin);
3. a 2-cell lattice with constant v0 has no power below k_L = pi/dx, so
   P_0 = -v0^(a+b) V/N there for every code: the shot-noise table.

A planted error in any of the oracle's tables (a field exponent, Re for Im,
the mean-subtracted set, a shot moment) fails at least one of them.
"""
import numpy as np
import pytest

NAMES = ["P00", "P01", "P11", "P02", "P12", "P22", "P03", "P13", "P04"]


def ab(code):
    return int(NAMES[code][1]), int(NAMES[code][2])


def form(code):
    a, b = ab(code)
    if a == b:
        return "auto"
    return "im" if (a + b) % 2 else "re"


def legendre(mu):
    return {0: np.ones_like(mu), 1: mu, 2: 1.5 * mu ** 2 - 0.5, 3: 2.5 * mu ** 3 - 1.5 * mu,
            4: 4.375 * mu ** 4 - 3.75 * mu ** 2 + 0.375}


def ftab(n, L):
    d = L / n
    s = 1.0 / (np.float64(np.float32(n)) * d)
    i = np.arange(n)
    return np.where(i <= n // 2, i.astype(np.float32).astype(np.float64) * s,
                    (i.astype(np.float32) - np.float32(n)).astype(np.float64) * s)


def stored_modes(n, L, nk, maxk, order):
    """Geometry of every stored mode (i, j, k <= n/2): bin, mu (LOS = z), NM,
    window deconvolution (no interlacing); the integer wavevector."""
    f = ftab(n, L)                   # f[k] = k sz >= 0 for k <= n/2, as fz is formed
    I, J, K = np.meshgrid(np.arange(n), np.arange(n), np.arange(n // 2 + 1), indexing="ij")
    FX, FY, FZ = f[I], f[J], f[K]
    fk = 2.0 * np.pi * np.sqrt(FX * FX + FY * FY + FZ * FZ)
    bw = maxk / nk
    kb = np.trunc(fk.astype(np.float32).astype(np.float64) / bw).astype(np.int64)
    d = L / n
    ok = (kb >= 0) & (kb < nk) & (fk <= np.pi / d)
    with np.errstate(divide="ignore", invalid="ignore"):
        mu = np.where(fk > 0, 2.0 * np.pi * FZ / fk, 0.0)

    def sinc(x):
        return np.where(x == 0, 1.0, np.sin(x) / np.where(x == 0, 1.0, x))

    w = sinc(np.pi * FX * d) * sinc(np.pi * FY * d) * sinc(np.pi * FZ * d)
    cor = (1.0 / w) ** (2 * order)
    nm = np.full(I.shape, 2, np.int64)
    plane = K == 0
    nm[plane & (J == 0) & ~((I == 0) | (I == n // 2))] = 1
    nm[plane & (J != 0) & (I != n // 2)] = 1
    return dict(I=I, J=J, K=K, kb=kb, ok=ok, mu=mu, cor=cor, nm=nm)


# three particles on grid points of a 16^3 grid, L = 16 (dx = 1)
N16, L16, NK16 = 16, 16.0, 8
PTS = np.array([[3.0, 5.0, 2.0], [3.0, 5.0, 5.0], [11.0, 1.0, 9.0]])
VZ = np.array([1.75, -0.625, 0.875])      # exact in float32: both forms see the same v


def closed_form_raw(code, plant=None):
    """Raw sums [5][nk] and mode counts of the three on-grid particles.
    plant (tests of this check only) overrides a rule: {"ab": (a, b)},
    {"form": "re"|"im"|"auto"} or {"meansub": bool}."""
    plant = plant or {}
    a, b = plant.get("ab", ab(code))
    m = stored_modes(N16, L16, NK16, NK16 * 0.3, 1)
    ph = np.zeros(m["I"].shape + (len(PTS),), np.complex128)
    for j, x in enumerate(PTS):
        ph[..., j] = np.exp(-2j * np.pi * (m["I"] * x[0] + m["J"] * x[1] + m["K"] * x[2]) / N16)
    A = (ph * VZ ** a).sum(-1)
    B = (ph * VZ ** b).sum(-1)
    dc = (m["I"] == 0) & (m["J"] == 0) & (m["K"] == 0)
    if plant.get("meansub", a == 0):
        A = np.where(dc, A - len(PTS), A)   # mean subtraction: alpha N^3 = N off the DC
    fm = plant.get("form", form(code))
    if fm == "auto":
        P = np.abs(A) ** 2
    elif fm == "re":
        P = (np.conj(A) * B).real
    else:
        P = (np.conj(A) * B).imag
    Ls = legendre(m["mu"])
    raw = np.zeros((5, NK16))
    ok = m["ok"]
    for l in range(5):
        raw[l] = np.bincount(m["kb"][ok], weights=(m["nm"] * m["cor"] * Ls[l] * P)[ok],
                             minlength=NK16)[:NK16]
    nmodes = np.bincount(m["kb"][ok], weights=m["nm"][ok], minlength=NK16)[:NK16]
    shot = len(PTS) if code == 0 else float((VZ ** (a + b)).sum())
    norm = len(PTS) ** 2 / L16 ** 3
    return raw, nmodes.astype(np.int64), shot, norm


def closed_form_finalized(raw, nmodes, shot, norm):
    """src/main.c:1014-1056: P0 shot subtraction, (2l+1)/(Nmodes norm),
    highest filled bin zeroed."""
    out = np.zeros_like(raw)
    filled = nmodes > 0
    nmn = np.where(filled, nmodes * norm, 1.0)
    out[0] = np.where(filled, (raw[0] - nmodes * shot) / nmn, 0.0)
    for l in (1, 2, 3, 4):
        out[l] = np.where(filled, (2 * l + 1) * raw[l] / nmn, 0.0)
    last = np.nonzero(filled)[0].max()
    out[:, last] = 0.0
    nm = nmodes.copy()
    nm[last] = 0
    return out, nm


def _three_particles():
    pos = PTS.copy()
    vel = np.zeros_like(pos)
    vel[:, 2] = VZ
    return pos, vel


def _params16(oracle, code):
    return oracle.params(nx=N16, ny=N16, nz=N16, xmax=L16, ymax=L16, zmax=L16, interp_order=1,
                         do_interlacing=0, los=0, momentum=code, nk=NK16, maxk=NK16 * 0.3,
                         odd_multipoles=1)


@pytest.mark.parametrize("code", range(9), ids=NAMES)
def test_oracle_codes_match_closed_form(code, oracle):
    pos, vel = _three_particles()
    p = _params16(oracle, code)
    rc, ob, st, alpha, shot, norm = oracle.run(p, pos, vel)
    assert rc == 0
    raw, nmodes, eshot, enorm = closed_form_raw(code)
    assert np.array_equal(ob.nmodes, nmodes)
    scale = np.abs(raw).max()
    for l in range(5):
        d = np.abs(ob.pk[l] - raw[l]).max()
        assert d <= 1e-12 * scale, (NAMES[code], l, d / scale)
    assert shot == pytest.approx(eshot, rel=1e-14)
    assert norm == pytest.approx(enorm, rel=1e-14)
    oracle.finalize(p, shot, norm, ob)
    fin, fnm = closed_form_finalized(raw, nmodes, eshot, enorm)
    assert np.array_equal(ob.nmodes, fnm)
    fs = np.abs(fin).max()
    for l in range(5):
        assert np.abs(ob.pk[l] - fin[l]).max() <= 1e-12 * fs, (NAMES[code], l)


@pytest.mark.parametrize("code", range(9), ids=NAMES)
def test_planted_table_errors_fail(code):
    """Every other split (a', b') of the field exponents (a' + b' <= 5, the
    transpose included), the other power forms, and the opposite
    mean-subtraction choice change this code's closed form wherever they
    could matter, so the comparison above sees each of them."""
    base = closed_form_raw(code)[0]
    s = np.abs(base).max()
    a, b = ab(code)

    def differs(**plant):
        return np.abs(closed_form_raw(code, plant)[0] - base).max() > 1e-6 * s

    for a2 in range(5):
        for b2 in range(5 - a2 + 1):
            if (a2, b2) != (a, b) and not (a == b and a2 == b2 == a):
                if form(code) == "auto" and a2 != b2:
                    continue     # an auto spectrum has one field: only a2 = b2 is a table error
                if form(code) == "re" and (b2, a2) == (a, b) and a * b != 0:
                    continue     # Re A* B = Re B* A, and neither field is mean-subtracted
                assert differs(ab=(a2, b2)), (NAMES[code], a2, b2)
    for f2 in ("re", "im", "auto"):
        # Re A* A is |A|^2: for an auto spectrum only Im is a different form
        if f2 != form(code) and not (form(code) == "auto" and f2 == "re"):
            assert differs(form=f2), (NAMES[code], f2)
    if form(code) != "im":       # Im A*B of real DC values is zero either way
        assert differs(meansub=a != 0), NAMES[code]


def _zeldovich_const_v(v0):
    from capow_amd import mocks
    pos, _ = mocks.zeldovich(24, 1000.0, 3)
    vel = np.zeros_like(pos)
    vel[:, 2] = v0
    return pos, vel


@pytest.mark.parametrize("code", range(1, 9), ids=NAMES[1:])
def test_oracle_constant_velocity_ratio(code, oracle):
    """Constant v0: Re/auto codes are v0^(a+b) x P00, Im codes vanish
    (every bin but the DC bin 0)."""
    v0 = 1.7
    pos, vel = _zeldovich_const_v(v0)
    kw = dict(nx=32, ny=32, nz=32, interp_order=3, do_interlacing=1, los=0, nk=20, maxk=0.2,
              odd_multipoles=1)
    rc0, b0, *_ = oracle.run(oracle.params(momentum=0, **kw), pos, vel)
    rc, bc, *_ = oracle.run(oracle.params(momentum=code, **kw), pos, vel)
    assert rc0 == 0 and rc == 0
    a, b = ab(code)
    scale = np.abs(b0.pk[0][1:]).max() * v0 ** (a + b)
    for l in range(5):
        exp = 0.0 if form(code) == "im" else v0 ** (a + b) * b0.pk[l][1:]
        assert np.abs(bc.pk[l][1:] - exp).max() <= 1e-12 * scale, (NAMES[code], l)


def _lattice(n, L, v0):
    dx = L / n
    q = np.arange(n // 2) * 2 * dx
    X, Y, Z = np.meshgrid(q + 0.3 * dx, q + 0.1 * dx, q + 0.45 * dx, indexing="ij")
    pos = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
    vel = np.zeros_like(pos)
    vel[:, 2] = v0
    return pos, vel


@pytest.mark.parametrize("code", range(9), ids=NAMES)
def test_oracle_lattice_shot_per_code(code, oracle):
    """P_0 = -v0^(a+b) V/N below k_L for every code: the shot table."""
    n, L, v0, nk, maxk = 32, 1000.0, 1.3, 20, 0.2
    pos, vel = _lattice(n, L, v0)
    p = oracle.params(nx=n, ny=n, nz=n, xmax=L, ymax=L, zmax=L, interp_order=2,
                      do_interlacing=1, momentum=code, nk=nk, maxk=maxk)
    rc, ob, st, alpha, shot, norm = oracle.run(p, pos, vel)
    assert rc == 0
    oracle.finalize(p, shot, norm, ob)
    a, b = ab(code)
    exp = -v0 ** (a + b) * L ** 3 / len(pos)
    edges_hi = (np.arange(nk) + 1) * (maxk / nk)
    sel = (edges_hi <= np.pi / (L / n)) & (ob.nmodes > 0)
    if a != 0 and form(code) != "im":
        sel[0] = False      # the DC mode of a field without mean subtraction is in bin 0
    assert sel.sum() >= 8
    np.testing.assert_allclose(ob.pk[0][sel], exp, rtol=1e-10)


# ---- the same closed forms through the HIP path -------------------------

@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("code", range(9), ids=NAMES)
def test_gpu_codes_match_closed_form(code, prec):
    from capow_amd import api, mocks
    print("gpu momentum closed form", NAMES[code], prec, flush=True)
    pos, vel = _three_particles()
    p = api.make_params(nx=N16, ny=N16, nz=N16, xmax=L16, ymax=L16, zmax=L16, interp_order=1,
                        do_interlacing=0, los=0, momentum=code, nk=NK16, maxk=NK16 * 0.3,
                        odd_multipoles=1, precision=api.FP64 if prec == "fp64" else api.FP32)
    ps = api.PowerSpectrum(p)
    ps.set_particles(mocks.xyzv(pos, vel, 0))
    res = ps.power()
    st = ps.stats
    ps.close()
    raw, nmodes, eshot, enorm = closed_form_raw(code)
    assert np.array_equal(np.asarray(res["nmodes"]), nmodes)
    tol = 1e-10 if prec == "fp64" else 1e-5
    scale = np.abs(raw).max()
    for l in range(5):
        d = np.abs(res[code]["pk"][l] - raw[l]).max()
        print("  l", l, "max rel", d / scale, flush=True)
        assert d <= tol * scale, (NAMES[code], l, d / scale)
    shot, norm = api.scalars(p, st, code)
    assert shot == pytest.approx(eshot, rel=1e-14)
    assert norm == pytest.approx(enorm, rel=1e-14)


@pytest.mark.gpu
@pytest.mark.parametrize("code", range(9), ids=NAMES)
def test_gpu_lattice_shot_per_code(code):
    from capow_amd import api, mocks
    print("gpu lattice shot", NAMES[code], flush=True)
    n, L, v0, nk, maxk = 32, 1000.0, 1.3, 20, 0.2
    pos, vel = _lattice(n, L, v0)
    p = api.make_params(nx=n, ny=n, nz=n, xmax=L, ymax=L, zmax=L, interp_order=2,
                        do_interlacing=1, momentum=code, nk=nk, maxk=maxk, precision=api.FP64)
    ps = api.PowerSpectrum(p)
    ps.set_particles(mocks.xyzv(pos, vel, 0))
    res = ps.power()
    shot, norm = api.scalars(p, ps.stats, code)
    pk5, nm, _, _, _ = api.finalize(p, shot, norm, res[code]["pk"].ravel(), res["nmodes"])
    ps.close()
    pk5 = pk5.reshape(5, -1)
    a, b = ab(code)
    exp = -v0 ** (a + b) * L ** 3 / len(pos)
    edges_hi = (np.arange(nk) + 1) * (maxk / nk)
    sel = (edges_hi <= np.pi / (L / n)) & (np.asarray(nm) > 0)
    if a != 0 and form(code) != "im":
        sel[0] = False      # the DC mode of a field without mean subtraction is in bin 0
    assert sel.sum() >= 8
    np.testing.assert_allclose(pk5[0][sel], exp, rtol=1e-10)
