"""Size-independent properties at the full benchmark configuration (1024^3
grid, 134,217,728 particles of a 512^3 Zel'dovich lattice, L = 2000 Mpc/h,
TSC + interlacing, density and momentum grids, P00/P11/P01, dk = 0.005,
NK = 321), where the oracle cannot run:

* the mode counts equal a brute-force count over the full k-space lattice
  (NumPy, the reference's float casts and bin formula, src/main.c:329-375)
  plus the reference's documented NM excess -- an independent formulation of
  the half-space NM rule (the same check pins the oracle at small sizes in
  test_oracle_power.py);
* the FP32 form agrees with the FP64 form within 1e-4 of the bin's scale
  for k < k_Nyq/2 on the same float32-representable catalogue;
* the FP32 form is bitwise reproducible.
"""
import numpy as np
import pytest

from capow_amd import api, mocks

from .test_gpu_parity import _finalized, say

pytestmark = pytest.mark.gpu

N = 1024
LAT = 512
LBOX = 2000.0
DK = 0.005
NK = int(np.floor(np.pi * N / LBOX / DK))   # 321


def _params(prec):
    return api.make_params(nx=N, ny=N, nz=N, xmax=LBOX, ymax=LBOX, zmax=LBOX, interp_order=3,
                           do_interlacing=1, los=0, momentum=1, nk=NK, maxk=NK * DK,
                           precision=prec, extra_spectra=api.SPEC_AUTO_A | api.SPEC_AUTO_B)


def _ftab(n, L):
    d = L / n
    s = 1.0 / (np.float64(np.float32(n)) * d)
    i = np.arange(n)
    return np.where(i <= n // 2, i.astype(np.float32).astype(np.float64) * s,
                    (i.astype(np.float32) - np.float32(n)).astype(np.float64) * s)


def _full_space_counts(n, L, nk, maxk):
    """One count per full-space mode, plane by plane (x), with the
    reference's evaluation order fx^2 + fy^2 + fz^2; planes i and n - i are
    identical (f(n - i) = -f(i) exactly), so half of them are counted twice.
    Plus the reference's NM excess: the kz = NZ/2 plane, the (NX/2, j, 0)
    line and DC carry NM = 2 for a single full-space mode."""
    f = _ftab(n, L)
    bw = maxk / nk
    nyq = np.pi / (L / n)
    fy2 = (f * f)[:, None]
    fz2 = (f * f)[None, :]
    tot = np.zeros(nk, np.int64)
    extra = np.zeros(nk, np.int64)

    def plane_bins(fx):
        fk = 2.0 * np.pi * np.sqrt((fx * fx + fy2) + fz2)
        kb = np.trunc(fk.astype(np.float32).astype(np.float64) / bw).astype(np.int64)
        ok = (kb >= 0) & (kb < nk) & (fk <= nyq)
        return kb, ok

    for i in range(n // 2 + 1):
        kb, ok = plane_bins(f[i])
        c = np.bincount(kb[ok], minlength=nk)[:nk]
        tot += c if i in (0, n // 2) else 2 * c
        # excess: the kz = n/2 column of every plane (and its mirror plane)
        col = kb[:, n // 2][ok[:, n // 2]]
        e = np.bincount(col, minlength=nk)[:nk]
        extra += e if i in (0, n // 2) else 2 * e
        if i == n // 2:
            line = kb[:, 0][ok[:, 0]]
            extra += np.bincount(line, minlength=nk)[:nk]
        if i == 0 and ok[0, 0]:
            extra[kb[0, 0]] += 1
    return tot + extra


@pytest.fixture(scope="module")
def c2():
    """FP32 form twice and the FP64 form once on the C2 catalogue."""
    import torch
    say("full size: generating the C2 catalogue")
    parts = mocks.zeldovich_torch(LAT, LBOX, seed=1, device="cuda", dtype=torch.float32)
    out = {}
    for tag, prec, dt in (("f32a", api.FP32, torch.float32), ("f32b", api.FP32, torch.float32),
                          ("f64", api.FP64, torch.float64)):
        say("full size run", tag)
        p = _params(prec)
        ps = api.PowerSpectrum(p)
        xyzv = parts if dt == torch.float32 else parts.to(dt)
        ps.set_particles(xyzv)
        res = ps.power()
        out[tag] = (p, res, ps.stats)
        ps.close()
        del xyzv
        torch.cuda.empty_cache()
    del parts
    torch.cuda.empty_cache()
    return out


def test_c2_mode_counts_match_full_space_brute_force(c2):
    say("full size: brute-force mode counts")
    exp = _full_space_counts(N, LBOX, NK, NK * DK)
    for tag in ("f32a", "f64"):
        got = np.asarray(c2[tag][1]["nmodes"], dtype=np.int64)
        assert np.array_equal(got, exp), (tag, np.nonzero(got != exp)[0][:10])


def test_c2_fp32_form_within_1e4_of_fp64(c2):
    fin = {}
    for tag in ("f32a", "f64"):
        p, res, st = c2[tag]
        fin[tag] = {c: _finalized(p, res, st, c) for c in res["codes"]}
    kc = (np.arange(NK) + 0.5) * DK
    sel = kc < (np.pi * N / LBOX) / 2
    f64 = fin["f64"]
    s00 = np.abs(f64[0][0][0]) + f64[0][2]
    s11 = np.abs(f64[2][0][0]) + f64[2][2]
    for code in (0, 1, 2):
        scale = s00 if code == 0 else (s11 if code == 2 else np.sqrt(s00 * s11))
        for l in (0, 2, 4):
            rel = np.abs(fin["f32a"][code][0][l] - f64[code][0][l]) / scale
            say("  code", code, "l", l, "max rel (k<kNyq/2)", rel[sel].max())
            assert rel[sel].max() <= 1e-4, (code, l)


def test_c2_fp32_form_is_bitwise_reproducible(c2):
    a, b = c2["f32a"][1], c2["f32b"][1]
    for code in a["codes"]:
        assert np.array_equal(a[code]["pk"].view(np.uint64), b[code]["pk"].view(np.uint64)), code
    assert np.array_equal(a["nmodes"], b["nmodes"])
