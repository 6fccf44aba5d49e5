"""The x pass fused with the binning (FP32 out-of-place transform, mirror-group
1-D bins): capow_fft leaves the x pass pending and capow_bin transforms and
bins the x rows without writing them (reference src/main.c:266-271 then
302-468).

Against the separate x pass + mirror binning (CAPOW_XBIN=0) on the same
catalogue: the transforms are the same register FFT compiled into two
kernels (the compiler may contract products into FMAs differently, so the
FP32 values can differ by an ulp), and the bin sums are accumulated
differently (exact int64 quanta of one quantum per spectrum instead of
per-wave FP64 copies).  Mode counts are bit-exact and the sums agree within
2e-6 of the bin's scale (FP32 rounding; the FP32 form's bar against FP64 is
1e-4); the k-space grids read afterwards (the pending x pass run on demand)
are bitwise those of the separate path.  Cases cover one,
two and four grids (16 and 8 rows per item), odd multipoles, log bins,
non-cubic grids and a grid whose x Nyquist frequency is the lowest (the x
Nyquist group binned at ky = kz = 0 only).
"""
import sys

import numpy as np
import pytest

from capow_amd import api, mocks

pytestmark = pytest.mark.gpu

L = 1000.0

CASES = [
    ("g4_tsc_int_m1", (64, 64, 64), dict(interp_order=3, do_interlacing=1, momentum=1,
                                          extra_spectra=3)),
    ("g4_odd_los_z", (64, 32, 128), dict(interp_order=3, do_interlacing=1, momentum=1,
                                         extra_spectra=3, odd_multipoles=1, los=3)),
    ("g2_cic_int_m0", (128, 96, 64), dict(interp_order=2, do_interlacing=1, momentum=0)),
    # x has the lowest Nyquist frequency (within maxk): the x Nyquist group
    # (NX/2, 0, 0) is binned, every other x Nyquist group is not
    ("g2_x_nyquist", (64, 96, 128), dict(interp_order=2, do_interlacing=1, momentum=0)),
    ("g2_hb_no_int_m3", (96, 64, 32), dict(interp_order=2, do_interlacing=0, momentum=3,
                                            extra_spectra=3)),
    ("g1_ngp_m2", (64, 128, 64), dict(interp_order=1, do_interlacing=0, momentum=2)),
    ("g4_log_bins", (128, 64, 64), dict(interp_order=3, do_interlacing=1, momentum=1,
                                         extra_spectra=3, output_log=1, mink=-2.2, maxk=-0.5,
                                         nk=25)),
    ("g4_256", (256, 256, 256), dict(interp_order=3, do_interlacing=1, momentum=1,
                                      extra_spectra=3, odd_multipoles=1)),
]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _run(p, xyzv, fused, monkeypatch):
    if not fused:
        monkeypatch.setenv("CAPOW_XBIN", "0")
    ps = api.PowerSpectrum(p)
    ps.set_timing(True)
    ps.set_particles(xyzv)
    ps.assign()
    ps.fft()
    res = ps.bin()
    names = list(ps.kernel_times())
    grids = [ps.get_grid(s).copy() for s in _slots(p)]
    ps.close()
    monkeypatch.delenv("CAPOW_XBIN", raising=False)
    return res, names, grids


def _slots(p):
    hb = p.momentum not in (0, 2, 5)
    return [s for s, u in enumerate((True, p.do_interlacing, hb, hb and p.do_interlacing)) if u]


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_fused_x_pass_binning_matches_separate(case, monkeypatch):
    name, (nx, ny, nz), kw = case
    say("fused x pass + binning case", name)
    base = dict(nx=nx, ny=ny, nz=nz, xmax=L, ymax=L, zmax=L, nk=30, maxk=0.3, precision=api.FP32)
    base.update(kw)
    p = api.make_params(**base)
    pos, vel = mocks.zeldovich(max(nx, ny, nz) // 2, L, 13)
    xyzv = mocks.xyzv(pos, vel, p.los).astype(np.float32)
    fr, fnames, fgrids = _run(p, xyzv, True, monkeypatch)
    sr, snames, sgrids = _run(p, xyzv, False, monkeypatch)
    assert "fft_x_bin" in fnames and "fft_x" not in fnames[:fnames.index("fft_x_bin")], fnames
    assert "fft_x_bin" not in snames
    assert np.array_equal(fr["nmodes"], sr["nmodes"])
    ls = (0, 1, 2, 3, 4) if p.odd_multipoles else (0, 2, 4)
    autos = [c for c in fr["codes"] if c in (0, 2, 5)]
    cs = np.sqrt(np.abs(sr[autos[0]]["pk"][0] * sr[autos[1]]["pk"][0])) if len(autos) == 2 else 0.0
    for code in fr["codes"]:
        scale = np.abs(sr[code]["pk"][0]) + cs
        for l in ls:
            d = np.abs(fr[code]["pk"][l] - sr[code]["pk"][l])
            rel = (d / np.maximum(scale, 1e-300)).max()
            say("  code", code, "l", l, "max rel", rel)
            assert rel <= 2e-6, (name, code, l, rel)
    # the pending x pass, run when the grids are read, is the separate pass
    for a, b in zip(fgrids, sgrids):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_fused_binning_twice_and_after_backward(monkeypatch):
    """capow_bin twice on one transform gives bitwise the same bins; a
    backward in between (which needs the k-space grids) leaves the forward
    path working."""
    say("fused binning twice")
    n = 64
    p = api.make_params(nx=n, ny=n, nz=n, xmax=L, ymax=L, zmax=L, interp_order=3,
                        do_interlacing=1, momentum=1, extra_spectra=3, nk=20, maxk=0.2,
                        precision=api.FP32)
    pos, vel = mocks.zeldovich(32, L, 4)
    xyzv = mocks.xyzv(pos, vel, 0).astype(np.float32)
    ps = api.PowerSpectrum(p)
    ps.set_particles(xyzv)
    ps.assign()
    ps.fft()
    a = ps.bin()
    b = ps.bin()
    for code in a["codes"]:
        assert np.array_equal(a[code]["pk"].view(np.uint64), b[code]["pk"].view(np.uint64))
    G = np.random.default_rng(1).normal(size=(3, 5, 20))
    ps.backward(G)
    c = ps.power()
    for code in a["codes"]:
        assert np.array_equal(a[code]["pk"].view(np.uint64), c[code]["pk"].view(np.uint64))
    ps.close()


@pytest.mark.parametrize("prec", [api.FP32, api.FP64])
def test_one_spare_pipeline_matches_full_scratch(prec, monkeypatch):
    """Grids too large for a second set (four FP32 grids at 2048^3) run the
    transposed FFT grid by grid through one spare buffer
    (CAPOW_FFT_ONE_SPARE=1 forces it here): the same kernels on the same
    rows, so the bins, the k-space grids (the pending x pass run grid by
    grid) and a backward after them are bitwise those of the full-scratch
    pipeline.  Four grids (TSC + interlacing, density and momentum)."""
    n = 64   # NZ/2 = 32: register sizes on every axis (the backward needs them)
    p = api.make_params(nx=n, ny=n, nz=n, interp_order=3, do_interlacing=1, momentum=1,
                        extra_spectra=3, precision=prec, nk=30, maxk=0.3)
    pos, vel = mocks.zeldovich(40, 1000.0, 17)
    xyzv = mocks.xyzv(pos, vel, 0)
    G = np.random.default_rng(2).normal(size=(3, 5, 30))

    def run():
        ps = api.PowerSpectrum(p)
        ps.set_particles(xyzv)
        res = ps.power()
        grids = [ps.get_grid(s) for s in range(4)]
        gx, gw = ps.backward(G)
        res2 = ps.power()
        ps.close()
        return res, grids, gx, gw, res2

    full = run()
    monkeypatch.setenv("CAPOW_FFT_ONE_SPARE", "1")
    one = run()
    monkeypatch.delenv("CAPOW_FFT_ONE_SPARE")
    u = np.uint32 if prec == api.FP32 else np.uint64
    for r in (0, 4):
        for code in full[r]["codes"]:
            assert np.array_equal(full[r][code]["pk"], one[r][code]["pk"]), (r, code)
        assert np.array_equal(full[r]["nmodes"], one[r]["nmodes"])
    for s in range(4):
        assert np.array_equal(full[1][s].view(u), one[1][s].view(u)), s
    assert np.array_equal(full[2], one[2]) and np.array_equal(full[3], one[3])
