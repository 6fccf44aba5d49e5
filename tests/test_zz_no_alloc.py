"""capow_bin_noalloc (the binning with every buffer in the library's static
device array) against capow_bin on the same k-space grids: mode counts
bit-exact, raw sums within 1e-12 of the bin's scale, on two calls in a row,
the second with smaller grids and fewer bins, in FP32 and FP64, including
(k, mu) bins with odd multipoles."""
import ctypes as C
import sys

import numpy as np
import pytest

from capow_amd import api, mocks

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _context_bins(p, seed):
    """Deposit + FFT + capow_bin through a context; the k-space grids back
    on the host in the reference layout."""
    pos, vel = mocks.zeldovich(p.nx // 2, p.xmax - p.xmin, seed)
    ps = api.PowerSpectrum(p)
    ps.set_particles(mocks.xyzv(pos, vel, p.los))
    ps.assign()
    ps.fft()
    ref = ps.bin()
    grids = []
    for s in range(4):
        try:
            grids.append(ps.get_grid(s).copy())
        except api.CapowError:
            grids.append(None)
    ps.close()
    return ref, grids


def _noalloc_bins(p, grids):
    import torch
    fp64 = p.precision == api.FP64
    dt = torch.float64 if fp64 else torch.float32
    nzh = p.nz // 2 + 1
    al = 8 if fp64 else 16
    pitch = (nzh + al - 1) // al * al
    dev = []
    ptrs = (C.c_void_p * 4)()
    for s, g in enumerate(grids):
        if g is None:
            continue
        h = np.zeros((p.nx, p.ny, 2 * pitch), dtype=np.float64 if fp64 else np.float32)
        h[:, :, :2 * nzh] = g.reshape(p.nx, p.ny, 2 * nzh)
        t = torch.from_numpy(h).to("cuda")
        dev.append(t)
        ptrs[s] = t.data_ptr()
    ns = len(api.spectra_codes(p))
    nmu = p.nmu if p.output_2d else 1
    out = dict(pk=np.zeros(ns * 5 * p.nk), nmodes=np.zeros(p.nk, dtype=np.int64),
               pk2d=np.zeros(ns * p.nk * nmu), nmodes2d=np.zeros(p.nk * nmu, dtype=np.int64))
    b = api.Bins(0, (C.c_int32 * 3)(), api._dptr(out["pk"]), api._iptr(out["nmodes"]),
                 api._dptr(out["pk2d"]), api._iptr(out["nmodes2d"]))
    torch.cuda.synchronize()
    rc = api.lib().capow_bin_noalloc(C.byref(p), ptrs, pitch, None, C.byref(b))
    if rc != 0:
        raise api.CapowError(rc, api.lib().capow_last_error(None).decode())
    del dev
    return b, out


CASES = [
    # larger grids and more bins first, then a smaller call
    ("fp32_96_then_48", api.FP32, dict(interp_order=3, do_interlacing=1, momentum=1,
                                       extra_spectra=3, odd_multipoles=1), (96, 60), (48, 20)),
    ("fp64_64_then_32", api.FP64, dict(interp_order=2, do_interlacing=1, momentum=0),
     (64, 40), (32, 12)),
    ("fp32_2d_64_then_32", api.FP32, dict(interp_order=3, do_interlacing=1, momentum=1,
                                          extra_spectra=3, output_2d=1, nmu=6), (64, 30), (32, 10)),
    ("fp64_log_64_then_32", api.FP64, dict(interp_order=2, do_interlacing=0, momentum=3,
                                           output_log=1, mink=-2.0, maxk=-0.5), (64, 25), (32, 9)),
]


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_bin_noalloc_matches_capow_bin(case):
    name, prec, kw, first, second = case
    for call, (n, nk) in enumerate((first, second)):
        say("noalloc case", name, "call", call, "grid", n, "nk", nk)
        base = dict(nx=n, ny=n, nz=n, xmax=1000.0, ymax=1000.0, zmax=1000.0, nk=nk,
                    maxk=0.01 * nk, precision=prec)
        base.update(kw)
        if base.get("output_log"):
            base["maxk"] = kw["maxk"]
        p = api.make_params(**base)
        ref, grids = _context_bins(p, 3 + call)
        b, out = _noalloc_bins(p, grids)
        codes = list(b.codes[:b.n_spectra])
        assert codes == ref["codes"]
        assert np.array_equal(out["nmodes"], ref["nmodes"])
        pk = out["pk"].reshape(len(codes), 5, nk)
        for s, code in enumerate(codes):
            scale = np.abs(ref[code]["pk"][0]).max()
            d = np.abs(pk[s] - ref[code]["pk"]).max()
            say("  code", code, "max |diff| / scale", d / scale)
            assert d <= 1e-12 * scale
        if p.output_2d:
            assert np.array_equal(out["nmodes2d"].reshape(nk, -1), ref["nmodes2d"])
            pk2 = out["pk2d"].reshape(len(codes), nk, -1)
            for s, code in enumerate(codes):
                r2 = ref[code]["pk2d"]
                assert np.abs(pk2[s] - r2).max() <= 1e-12 * np.abs(r2).max()


def test_bin_noalloc_rejects_bad_layouts():
    """Argument checks that need no GPU work: a row pitch that is not whole
    128-byte lines, a missing grid, survey parameters."""
    p = api.make_params(nx=32, ny=32, nz=32, interp_order=2, momentum=1, precision=api.FP32)
    ptrs = (C.c_void_p * 4)()
    b = api.Bins()
    assert api.lib().capow_bin_noalloc(C.byref(p), ptrs, 17, None, C.byref(b)) == 1
    assert api.lib().capow_bin_noalloc(C.byref(p), ptrs, 32, None, C.byref(b)) == 1
    assert b"grid 0" in api.lib().capow_last_error(None)
    ps = api.make_params(nx=32, ny=32, nz=32, momentum=0, edge_mode=api.EDGE_SURVEY)
    assert api.lib().capow_bin_noalloc(C.byref(ps), ptrs, 32, None, C.byref(b)) == 5
