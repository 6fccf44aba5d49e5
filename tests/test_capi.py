"""CPU checks of the C ABI: the library loads and exports every entry point
include/capow.h declares, parameter validation mirrors check_inputs
(reference src/read_params.c:294-428), and the host-side output math
(capow_scalars / capow_finalize / capow_write_output) reproduces the
oracle's restatement of src/main.c:103-148 and 1014-1101 byte for byte."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from capow_amd import api

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "capow.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(capow_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = api.lib()
    syms = declared_symbols()
    assert len(syms) >= 20
    for s in syms:
        assert hasattr(L, s), s
    # and the ctypes layer binds exactly these
    assert sorted(api.SIGNATURES) == syms


def test_struct_layouts_match_header():
    # capow_params: 3 int64 + 6 double + 9 int32 (+4 pad) + 4 double + 4 int32
    assert C.sizeof(api.Params) == 3 * 8 + 6 * 8 + 9 * 4 + 4 + 4 * 8 + 4 * 4
    assert C.sizeof(api.AssignStats) == 9 * 8
    assert C.sizeof(api.Bins) == 4 + 12 + 4 * 8


@pytest.mark.parametrize("bad,msg", [
    (dict(interp_order=4), "InterpOrder"), (dict(los=5), "LOS"), (dict(do_interlacing=2), "DoInterlacing"),
    (dict(output_2d=1, nmu=0), "NMU"), (dict(output_log=3), "OutputLog"), (dict(nk=0), "NK"),
    (dict(nx=0), "NX"), (dict(xmax=-1.0), "XMAX"), (dict(maxk=-1.0), "Maxk"),
    (dict(momentum=9), "Momentum"), (dict(nz=33), "NZ"), (dict(nx=14 * 11), "2, 3 and 5"),
])
def test_create_rejects_invalid_parameters(bad, msg):
    p = api.make_params(**bad)
    with pytest.raises(api.CapowError) as e:
        api.PowerSpectrum(p)
    assert e.value.code in (1, 5)
    assert msg in str(e.value)


def test_create_without_gpu_fails_cleanly():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    with pytest.raises(api.CapowError) as e:
        api.PowerSpectrum(api.make_params(nx=16, ny=16, nz=16))
    assert e.value.code == 2


def _stats(rng):
    st = api.AssignStats()
    st.ngrid = 12345.0
    st.nsq = 13000.0
    st.vr, st.vr2, st.vr3, st.vr4 = rng.normal(size=4) * np.array([1e3, 1e7, 1e10, 1e13])
    return st


@pytest.mark.parametrize("code", range(9))
def test_scalars_match_oracle(code, oracle):
    rng = np.random.default_rng(code)
    p = api.make_params(momentum=code, xmin=-5.0, xmax=995.0, ymax=800.0, zmin=10.0, zmax=1210.0)
    st = _stats(rng)
    shot, norm = api.scalars(p, st, code)
    op = oracle.params(momentum=code, xmin=-5.0, xmax=995.0, ymax=800.0, zmin=10.0, zmax=1210.0)
    ost = oracle.OrcStats(ngrid=st.ngrid, nsq=st.nsq, vr=st.vr, vr2=st.vr2, vr3=st.vr3, vr4=st.vr4)
    a, s2, n2 = oracle.scalars(op, ost)
    assert shot == s2 and norm == n2


@pytest.mark.parametrize("odd,out2d,logk", [(0, 0, 0), (1, 0, 0), (0, 1, 0), (1, 1, 1)])
def test_finalize_and_output_match_oracle(odd, out2d, logk, oracle, tmp_path):
    rng = np.random.default_rng(odd * 4 + out2d * 2 + logk)
    nk, nmu = 13, 5
    kw = dict(nk=nk, nmu=nmu, odd_multipoles=odd, output_2d=out2d, output_log=logk,
              mink=-2.0 if logk else 0.0, maxk=-0.3 if logk else 0.26)
    p = api.make_params(**kw)
    op = oracle.params(**kw)
    pk5 = rng.normal(size=(5, nk)) * 1e6
    nm = rng.integers(0, 50, size=nk)
    nm[-2:] = 0
    pk2d = rng.normal(size=nk * nmu)
    nm2d = rng.integers(0, 9, size=nk * nmu)
    shot, norm = 1234.5, 0.0123
    g5, gnm, g2d, gnm2d, last = api.finalize(p, shot, norm, pk5.ravel(), nm, pk2d, nm2d)
    b = oracle.Bins(op)
    for l in range(5):
        b.pk[l][:] = pk5[l]
    b.nmodes[:] = nm
    b.pk2d[:] = pk2d
    b.nmodes2d[:] = nm2d
    olast = oracle.finalize(op, shot, norm, b)
    assert last == olast
    for l in ((0, 1, 2, 3, 4) if odd else (0, 2, 4)):
        assert np.array_equal(g5.reshape(5, nk)[l], b.pk[l])
    assert np.array_equal(gnm, b.nmodes)
    if out2d:
        assert np.array_equal(g2d, b.pk2d)
    for precise in (False, True):
        a, o = str(tmp_path / ("a%d" % precise)), str(tmp_path / ("o%d" % precise))
        api.write_output(p, shot, norm, g5, gnm, a, g2d if out2d else None,
                         gnm2d if out2d else None, precise=precise)
        oracle.write_output(op, shot, norm, b, o, precise=precise)
        assert open(a).read() == open(o).read()
        if out2d:
            assert open(a + "_2D").read() == open(o + "_2D").read()
