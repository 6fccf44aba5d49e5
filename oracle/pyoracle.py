"""ctypes binding of the FP64 CPU oracle (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, and only as the checker.  The product library never loads it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")


class OrcParams(C.Structure):
    _fields_ = [
        ("nx", C.c_longlong), ("ny", C.c_longlong), ("nz", C.c_longlong),
        ("xmin", C.c_double), ("xmax", C.c_double),
        ("ymin", C.c_double), ("ymax", C.c_double),
        ("zmin", C.c_double), ("zmax", C.c_double),
        ("interp_order", C.c_int), ("do_interlacing", C.c_int),
        ("los", C.c_int), ("momentum", C.c_int),
        ("nk", C.c_int), ("nmu", C.c_int), ("output_2d", C.c_int),
        ("output_log", C.c_int), ("odd_multipoles", C.c_int),
        ("mink", C.c_double), ("maxk", C.c_double),
        ("omega_m", C.c_double), ("redshift", C.c_double),
    ]


class OrcStats(C.Structure):
    _fields_ = [(n, C.c_double) for n in
                ("nread", "nout", "ngrid", "npv", "nsq", "vr", "vr2", "vr3", "vr4")]


_D = C.POINTER(C.c_double)
_L = C.POINTER(C.c_longlong)


class OrcBins(C.Structure):
    _fields_ = [("pk0", _D), ("pk1", _D), ("pk2", _D), ("pk3", _D), ("pk4", _D),
                ("nmodes", _L), ("pk2d", _D), ("nmodes2d", _L)]


_lib = None


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        P = C.POINTER(OrcParams)
        L.orc_grid_size.argtypes = [P]; L.orc_grid_size.restype = C.c_longlong
        L.orc_add_to_grid.argtypes = [P] + [C.c_double] * 5 + [_D]
        L.orc_add_to_grid.restype = C.c_double
        L.orc_assign.argtypes = [P, _D, _D, _D, C.c_longlong, C.POINTER(_D), C.POINTER(OrcStats)]
        L.orc_assign.restype = C.c_int
        L.orc_scalars.argtypes = [P, C.POINTER(OrcStats), _D, _D, _D]
        L.orc_mean_subtract.argtypes = [P, C.c_double, C.POINTER(_D)]
        L.orc_fft_r2c_3d.argtypes = [_D, C.c_longlong, C.c_longlong, C.c_longlong]
        L.orc_fft_c2c_1d.argtypes = [_D, C.c_longlong]
        L.orc_compute_power.argtypes = [P, C.POINTER(_D), C.POINTER(OrcBins)]
        L.orc_finalize.argtypes = [P, C.c_double, C.c_double, C.POINTER(OrcBins)]
        L.orc_finalize.restype = C.c_int
        L.orc_write_output.argtypes = [P, C.c_double, C.c_double, C.POINTER(OrcBins), C.c_char_p, C.c_int]
        L.orc_write_output.restype = C.c_int
        L.orc_run.argtypes = [P, _D, _D, _D, C.c_longlong, C.POINTER(OrcBins), C.POINTER(OrcStats),
                              _D, _D, _D]
        L.orc_run.restype = C.c_int
        L.orc_survey_power.argtypes = [P, _D, C.POINTER(_D), C.POINTER(OrcBins)]
        L.orc_survey_power.restype = C.c_int
        L.orc_survey_combos.argtypes = [C.c_int, C.POINTER(C.c_int)]
        L.orc_survey_combos.restype = C.c_int
        _lib = L
    return _lib


DEFAULTS = dict(nx=32, ny=32, nz=32, xmin=0.0, xmax=1000.0, ymin=0.0, ymax=1000.0,
                zmin=0.0, zmax=1000.0, interp_order=2, do_interlacing=0, los=0,
                momentum=0, nk=20, nmu=1, output_2d=0, output_log=0, odd_multipoles=0,
                mink=0.0, maxk=0.2, omega_m=0.3, redshift=0.5)


def params(**kw):
    d = dict(DEFAULTS)
    d.update(kw)
    return OrcParams(**d)


def _dp(a):
    return a.ctypes.data_as(_D) if a is not None else None


def grid_size(p):
    return lib().orc_grid_size(C.byref(p))


def new_grid(p):
    return np.zeros(grid_size(p), dtype=np.float64)


def add_to_grid(p, x, y, z, w, vr, grid):
    return lib().orc_add_to_grid(C.byref(p), x, y, z, w, vr, _dp(grid))


def real_view(p, grid):
    """[NX+InterpOrder][NY][2(NZ/2+1)] view of a padded grid buffer."""
    return grid.reshape(p.nx + p.interp_order, p.ny, 2 * (p.nz // 2 + 1))


class Bins:
    def __init__(self, p):
        nk, nmu = p.nk, max(p.nmu, 1)
        self.pk = {l: np.zeros(nk) for l in range(5)}
        self.nmodes = np.zeros(nk, dtype=np.int64)
        self.pk2d = np.zeros(nk * nmu)
        self.nmodes2d = np.zeros(nk * nmu, dtype=np.int64)
        self.c = OrcBins(_dp(self.pk[0]), _dp(self.pk[1]), _dp(self.pk[2]),
                         _dp(self.pk[3]), _dp(self.pk[4]),
                         self.nmodes.ctypes.data_as(_L), _dp(self.pk2d),
                         self.nmodes2d.ctypes.data_as(_L))


def _grid_ptrs(grids):
    arr = (_D * 4)()
    for i in range(4):
        arr[i] = _dp(grids[i]) if grids[i] is not None else None
    return arr


def assign(p, pos, vel, w=None):
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    vel = np.ascontiguousarray(vel, dtype=np.float64)
    w = None if w is None else np.ascontiguousarray(w, dtype=np.float64)
    m, inter = p.momentum, p.do_interlacing
    has_b = m not in (0, 2, 5)
    grids = [new_grid(p), new_grid(p) if inter else None,
             new_grid(p) if has_b else None, new_grid(p) if (has_b and inter) else None]
    st = OrcStats()
    rc = lib().orc_assign(C.byref(p), _dp(pos), _dp(vel), _dp(w), len(pos),
                          _grid_ptrs(grids), C.byref(st))
    return rc, grids, st


def scalars(p, st):
    a, s, n = C.c_double(), C.c_double(), C.c_double()
    lib().orc_scalars(C.byref(p), C.byref(st), C.byref(a), C.byref(s), C.byref(n))
    return a.value, s.value, n.value


def mean_subtract(p, alpha, grids):
    lib().orc_mean_subtract(C.byref(p), alpha, _grid_ptrs(grids))


def fft_r2c_3d(grid, nx, ny, nz):
    lib().orc_fft_r2c_3d(_dp(grid), nx, ny, nz)


def fft_c2c_1d(x):
    a = np.ascontiguousarray(x, dtype=np.complex128).copy()
    lib().orc_fft_c2c_1d(a.view(np.float64).ctypes.data_as(_D), len(a))
    return a


def compute_power(p, grids):
    b = Bins(p)
    lib().orc_compute_power(C.byref(p), _grid_ptrs(grids), C.byref(b.c))
    return b


def survey_power(p, origin, grids):
    """Survey (Yamamoto) estimator on real-space grids (not modified)."""
    b = Bins(p)
    o = np.ascontiguousarray(origin, dtype=np.float64)
    rc = lib().orc_survey_power(C.byref(p), _dp(o), _grid_ptrs(grids), C.byref(b.c))
    if rc != 0:
        raise ValueError("survey estimator: Momentum %d has no ddg_mom grid" % p.momentum)
    return b


def survey_combos(multipole):
    idx = (C.c_int * 45)()
    n = lib().orc_survey_combos(multipole, idx)
    return [tuple(idx[3 * c:3 * c + 3]) for c in range(n)]


def finalize(p, shot, norm, b):
    return lib().orc_finalize(C.byref(p), shot, norm, C.byref(b.c))


def write_output(p, shot, norm, b, path, precise=False):
    return lib().orc_write_output(C.byref(p), shot, norm, C.byref(b.c),
                                  path.encode(), 1 if precise else 0)


def run(p, pos, vel, w=None):
    """Whole periodic pipeline; returns (rc, raw bins, stats, alpha, shot, norm)."""
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    vel = np.ascontiguousarray(vel, dtype=np.float64)
    w = None if w is None else np.ascontiguousarray(w, dtype=np.float64)
    b = Bins(p)
    st = OrcStats()
    a, s, n = C.c_double(), C.c_double(), C.c_double()
    rc = lib().orc_run(C.byref(p), _dp(pos), _dp(vel), _dp(w), len(pos), C.byref(b.c),
                       C.byref(st), C.byref(a), C.byref(s), C.byref(n))
    return rc, b, st, a.value, s.value, n.value
