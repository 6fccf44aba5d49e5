"""ctypes binding of libcapow.so (include/capow.h).

This is the checking/benchmark layer: the product is the C ABI and its HIP
kernels.  Loading fails loudly when the library has not been built; there is
no CPU fallback.

`PowerSpectrum` follows the reference's periodic call sequence
(create_grids -> read + add_to_grid -> mean subtraction -> fftw_execute ->
compute_periodic_power -> output_power, reference src/main.c:28-298) with
the same parameter names (src/read_params.c:65-227).
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcapow.so")

CAPOW_OK = 0
CAPOW_ERR_IO, CAPOW_ERR_PARSE, CAPOW_ERR_STALE, CAPOW_ERR_OUT_OF_BOX = 7, 8, 9, 110
ERRORS = {1: "CAPOW_ERR_ARG", 2: "CAPOW_ERR_HIP", 3: "CAPOW_ERR_NOMEM", 4: "CAPOW_ERR_STATE",
          5: "CAPOW_ERR_UNSUPPORTED", 6: "CAPOW_ERR_INTERNAL", 7: "CAPOW_ERR_IO",
          8: "CAPOW_ERR_PARSE", 110: "CAPOW_ERR_OUT_OF_BOX"}
FP32, FP64 = 0, 1
EDGE_REFERENCE, EDGE_PERIODIC, EDGE_SURVEY = 0, 1, 2
SPEC_AUTO_A, SPEC_AUTO_B = 1, 2
HOST, DEVICE = 0, 1


class CapowError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s (%d): %s" % (ERRORS.get(code, "?"), code, msg))
        self.code = code


class Params(C.Structure):
    _fields_ = [
        ("nx", C.c_int64), ("ny", C.c_int64), ("nz", C.c_int64),
        ("xmin", C.c_double), ("xmax", C.c_double), ("ymin", C.c_double),
        ("ymax", C.c_double), ("zmin", C.c_double), ("zmax", C.c_double),
        ("interp_order", C.c_int32), ("do_interlacing", C.c_int32), ("los", C.c_int32),
        ("momentum", C.c_int32), ("nk", C.c_int32), ("nmu", C.c_int32),
        ("output_2d", C.c_int32), ("output_log", C.c_int32), ("odd_multipoles", C.c_int32),
        ("mink", C.c_double), ("maxk", C.c_double), ("omega_m", C.c_double),
        ("redshift", C.c_double),
        ("precision", C.c_int32), ("edge_mode", C.c_int32), ("extra_spectra", C.c_int32),
        ("reserved", C.c_int32),
    ]


class AssignStats(C.Structure):
    _fields_ = [(n, C.c_double) for n in
                ("nread", "nout", "ngrid", "nsq", "vr", "vr2", "vr3", "vr4", "alpha")]


class Bins(C.Structure):
    _fields_ = [("n_spectra", C.c_int32), ("codes", C.c_int32 * 3),
                ("pk", C.POINTER(C.c_double)), ("nmodes", C.POINTER(C.c_int64)),
                ("pk2d", C.POINTER(C.c_double)), ("nmodes2d", C.POINTER(C.c_int64))]


class SurveyObj(C.Structure):
    _fields_ = [(k, C.c_double) for k in
                ("x", "y", "z", "redshift", "weight", "weight_pv", "nbar", "pv", "pverr")]


class SurveyParams(C.Structure):
    _fields_ = [("coord_type", C.c_int32), ("momentum", C.c_int32), ("nbar_column", C.c_int32),
                ("fkp_column", C.c_int32), ("omega_m", C.c_double), ("redmin_in", C.c_double),
                ("redmax_in", C.c_double), ("sky_area", C.c_double), ("fkp_pk", C.c_double),
                ("fkp_pk_mom", C.c_double), ("fkp_pk_cross", C.c_double)]


class SurveyTotals(C.Structure):
    _fields_ = [(k, C.c_double) for k in
                ("data_nbw", "data_nbwsq", "data_nbsqwsq", "data_nbsqwsq_pv", "data_vr",
                 "data_vrsq", "rand_nbw", "rand_nbwsq", "rand_nbsqwsq", "rand_nbsqwsq_pv",
                 "alpha", "shot", "norm")]


class Mock(C.Structure):
    _fields_ = [("xyzv", C.c_void_p), ("w", C.c_void_p), ("n", C.c_int64),
                ("config", C.c_int32), ("reserved", C.c_int32)]


# Exported symbols of include/capow.h with their ctypes signatures.
_P = C.POINTER(Params)
_V = C.c_void_p
_D = C.POINTER(C.c_double)
_I64 = C.POINTER(C.c_int64)
SIGNATURES = {
    "capow_create": ([_P, C.POINTER(_V)], C.c_int),
    "capow_destroy": ([_V], C.c_int),
    "capow_last_error": ([_V], C.c_char_p),
    "capow_set_stream": ([_V, _V], C.c_int),
    "capow_synchronize": ([_V], C.c_int),
    "capow_set_timing": ([_V, C.c_int], C.c_int),
    "capow_set_particles": ([_V, _V, _V, C.c_int64, C.c_int32, C.c_int32], C.c_int),
    "capow_assign": ([_V, C.POINTER(AssignStats)], C.c_int),
    "capow_fft": ([_V], C.c_int),
    "capow_bin": ([_V, C.POINTER(Bins)], C.c_int),
    "capow_bin_noalloc": ([_P, C.POINTER(_V), C.c_int64, _V, C.POINTER(Bins)], C.c_int),
    "capow_power": ([_V, C.POINTER(AssignStats), C.POINTER(Bins)], C.c_int),
    "capow_spectra": ([_P, C.POINTER(C.c_int32), C.POINTER(C.c_int32)], C.c_int),
    "capow_scalars": ([_P, C.POINTER(AssignStats), C.c_int32, _D, _D], C.c_int),
    "capow_finalize": ([_P, C.c_double, C.c_double, _D, _I64, _D, _I64,
                        C.POINTER(C.c_int32)], C.c_int),
    "capow_write_output": ([_P, C.c_double, C.c_double, _D, _I64, _D, _I64, C.c_char_p,
                            C.c_int32], C.c_int),
    "capow_write_sidecar": ([C.c_char_p, C.c_char_p, C.c_int32, C.c_char_p, C.c_int32], C.c_int),
    "capow_read_sidecar": ([C.c_char_p, C.c_char_p, C.c_int32, _D, _D, C.c_int32, C.c_int32,
                            C.POINTER(_V), C.POINTER(C.c_int64), _D], C.c_int),
    "capow_read_ascii": ([C.c_char_p, C.c_int32, _D, _D, C.c_int32, C.c_int32, C.POINTER(_V),
                          _I64, C.c_char_p, C.c_int32, _D], C.c_int),
    "capow_free_rows": ([_V], None),
    "capow_grid_elems": ([_V, _I64], C.c_int),
    "capow_get_grid": ([_V, C.c_int32, _V, C.c_int64], C.c_int),
    "capow_set_grid": ([_V, C.c_int32, _V, C.c_int64], C.c_int),
    "capow_mark_assigned": ([_V, C.POINTER(AssignStats)], C.c_int),
    "capow_mark_transformed": ([_V], C.c_int),
    "capow_survey_power": ([_V, _D, _D, _I64], C.c_int),
    "capow_kernel_times": ([_V, C.c_char_p, C.c_int32, _D, C.POINTER(C.c_int32), C.c_int32,
                            C.POINTER(C.c_int32)], C.c_int),
    "capow_reset_kernel_times": ([_V], C.c_int),
    "capow_backward": ([_V, _D, _V, _V, C.c_int32], C.c_int),
    "capow_keep_forward": ([_V, C.c_int32], C.c_int),
    "capow_backward_reused": ([_V, C.POINTER(C.c_int32)], C.c_int),
    "capow_batch_create": ([_P, C.c_int32, C.c_int64, C.c_int32, C.POINTER(_V)], C.c_int),
    "capow_batch_run": ([_V, C.POINTER(Mock), C.c_int32, C.POINTER(Bins),
                         C.POINTER(AssignStats), C.POINTER(C.c_float)], C.c_int),
    "capow_batch_destroy": ([_V], C.c_int),
    "capow_batch_set_graph": ([_V, C.c_int32], C.c_int),
    "capow_batch_set_stack": ([_V, C.c_int32], C.c_int),
    "capow_autotune": ([_V, C.c_int32, C.c_char_p, C.c_int32], C.c_int),
    "capow_batch_autotune": ([_V, C.POINTER(Mock), C.c_int32, C.c_int32, C.c_char_p, C.c_int32],
                             C.c_int),
    "capow_batch_last_error": ([_V], C.c_char_p),
    "capow_cspline_init": ([C.c_int32, _D, _D, _D], C.c_int),
    "capow_cspline_eval": ([C.c_int32, _D, _D, _D, C.c_double, _D], C.c_int),
    "capow_zr_table": ([C.c_double, C.c_int32, C.c_double, _D, _D], C.c_int),
    "capow_cosmo_create": ([C.c_double, C.POINTER(_V)], C.c_int),
    "capow_cosmo_destroy": ([_V], C.c_int),
    "capow_cosmo_r_of_z": ([_V, C.c_double, _D], C.c_int),
    "capow_cosmo_z_of_r": ([_V, C.c_double, _D], C.c_int),
    "capow_read_survey": ([C.c_char_p, C.c_int32, _P, C.POINTER(SurveyParams), _V, C.c_int64,
                           C.POINTER(C.POINTER(SurveyObj)), _I64, _I64, _D, _D, C.c_char_p,
                           C.c_int32, _D], C.c_int),
    "capow_free_survey": ([C.POINTER(SurveyObj)], None),
    "capow_survey_nbar_bins": ([C.POINTER(SurveyParams)], C.c_int32),
    "capow_survey_nbar": ([C.POINTER(SurveyParams), _V, _D, C.POINTER(SurveyObj), C.c_int64,
                           C.POINTER(SurveyObj), C.c_int64, _D, _D, C.c_int32], C.c_int),
    "capow_survey_fkp": ([C.POINTER(SurveyParams), C.POINTER(SurveyObj), C.c_int64], C.c_int),
    "capow_survey_sums": ([C.POINTER(SurveyParams), C.POINTER(SurveyObj), C.c_int64,
                           C.POINTER(SurveyObj), C.c_int64, C.POINTER(SurveyTotals)], C.c_int),
    "capow_survey_pack": ([C.POINTER(SurveyParams), C.POINTER(SurveyObj), C.c_int64,
                           C.POINTER(SurveyObj), C.c_int64, C.c_double, C.c_int32, _V, _V,
                           _I64], C.c_int),
}

_lib = None


def _init_torch_device():
    """torch ships its own HIP runtime beside the system one libcapow links;
    in one process both work only when torch's initializes the GPU first, so
    the first use of the library does that (a no-op without torch or GPU)."""
    try:
        import torch
    except ImportError:
        return
    try:
        if torch.cuda.is_available():
            torch.cuda.init()
    except Exception:
        pass


def lib():
    """Loads libcapow.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        _init_torch_device()
        if not os.path.exists(LIB_PATH):
            raise ImportError("libcapow.so not built: run __graft_entry__.build() or "
                              "python -m capow_amd.build_lib")
        L = C.CDLL(LIB_PATH)
        for name, (args, res) in SIGNATURES.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _lib = L
    return _lib


DEFAULTS = dict(nx=64, ny=64, nz=64, xmin=0.0, xmax=1000.0, ymin=0.0, ymax=1000.0,
                zmin=0.0, zmax=1000.0, interp_order=2, do_interlacing=0, los=0, momentum=0,
                nk=40, nmu=1, output_2d=0, output_log=0, odd_multipoles=0, mink=0.0,
                maxk=0.4, omega_m=0.3, redshift=0.5, precision=FP64,
                edge_mode=EDGE_REFERENCE, extra_spectra=0, reserved=0)


def make_params(**kw):
    d = dict(DEFAULTS)
    unknown = set(kw) - set(d)
    if unknown:
        raise TypeError("unknown parameters: %s" % sorted(unknown))
    d.update(kw)
    return Params(**d)


def spectra_codes(p):
    n = C.c_int32()
    codes = (C.c_int32 * 3)()
    _check(None, lib().capow_spectra(C.byref(p), C.byref(n), codes))
    return list(codes[:n.value])


def _check(ctx, rc):
    if rc != CAPOW_OK:
        msg = lib().capow_last_error(ctx)
        raise CapowError(rc, msg.decode() if msg else "")


def _dptr(a):
    return a.ctypes.data_as(_D)


def _iptr(a):
    return a.ctypes.data_as(_I64)


class PowerSpectrum:
    """One context: device grids, sort buffers and tables for fixed params."""

    def __init__(self, params=None, **kw):
        self.p = params if params is not None else make_params(**kw)
        self.ctx = _V()
        _check(None, lib().capow_create(C.byref(self.p), C.byref(self.ctx)))
        self.fp64 = self.p.precision == FP64
        self.dtype = np.float64 if self.fp64 else np.float32
        self.codes = spectra_codes(self.p)
        self._keep = None
        self.stats = None

    def close(self):
        if self.ctx:
            lib().capow_destroy(self.ctx)
            self.ctx = _V()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def check(self, rc):
        _check(self.ctx, rc)

    def set_timing(self, on=True):
        self.check(lib().capow_set_timing(self.ctx, 1 if on else 0))

    def set_stream(self, stream_ptr):
        self.check(lib().capow_set_stream(self.ctx, _V(stream_ptr)))

    def set_particles(self, xyzv, w=None):
        """xyzv: (n, 4) rows of x y z v_los; numpy (copied) or a torch CUDA
        tensor (used in place)."""
        dt = 1 if self.fp64 else 0
        # every call starts a new generation (the autograd wrapper checks it)
        self.generation = getattr(self, "generation", 0) + 1
        if hasattr(xyzv, "data_ptr") and getattr(xyzv, "is_cuda", False):
            assert xyzv.is_contiguous() and xyzv.shape[-1] == 4
            wp = _V(w.data_ptr()) if w is not None else None
            self._keep = (xyzv, w)
            self.check(lib().capow_set_particles(self.ctx, _V(xyzv.data_ptr()), wp,
                                                 xyzv.shape[0], dt, DEVICE))
        else:
            a = np.ascontiguousarray(xyzv, dtype=self.dtype)
            assert a.ndim == 2 and a.shape[1] == 4
            wa = None if w is None else np.ascontiguousarray(w, dtype=self.dtype)
            self._keep = (a, wa)
            self.check(lib().capow_set_particles(
                self.ctx, a.ctypes.data_as(_V), wa.ctypes.data_as(_V) if wa is not None else None,
                a.shape[0], dt, HOST))

    def assign(self):
        st = AssignStats()
        self.check(lib().capow_assign(self.ctx, C.byref(st)))
        self.stats = st
        return st

    def fft(self):
        self.check(lib().capow_fft(self.ctx))

    def _bins_struct(self):
        p = self.p
        ns = len(self.codes)
        nmu = p.nmu if p.output_2d else 1
        out = dict(pk=np.zeros(ns * 5 * p.nk), nmodes=np.zeros(p.nk, dtype=np.int64),
                   pk2d=np.zeros(ns * p.nk * nmu), nmodes2d=np.zeros(p.nk * nmu, dtype=np.int64))
        b = Bins(0, (C.c_int32 * 3)(), _dptr(out["pk"]), _iptr(out["nmodes"]),
                 _dptr(out["pk2d"]), _iptr(out["nmodes2d"]))
        return b, out

    def bin(self):
        """Raw per-bin sums: dict code -> {'pk': (5, nk), 'pk2d'}, plus 'nmodes'."""
        b, out = self._bins_struct()
        self.check(lib().capow_bin(self.ctx, C.byref(b)))
        return self._unpack(b, out)

    def _unpack(self, b, out):
        p = self.p
        ns = b.n_spectra
        nmu = p.nmu if p.output_2d else 1
        res = {"nmodes": out["nmodes"], "nmodes2d": out["nmodes2d"].reshape(p.nk, nmu),
               "codes": list(b.codes[:ns])}
        pk = out["pk"].reshape(ns, 5, p.nk)
        pk2d = out["pk2d"].reshape(ns, p.nk, nmu)
        for s in range(ns):
            res[b.codes[s]] = {"pk": pk[s].copy(), "pk2d": pk2d[s].copy()}
        return res

    def power(self):
        st = AssignStats()
        b, out = self._bins_struct()
        self.check(lib().capow_power(self.ctx, C.byref(st), C.byref(b)))
        self.stats = st
        return self._unpack(b, out)

    def backward(self, dL_draw, want_w=True):
        """dL/d(raw bins) (array [n_spectra][5][nk], PowerSpectrum.bin()'s
        order) -> (dL/d(x, y, z, v_los) as (n, 4), dL/dw as (n,) or None) for
        the particles last set: torch tensors on the device for device
        particles, numpy arrays otherwise.  Recomputes the forward unless
        keep_forward(True) kept it (see backward_reused())."""
        G = np.ascontiguousarray(dL_draw, dtype=np.float64).ravel()
        ns = len(self.codes)
        if G.size != ns * 5 * self.p.nk:
            raise ValueError("dL_draw must hold n_spectra * 5 * nk values")
        xyzv = self._keep[0]
        n = xyzv.shape[0]
        if hasattr(xyzv, "data_ptr") and getattr(xyzv, "is_cuda", False):
            import torch
            gx = torch.empty((n, 4), dtype=xyzv.dtype, device=xyzv.device)
            gw = torch.empty((n,), dtype=xyzv.dtype, device=xyzv.device) if want_w else None
            self.check(lib().capow_backward(self.ctx, _dptr(G), _V(gx.data_ptr()),
                                            _V(gw.data_ptr()) if gw is not None else None, DEVICE))
            return gx, gw
        gx = np.empty((n, 4), dtype=self.dtype)
        gw = np.empty((n,), dtype=self.dtype) if want_w else None
        self.check(lib().capow_backward(self.ctx, _dptr(G), gx.ctypes.data_as(_V),
                                        gw.ctypes.data_as(_V) if gw is not None else None, HOST))
        return gx, gw

    def keep_forward(self, on=True):
        """capow_keep_forward: the forward keeps its sort (with the rows'
        input indices) and transform for the next backward."""
        self.check(lib().capow_keep_forward(self.ctx, 1 if on else 0))

    def backward_reused(self):
        """True when the last backward reused the forward."""
        r = C.c_int32()
        self.check(lib().capow_backward_reused(self.ctx, C.byref(r)))
        return bool(r.value)

    def grid_elems(self):
        n = C.c_int64()
        self.check(lib().capow_grid_elems(self.ctx, C.byref(n)))
        return n.value

    def get_grid(self, slot):
        a = np.empty(self.grid_elems(), dtype=self.dtype)
        self.check(lib().capow_get_grid(self.ctx, slot, a.ctypes.data_as(_V), a.size))
        return a

    def set_grid(self, slot, arr):
        a = np.ascontiguousarray(arr, dtype=self.dtype).ravel()
        self.check(lib().capow_set_grid(self.ctx, slot, a.ctypes.data_as(_V), a.size))

    def mark_assigned(self, stats=None):
        self.check(lib().capow_mark_assigned(self.ctx, C.byref(stats) if stats else None))

    def mark_transformed(self):
        self.check(lib().capow_mark_transformed(self.ctx))

    def survey_power(self, origin):
        """Survey (Yamamoto) estimator on the real-space grids (capow_survey_power):
        {'pk': (5, nk) raw sums P0..P4, 'nmodes': (nk,)}."""
        o = np.ascontiguousarray(origin, dtype=np.float64)
        nk = self.p.nk
        pk = np.zeros(5 * nk)
        nm = np.zeros(nk, dtype=np.int64)
        self.check(lib().capow_survey_power(self.ctx, _dptr(o), _dptr(pk), _iptr(nm)))
        return {"pk": pk.reshape(5, nk), "nmodes": nm}

    def synchronize(self):
        self.check(lib().capow_synchronize(self.ctx))

    def kernel_times(self):
        names = C.create_string_buffer(4096)
        ms = (C.c_double * 64)()
        la = (C.c_int32 * 64)()
        n = C.c_int32()
        self.check(lib().capow_kernel_times(self.ctx, names, 4096, ms, la, 64, C.byref(n)))
        labels = names.raw.split(b"\0")
        return {labels[i].decode(): (ms[i], la[i]) for i in range(n.value)}

    def reset_kernel_times(self):
        self.check(lib().capow_reset_kernel_times(self.ctx))

    def autotune(self, reps=3):
        """Times the tunables on the particles set (capow_autotune); returns
        the report text (every timing and the choice)."""
        buf = C.create_string_buffer(8192)
        self.check(lib().capow_autotune(self.ctx, int(reps), buf, 8192))
        return buf.value.decode()


def scalars(p, stats, code):
    shot, norm = C.c_double(), C.c_double()
    _check(None, lib().capow_scalars(C.byref(p), C.byref(stats), code, C.byref(shot),
                                     C.byref(norm)))
    return shot.value, norm.value


def finalize(p, shot, norm, pk5, nmodes, pk2d=None, nmodes2d=None):
    """In place on copies; returns (pk5, nmodes, pk2d, nmodes2d, last_bin)."""
    pk5 = np.ascontiguousarray(pk5, dtype=np.float64).copy()
    nmodes = np.ascontiguousarray(nmodes, dtype=np.int64).copy()
    pk2d = None if pk2d is None else np.ascontiguousarray(pk2d, dtype=np.float64).copy()
    nmodes2d = None if nmodes2d is None else np.ascontiguousarray(nmodes2d, dtype=np.int64).copy()
    last = C.c_int32()
    _check(None, lib().capow_finalize(C.byref(p), shot, norm, _dptr(pk5), _iptr(nmodes),
                                      _dptr(pk2d) if pk2d is not None else None,
                                      _iptr(nmodes2d) if nmodes2d is not None else None,
                                      C.byref(last)))
    return pk5, nmodes, pk2d, nmodes2d, last.value


def write_output(p, shot, norm, pk5, nmodes, path, pk2d=None, nmodes2d=None, precise=False):
    pk5 = np.ascontiguousarray(pk5, dtype=np.float64)
    nmodes = np.ascontiguousarray(nmodes, dtype=np.int64)
    _check(None, lib().capow_write_output(
        C.byref(p), shot, norm, _dptr(pk5), _iptr(nmodes),
        _dptr(np.ascontiguousarray(pk2d, dtype=np.float64)) if pk2d is not None else None,
        _iptr(np.ascontiguousarray(nmodes2d, dtype=np.int64)) if nmodes2d is not None else None,
        path.encode(), 1 if precise else 0))


def write_sidecar(text_path, bin_path, nthreads=0):
    """capow_write_sidecar: parse the periodic ASCII catalogue at text_path
    and write its binary sidecar to bin_path."""
    bad = C.create_string_buffer(2048)
    rc = lib().capow_write_sidecar(os.fsencode(text_path), os.fsencode(bin_path), nthreads, bad,
                                   len(bad))
    if rc != CAPOW_OK:
        msg = ("error reading file: %s" % bad.value.decode(errors="replace")
               if rc == CAPOW_ERR_PARSE else "cannot write the sidecar of %s" % text_path)
        raise CapowError(rc, msg)


def read_sidecar(bin_path, text_path=None, los=0, lo=(0.0, 0.0, 0.0), hi=(1000.0, 1000.0, 1000.0),
                 precision=FP64, nthreads=0):
    """capow_read_sidecar: the rows read_ascii gives for the text catalogue,
    from its binary sidecar (refused as stale when text_path is given and
    the text changed since the sidecar was written)."""
    lo_a = (C.c_double * 3)(*lo)
    hi_a = (C.c_double * 3)(*hi)
    rows = _V()
    n = C.c_int64()
    xyz = (C.c_double * 3)()
    fp64 = precision == FP64
    rc = lib().capow_read_sidecar(os.fsencode(bin_path),
                                  os.fsencode(text_path) if text_path is not None else None, los,
                                  lo_a, hi_a, 1 if fp64 else 0, nthreads, C.byref(rows),
                                  C.byref(n), xyz)
    if rc != CAPOW_OK:
        if rc == CAPOW_ERR_OUT_OF_BOX:
            msg = "particle out of grid bounds: x=%f, y=%f, z=%f" % tuple(xyz)
        elif rc == CAPOW_ERR_STALE:
            msg = "sidecar %s is older than %s" % (bin_path, text_path)
        elif rc == CAPOW_ERR_PARSE:
            msg = "%s is not a catalogue sidecar" % bin_path
        else:
            msg = "cannot read %s" % bin_path
        raise CapowError(rc, msg)
    try:
        dt = np.float64 if fp64 else np.float32
        count = n.value * 4
        buf = (C.c_char * max(count * np.dtype(dt).itemsize, 1)).from_address(rows.value)
        out = np.frombuffer(buf, dtype=dt, count=count).reshape(n.value, 4).copy()
    finally:
        lib().capow_free_rows(rows)
    return out


def read_ascii(path, los=0, lo=(0.0, 0.0, 0.0), hi=(1000.0, 1000.0, 1000.0), precision=FP64,
               nthreads=0):
    """The periodic ASCII catalogue at `path` as an (n, 4) array of
    (x, y, z, v_los) rows (capow_read_ascii: the reference's fgets/sscanf
    semantics, parsed over `nthreads` threads).  Raises CapowError with the
    first malformed line, or the first particle outside [lo, hi] for LOS 0."""
    lo_a = (C.c_double * 3)(*lo)
    hi_a = (C.c_double * 3)(*hi)
    rows = _V()
    n = C.c_int64()
    bad = C.create_string_buffer(2048)
    xyz = (C.c_double * 3)()
    fp64 = precision == FP64
    rc = lib().capow_read_ascii(os.fsencode(path), los, lo_a, hi_a, 1 if fp64 else 0, nthreads,
                                C.byref(rows), C.byref(n), bad, len(bad), xyz)
    if rc != CAPOW_OK:
        if rc == CAPOW_ERR_OUT_OF_BOX:
            msg = "particle out of grid bounds: x=%f, y=%f, z=%f" % tuple(xyz)
        elif rc == CAPOW_ERR_PARSE:
            msg = "error reading file: %s" % bad.value.decode(errors="replace")
        else:
            msg = "cannot read %s" % path
        raise CapowError(rc, msg)
    try:
        dt = np.float64 if fp64 else np.float32
        count = n.value * 4
        buf = (C.c_char * (count * np.dtype(dt).itemsize)).from_address(rows.value)
        out = np.frombuffer(buf, dtype=dt, count=count).reshape(n.value, 4).copy()
    finally:
        lib().capow_free_rows(rows)
    return out


class Batch:
    """Many catalogues through one batch runner (capow_batch_*): catalogue i
    runs on stream i % nstreams with no host synchronisation between
    catalogues.  `configs` are parameter dicts or Params (e.g. one per grid
    size); catalogues are device tensors of (n, 4) rows x y z v_los."""

    def __init__(self, configs, max_particles, nstreams=4):
        ps = [c if isinstance(c, Params) else make_params(**c) for c in configs]
        self.configs = ps
        arr = (Params * len(ps))(*ps)
        h = _V()
        rc = lib().capow_batch_create(arr, len(ps), int(max_particles), int(nstreams), C.byref(h))
        if rc:
            raise CapowError(rc, lib().capow_batch_last_error(None).decode())
        self.h = h

    def close(self):
        if self.h:
            lib().capow_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_graph(self, on=True):
        """capow_batch_set_graph: capture a catalogue list into one HIP graph
        after its first run and replay it while the list is unchanged."""
        rc = lib().capow_batch_set_graph(self.h, 1 if on else 0)
        if rc:
            raise CapowError(rc, lib().capow_batch_last_error(self.h).decode())

    def set_stack(self, m):
        """capow_batch_set_stack: run each parameter set's catalogues in
        stacks of up to m, one launch per stage per stack (0: one catalogue
        per launch)."""
        rc = lib().capow_batch_set_stack(self.h, int(m))
        if rc:
            raise CapowError(rc, lib().capow_batch_last_error(self.h).decode())

    def autotune(self, catalogues, configs, reps=3):
        """capow_batch_autotune on these catalogues; returns the report."""
        n = len(catalogues)
        mocks = (Mock * n)()
        for i, (t, q) in enumerate(zip(catalogues, configs)):
            mocks[i] = Mock(t.data_ptr(), None, int(t.shape[0]), int(q), 0)
        buf = C.create_string_buffer(16384)
        rc = lib().capow_batch_autotune(self.h, mocks, n, int(reps), buf, 16384)
        if rc:
            raise CapowError(rc, lib().capow_batch_last_error(self.h).decode())
        return buf.value.decode()

    def run(self, catalogues, configs, weights=None):
        """catalogues: list of device tensors; configs: list of config indices.
        Returns (results, stats, ms): per-mock raw bins in PowerSpectrum.bin()'s
        form, AssignStats and GPU ms per mock."""
        n = len(catalogues)
        mocks = (Mock * n)()
        outs, bins = [], (Bins * n)()
        keep = []
        for i, (t, q) in enumerate(zip(catalogues, configs)):
            w = weights[i] if weights is not None else None
            mocks[i] = Mock(t.data_ptr(), w.data_ptr() if w is not None else None,
                            int(t.shape[0]), int(q), 0)
            p = self.configs[q]
            ns = C.c_int32()
            codes = (C.c_int32 * 3)()
            lib().capow_spectra(C.byref(p), C.byref(ns), codes)
            nmu = p.nmu if p.output_2d else 1
            out = {"pk": np.zeros(ns.value * 5 * p.nk), "nmodes": np.zeros(p.nk, dtype=np.int64),
                   "pk2d": np.zeros(ns.value * p.nk * nmu),
                   "nmodes2d": np.zeros(p.nk * nmu, dtype=np.int64)}
            keep.append(out)
            bins[i] = Bins(0, (C.c_int32 * 3)(), _dptr(out["pk"]), _iptr(out["nmodes"]),
                           _dptr(out["pk2d"]) if p.output_2d else None,
                           _iptr(out["nmodes2d"]) if p.output_2d else None)
        stats = (AssignStats * n)()
        ms = (C.c_float * n)()
        rc = lib().capow_batch_run(self.h, mocks, n, bins, stats, ms)
        if rc:
            raise CapowError(rc, lib().capow_batch_last_error(self.h).decode())
        res = []
        for i in range(n):
            p = self.configs[configs[i]]
            b, out = bins[i], keep[i]
            nsp = b.n_spectra
            nmu = p.nmu if p.output_2d else 1
            r = {"nmodes": out["nmodes"], "nmodes2d": out["nmodes2d"].reshape(p.nk, nmu),
                 "codes": list(b.codes[:nsp])}
            pk = out["pk"].reshape(nsp, 5, p.nk)
            pk2d = out["pk2d"].reshape(nsp, p.nk, nmu)
            for s in range(nsp):
                r[b.codes[s]] = {"pk": pk[s].copy(), "pk2d": pk2d[s].copy()}
            res.append(r)
        return res, list(stats), np.array(ms[:], dtype=np.float64)
