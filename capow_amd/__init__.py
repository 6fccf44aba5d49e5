"""MI355X (gfx950) implementation of CAPOW's periodic-box P_l(k) path.

The product is the C ABI in include/capow.h, implemented by the HIP kernels
in capow_amd/csrc (built into capow_amd/libcapow.so) and driven by the C
host program capow_amd/bin/capow.  `capow_amd.api` is a thin ctypes layer
for checking and benchmarking.
"""
from .api import (PowerSpectrum, CapowError, make_params, scalars, finalize,  # noqa: F401
                  write_output, FP32, FP64, EDGE_REFERENCE, EDGE_PERIODIC, SPEC_AUTO_A,
                  SPEC_AUTO_B)
