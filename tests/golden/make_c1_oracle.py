"""Writes tests/golden/c1_oracle.npz: the FP64 oracle's raw bins for the
512^3 / 16,777,216-particle Zel'dovich workload (256^3 lattice, L = 1000
Mpc/h, seed 1, TSC + interlacing, LOS 0, odd multipoles, dk = 0.01, NK =
160), spectra P00, P11 and P01 (Momentum 0, 2, 1).

The oracle takes minutes at this size on one core, so its output is kept
as a small fixture (raw sums, mode counts, assignment sums, shot and norm)
together with checksums of the catalogue it was computed from;
tests/test_gpu_configs.py rebuilds the same catalogue, checks the
checksums and compares the GPU's FP64 and FP32 forms with these numbers.

Run from the repository root:  python tests/golden/make_c1_oracle.py
"""
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from capow_amd import mocks  # noqa: E402
from oracle import pyoracle as O  # noqa: E402

NL, L, SEED, N, NK, DK = 256, 1000.0, 1, 512, 160, 0.01
CODES = (0, 2, 1)


def catalogue():
    return mocks.zeldovich(NL, L, SEED)


def checksums(pos, vel):
    return np.array([pos.sum(), (pos * pos).sum(), vel[:, 2].sum(), (vel[:, 2] ** 2).sum()])


def params(code):
    return O.params(nx=N, ny=N, nz=N, xmax=L, ymax=L, zmax=L, interp_order=3, do_interlacing=1,
                    los=0, momentum=code, nk=NK, maxk=NK * DK, odd_multipoles=1)


def main():
    pos, vel = catalogue()
    out = dict(checksums=checksums(pos, vel), codes=np.array(CODES))
    for code in CODES:
        t = time.time()
        rc, b, st, alpha, shot, norm = O.run(params(code), pos, vel)
        assert rc == 0
        out["raw_%d" % code] = np.stack([b.pk[l] for l in range(5)])
        out["nmodes_%d" % code] = b.nmodes.copy()
        out["scalars_%d" % code] = np.array([st.ngrid, st.nread, alpha, shot, norm])
        print("code %d: %.1f s" % (code, time.time() - t), flush=True)
    np.savez_compressed(os.path.join(os.path.dirname(os.path.abspath(__file__)), "c1_oracle.npz"),
                        **out)


if __name__ == "__main__":
    main()
