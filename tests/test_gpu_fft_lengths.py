"""GPU FFT at every register length against the oracle (reference FFT call
src/main.c:266-271, plans src/main.c:1200-1217).

The register kernels serve NX, NY and NZ/2 in {16, 32, 64, 96, 128, 192,
256, 384, 512, 1024} and, in FP32, 2048, each as a two-level split N1 x N2
(4x4 ... 32x32, 64x32; the 24-, 32- and 64-point in-register DFTs appear in
8x24, 16x24, 16x32, 32x32, 64x32).  FP64 takes the Stockham kernels at 2048
(and the backward, which needs register sizes, refuses it).
Thin grids put every length on every axis -- x (column pass), y (column
pass) and z (r2c row pass) -- so each kernel template the 256^3-2048^3
workloads run is compared with the oracle's FFT at least once, through
both FFT forms:

* out of place (three transposing row passes, the default when scratch
  fits), and
* in place (row pass + strided column passes; CAPOW_FFT_TRANSPOSED=0, the
  form the largest grids take),

in FP64 (<= 1e-12 of max|X|) and FP32 (<= 2e-5 of max|X|).  The adjoint's
inverse transform (conjugated column passes and the z-row adjoint) is
compared through capow_backward against the FP64 reference gradient on the
same grids (FP64 within 1e-9, FP32 within 2e-3 of max|gradient|).
"""
import sys

import numpy as np
import pytest

from capow_amd import api, mocks
from oracle import adjoint_ref as AR

pytestmark = pytest.mark.gpu

REG = (16, 32, 64, 96, 128, 192, 256, 384, 512, 1024, 2048)

# (NX, NY, NZ): x, y and z = NZ/2 lengths
SHAPES = [
    (1024, 16, 192),   # x 1024 (32x32), y 16, z 96 (8x12)
    (16, 1024, 384),   # y 1024, z 192 (8x24)
    (96, 16, 2048),    # x 96, z 1024 (32x32 r2c)
    (192, 384, 128),   # x 192, y 384 (16x24), z 64
    (384, 96, 64),     # x 384, y 96, z 32
    (512, 192, 32),    # x 512 (16x32), y 192, z 16
    (256, 512, 32),    # x 256, y 512, z 16
    (32, 256, 768),    # x 32, y 256 (16x16), z 384
    (64, 32, 1024),    # x 64, y 32, z 512
    (128, 64, 512),    # x 128, y 64, z 256
    (16, 128, 256),    # y 128 (8x16), z 128
    (2048, 16, 32),    # x 2048 (64x32), y 16, z 16
    (16, 2048, 64),    # y 2048, z 32
    (32, 16, 4096),    # z 2048 (64x32 r2c)
]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_shapes_cover_every_length_on_every_axis():
    for axis in range(3):
        got = {s[axis] if axis < 2 else s[2] // 2 for s in SHAPES}
        assert set(REG) <= got, (axis, sorted(set(REG) - got))


@pytest.mark.parametrize("form", ["transposed", "in_place"])
@pytest.mark.parametrize("prec", [api.FP64, api.FP32])
@pytest.mark.parametrize("shape", SHAPES, ids=["x".join(map(str, s)) for s in SHAPES])
def test_fft_length_matches_oracle(shape, prec, form, oracle, monkeypatch):
    nx, ny, nz = shape
    say("fft length case", shape, "fp64" if prec == api.FP64 else "fp32", form)
    if form == "in_place":
        monkeypatch.setenv("CAPOW_FFT_TRANSPOSED", "0")
    p = api.make_params(nx=nx, ny=ny, nz=nz, interp_order=1, precision=prec)
    ps = api.PowerSpectrum(p)
    rng = np.random.default_rng(nx + 3 * ny + 7 * nz)
    rowlen = 2 * (nz // 2 + 1)
    g = np.zeros((nx, ny, rowlen))
    # float32-representable input: both precisions transform the same numbers
    g[:, :, :nz] = rng.standard_normal((nx, ny, nz)).astype(np.float32)
    ps.set_grid(0, g)
    ps.mark_assigned()
    ps.fft()
    got = ps.get_grid(0).astype(np.float64).reshape(nx, ny, rowlen)
    ps.close()
    ref = g.ravel().copy()
    oracle.fft_r2c_3d(ref, nx, ny, nz)
    ref = ref.reshape(nx, ny, rowlen)
    scale = np.abs(ref).max()
    err = np.abs(got - ref).max() / scale
    say("  rel err", err)
    assert err <= (1e-12 if prec == api.FP64 else 2e-5)


@pytest.mark.parametrize("prec", [api.FP64, api.FP32])
@pytest.mark.parametrize("shape", SHAPES, ids=["x".join(map(str, s)) for s in SHAPES])
def test_inverse_fft_length_through_backward(shape, prec, oracle):
    """capow_backward = binning backward, inverse FFT (x, y column adjoints
    and the z-row adjoint at these lengths), transposed deposit."""
    nx, ny, nz = shape
    if prec == api.FP64 and max(nx, ny, nz // 2) > 1024:
        pytest.skip("FP64 register kernels stop at 1024; the backward needs register sizes")
    say("ifft length case", shape, "fp64" if prec == api.FP64 else "fp32")
    L = 1000.0
    rng = np.random.default_rng(nx * 5 + ny)
    n = 1500
    pos = rng.uniform(0, L, size=(n, 3)).astype(np.float32).astype(np.float64)
    vel = rng.normal(0, 100.0, size=(n, 3)).astype(np.float32).astype(np.float64)
    dmin = L / max(nx, ny, nz)
    nk = 10
    maxk = np.pi / dmin * 0.9
    kw = dict(nx=nx, ny=ny, nz=nz, xmax=L, ymax=L, zmax=L, interp_order=2, do_interlacing=1,
              momentum=1, nk=nk, maxk=maxk, odd_multipoles=1)
    G = np.random.default_rng(7).normal(size=(1, 5, nk))
    ref_x, ref_w = AR.gradient(oracle.params(**kw), pos, vel, None, G, 0)
    ps = api.PowerSpectrum(api.make_params(precision=prec, **kw))
    ps.set_particles(mocks.xyzv(pos, vel, 0))
    gx, gw = ps.backward(G)
    ps.close()
    tol = 1e-9 if prec == api.FP64 else 2e-3
    keep = np.ones(n, dtype=bool)
    if prec == api.FP32:
        # The FP32 form scales float32 coordinates in FP32: s = (x - MIN)/dx
        # carries an ulp of s (2.4e-4 cells at s ~ 2048), so a particle
        # that close to a cell edge (s, or s + 1/2 when interlaced, an
        # integer) can take the neighbouring cell, where the CIC weights'
        # derivative jumps; the FP64 reference, on the same float32
        # positions, does not.  Those particles are left out.
        for a, na in enumerate((nx, ny, nz)):
            sa = pos[:, a] / (L / na)
            for sh in (0.0, 0.5):
                fr = (sa + sh) % 1.0
                keep &= np.minimum(fr, 1.0 - fr) > 1e-3
        say("  particles within 1e-3 cells of an edge, left out:", int((~keep).sum()))
    for c in range(4):
        s = np.abs(ref_x[keep, c]).max()
        assert s > 0, c
        err = np.abs(gx[keep, c] - ref_x[keep, c]).max() / s
        say("  component", c, "max rel err", err)
        assert err <= tol, (c, err)
    err = np.abs(gw[keep] - ref_w[keep]).max() / np.abs(ref_w[keep]).max()
    say("  weights max rel err", err)
    assert err <= tol
