"""Pins the oracle FFT (stand-in for FFTW's forward r2c, reference
src/main.c:266-271) against the DFT definition, numpy's pocketfft and
Parseval's theorem."""
import numpy as np
import pytest


def _dft_matrix(n):
    j = np.arange(n)
    return np.exp(-2j * np.pi * np.outer(j, j) / n)


def _pad(r):
    nx, ny, nz = r.shape
    g = np.zeros((nx, ny, 2 * (nz // 2 + 1)))
    g[:, :, :nz] = r
    return g


def _unpad(g, nz):
    return g[:, :, 0::2] + 1j * g[:, :, 1::2]


@pytest.mark.parametrize("shape", [(4, 4, 4), (3, 5, 6), (8, 6, 7), (2, 12, 9)])
def test_r2c_matches_bruteforce_dft(shape, oracle):
    rng = np.random.default_rng(sum(shape))
    r = rng.standard_normal(shape)
    g = _pad(r).ravel()
    oracle.fft_r2c_3d(g, *shape)
    got = _unpad(g.reshape(shape[0], shape[1], -1), shape[2])
    Fx, Fy, Fz = (_dft_matrix(n) for n in shape)
    full = np.einsum("ai,bj,ck,ijk->abc", Fx, Fy, Fz, r)
    ref = full[:, :, : shape[2] // 2 + 1]
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()


@pytest.mark.parametrize("shape", [(16, 16, 16), (32, 24, 20), (30, 42, 50), (7, 11, 26)])
def test_r2c_matches_numpy(shape, oracle):
    rng = np.random.default_rng(7)
    r = rng.standard_normal(shape)
    g = _pad(r).ravel()
    oracle.fft_r2c_3d(g, *shape)
    got = _unpad(g.reshape(shape[0], shape[1], -1), shape[2])
    ref = np.fft.rfftn(r)
    assert np.abs(got - ref).max() <= 1e-13 * np.abs(ref).max() * np.log2(np.prod(shape))


def test_parseval(oracle):
    shape = (24, 24, 24)
    r = np.random.default_rng(3).standard_normal(shape)
    g = _pad(r).ravel()
    oracle.fft_r2c_3d(g, *shape)
    c = _unpad(g.reshape(24, 24, -1), 24)
    # half-space weights: kz = 0 and kz = N/2 planes once, others twice
    w = np.full(c.shape, 2.0)
    w[:, :, 0] = 1.0
    w[:, :, -1] = 1.0
    lhs = (np.abs(c) ** 2 * w).sum() / np.prod(shape)
    assert lhs == pytest.approx((r ** 2).sum(), rel=1e-12)


@pytest.mark.parametrize("n", [1, 2, 3, 5, 7, 8, 12, 49, 60, 97, 128, 210, 1024])
def test_c2c_1d(n, oracle):
    x = np.random.default_rng(n).standard_normal(n) + 1j * np.random.default_rng(n + 1).standard_normal(n)
    got = oracle.fft_c2c_1d(x)
    ref = np.fft.fft(x)
    assert np.abs(got - ref).max() <= 1e-13 * max(1.0, np.abs(ref).max()) * max(1, np.log2(n))
