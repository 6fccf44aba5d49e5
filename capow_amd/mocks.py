"""Seeded synthetic catalogues for tests and benchmarks.

Rows are (x, y, z, vx, vy, vz) like the reference's periodic ASCII format
(src/read_data.c:64).  Two constructions:

uniform(n, L, seed)
    positions uniform in [0, L)^3 (Philox4x32 stream), zero velocities.
zeldovich(nl, L, seed)
    an nl^3 lattice displaced by Zel'dovich displacements of a Gaussian
    random field with P(k) = amp (k/k0)^slope, zero at k = 0 and above the
    lattice Nyquist frequency; velocities = f aH psi.  White noise w(q) is
    drawn on the lattice, delta_k = rfftn(w) sqrt(P(k)/V_cell),
    psi_k = i k delta_k / k^2, psi = irfftn(psi_k).
Both backends (numpy on the host, torch on a GPU) implement the same
construction; numbers differ between backends (different generators).
"""
import numpy as np

AMP, K0, SLOPE = 2.0e4, 0.1, -1.5
F_GROWTH, AH = 0.5, 100.0


def _wrap(x, L):
    x = np.mod(x, L)
    x[x >= L] = 0.0
    return x


def uniform(n, L=1000.0, seed=1, dtype=np.float64):
    rng = np.random.Generator(np.random.Philox(seed))
    pos = rng.random((n, 3)) * L
    pos = _wrap(pos, L).astype(dtype)
    pos[pos >= dtype(L)] = 0
    vel = np.zeros((n, 3), dtype=dtype)
    return pos, vel


def _kgrid(nl, L):
    kf = 2 * np.pi / L
    k1 = np.fft.fftfreq(nl, d=1.0 / nl) * kf
    kz = np.fft.rfftfreq(nl, d=1.0 / nl) * kf
    return k1, kz


def zeldovich(nl, L=1000.0, seed=1, dtype=np.float64, amp=AMP, slope=SLOPE):
    """Numpy construction on the host (for sizes up to ~256^3)."""
    rng = np.random.Generator(np.random.Philox(seed))
    w = rng.standard_normal((nl, nl, nl))
    wk = np.fft.rfftn(w)
    del w
    k1, kz = _kgrid(nl, L)
    kx = k1[:, None, None]
    ky = k1[None, :, None]
    kzz = kz[None, None, :]
    k2 = kx * kx + ky * ky + kzz * kzz
    k = np.sqrt(k2)
    vcell = (L / nl) ** 3
    knyq = np.pi * nl / L
    with np.errstate(divide="ignore", invalid="ignore"):
        pk = np.where((k > 0) & (k <= knyq), amp * (k / K0) ** slope, 0.0)
        dk = wk * np.sqrt(pk / vcell)
        inv = np.where(k2 > 0, 1.0 / k2, 0.0)
    q = (np.arange(nl) + 0.5) * (L / nl)
    pos = np.empty((nl, nl, nl, 3))
    vel = np.empty((nl, nl, nl, 3))
    for a, ka in enumerate((kx, ky, kzz)):
        psi = np.fft.irfftn(1j * ka * dk * inv, s=(nl, nl, nl), axes=(0, 1, 2))
        qa = np.broadcast_to(q.reshape([nl if b == a else 1 for b in range(3)]), (nl, nl, nl))
        pos[..., a] = qa + psi
        vel[..., a] = F_GROWTH * AH * psi
    pos = _wrap(pos.reshape(-1, 3), L).astype(dtype)
    pos[pos >= dtype(L)] = 0
    return pos, vel.reshape(-1, 3).astype(dtype)


def zeldovich_torch(nl, L=1000.0, seed=1, device="cuda", dtype=None, amp=AMP, slope=SLOPE):
    """Same construction on a GPU with torch (benchmark inputs at 512^3).
    Returns (n, 4) rows x y z v_z on the device (LOS = z)."""
    import torch
    dtype = dtype or torch.float32
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    w = torch.randn((nl, nl, nl), generator=g, device=device, dtype=torch.float64)
    wk = torch.fft.rfftn(w)
    del w
    kf = 2 * np.pi / L
    k1 = torch.fft.fftfreq(nl, d=1.0 / nl, device=device, dtype=torch.float64) * kf
    kz = torch.fft.rfftfreq(nl, d=1.0 / nl, device=device, dtype=torch.float64) * kf
    kx, ky, kzz = k1[:, None, None], k1[None, :, None], kz[None, None, :]
    k2 = kx * kx + ky * ky + kzz * kzz
    k = torch.sqrt(k2)
    knyq = np.pi * nl / L
    vcell = (L / nl) ** 3
    pk = torch.where((k > 0) & (k <= knyq), amp * (k / K0) ** slope, torch.zeros_like(k))
    dk = wk * torch.sqrt(pk / vcell)
    del wk, pk, k
    inv = torch.where(k2 > 0, 1.0 / k2, torch.zeros_like(k2))
    dk = dk * inv
    del inv, k2
    q = (torch.arange(nl, device=device, dtype=torch.float64) + 0.5) * (L / nl)
    out = torch.empty((nl, nl, nl, 4), device=device, dtype=dtype)
    for a, ka in enumerate((kx, ky, kzz)):
        psi = torch.fft.irfftn(1j * ka * dk, s=(nl, nl, nl))
        shape = [nl if b == a else 1 for b in range(3)]
        x = torch.remainder(q.reshape(shape) + psi, L)
        x = x.to(dtype)
        x = torch.where(x >= L, torch.zeros_like(x), x)
        out[..., a] = x
        if a == 2:
            out[..., 3] = (F_GROWTH * AH * psi).to(dtype)
        del psi, x
    return out.reshape(-1, 4)


def xyzv(pos, vel, los=0):
    """(n, 4) rows x y z v_los with the reference's column choice
    (src/read_data.c:65-76: vz for LOS 0 and 3, vx for 1, vy for 2)."""
    col = {0: 2, 1: 0, 2: 1, 3: 2}[los]
    return np.concatenate([pos, vel[:, col:col + 1]], axis=1)


def lognormal_counts(nmocks, seed, median=5e5, sigma_ln=0.6, lo=1e5, hi=3e6):
    """Per-mock particle counts: lognormal with the given median and sigma of
    ln n, clipped to [lo, hi] (seeded, numpy Philox)."""
    rng = np.random.Generator(np.random.Philox(seed))
    n = np.exp(np.log(median) + sigma_ln * rng.standard_normal(nmocks))
    return np.clip(np.round(n), lo, hi).astype(np.int64)


def mixed_mock_torch(n, L, seed, device="cuda", dtype=None, frac_clustered=0.5):
    """One mock of n points on the GPU: a Zel'dovich lattice of nl^3 points
    (power law as above; nl = (frac_clustered n)^(1/3) rounded to a multiple
    of 8, so that a few FFT sizes serve every mock) plus uniform points for
    the rest (zero velocity).  Rows x y z v_z."""
    import torch
    dtype = dtype or torch.float32
    nl = int(round((frac_clustered * n) ** (1.0 / 3.0) / 8.0)) * 8
    nl = max(8, min(nl, int(n ** (1.0 / 3.0)) // 8 * 8 if n >= 512 else 8))
    z = zeldovich_torch(nl, L, seed=seed, device=device, dtype=dtype)
    nu = max(0, int(n) - z.shape[0])
    g = torch.Generator(device=device)
    g.manual_seed(seed + 7919)
    u = torch.rand((nu, 3), generator=g, device=device, dtype=torch.float64) * L
    u = torch.where(u >= L, torch.zeros_like(u), u).to(dtype)
    u = torch.cat([u, torch.zeros((nu, 1), device=device, dtype=dtype)], dim=1)
    return torch.cat([z, u], dim=0).contiguous()
