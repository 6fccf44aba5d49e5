"""FP64 reference gradient of the raw binned sums -- TEST INFRASTRUCTURE ONLY.

The checker for capow_backward: the derivative of
    L = sum_{s, l, k} G[s][l][k] raw_{s,l}(k)
with respect to the particle positions, line-of-sight velocities and
weights, where raw_{s,l}(k) are the binned sums of compute_periodic_power
(reference src/main.c:302-468) of spectrum s for the catalogue.  The reference
has no derivatives; this restates the chain rule on the reference's forward:

* forward grids: the C oracle (orc_assign = add_to_grid at one rank with the
  halo add, src/read_data.c:32-179, 514-745; orc_mean_subtract =
  src/main.c:211-225), then numpy.fft.rfftn;
* binning geometry per mode with the reference's float casts
  (src/main.c:329-429): frequency, |k|, bin, mu, NM, window deconvolution,
  interlace phase;
* k-space gradients of the grids (dL/dRe + i dL/dIm), the inverse transform
  g(x) = Re sum over stored modes G(k) e^{+i k.x} (numpy ifft), the mean
  subtraction term and the transposed deposit with the derivatives of the
  NGP/CIC/TSC weights.

Scope: periodic box with LOS = 0 (no RSD; the catalogues of the gradient
configuration), reference edge rules.  Only tests import this module.
Pinned in tests/test_adjoint_ref.py by central finite differences of the C
oracle's forward (the mathematics fixes the gradient of a linear functional).
"""
import numpy as np

from . import pyoracle as O

AUTO_A, AUTO_B, CROSS_IM, CROSS_RE = 0, 1, 2, 3
_FA = (0, 0, 1, 0, 1, 2, 0, 1, 0)          # field A = w v^a per Momentum (src/read_data.c:91-130)
_FB = (-1, 1, -1, 2, 2, -1, 3, 3, 4)       # field B = w v^b (-1: none)


def _auto_code(e):
    return {0: 0, 1: 2, 2: 5}.get(e, -1)


def spectra(momentum, extra):
    """(code, kind) per spectrum, in the order of capow_spectra."""
    fa, fb = _FA[momentum], _FB[momentum]
    out = []
    if momentum in (0, 2, 5):
        out.append((momentum, AUTO_A))
    elif momentum in (1, 4, 6):
        out.append((momentum, CROSS_IM))
    else:
        out.append((momentum, CROSS_RE))
    if fb >= 0 and (extra & 1):
        out.append((_auto_code(fa), AUTO_A))
    if fb >= 0 and fb <= 2 and (extra & 2):
        out.append((_auto_code(fb), AUTO_B))
    return out


def axis_taps(u, lo, d, L, n, order, is_x):
    """Per-axis taps at LOS = 0 (periodic): first tap, weights, d weight /
    d s and acceptance, as add_to_grid (src/read_data.c:519-743)."""
    u = np.asarray(u, dtype=np.float64)
    hi = lo + L
    w = np.zeros((3,) + u.shape)
    dw = np.zeros((3,) + u.shape)
    if order == 3:
        if is_x:
            u = np.where(u - lo < d, u + L, u)
            ok = ~((u < lo + d) | (u >= hi + d))
        else:
            ok = ~((u < lo) | (u >= hi))
        s = (u - lo) / d
        c = np.trunc(s + 0.5).astype(np.int64)
        f = s - c
        keep0 = c != 1          # index-0 low tap dropped (the "> 0" guards)
        w[0] = np.where(keep0, 0.5 * (0.5 - f) ** 2, 0.0)
        w[1] = 0.75 - f * f
        w[2] = 0.5 * (0.5 + f) ** 2
        dw[0] = np.where(keep0, -(0.5 - f), 0.0)
        dw[1] = -2.0 * f
        dw[2] = 0.5 + f
        first = c - 1
    elif order == 2:
        ok = ~((u < lo) | (u >= hi))
        s = (u - lo) / d
        c = np.trunc(s).astype(np.int64)
        f = s - c
        w[0], w[1] = 1.0 - f, f
        dw[0], dw[1] = -1.0, 1.0
        first = c
    else:
        ok = ~((u < lo) | (u >= hi))
        s = (u - lo) / d
        c = np.trunc(s).astype(np.int64)
        if not is_x:
            c = np.where(c == n, n - 1, c)
        w[0] = 1.0
        first = c
    return first, w, dw, ok


def mode_geometry(p):
    """Arrays over the stored modes [NX][NY][NZ/2+1]: kbin, valid, mu, NM,
    cor (x0.25 with interlacing), phase (cos, sin)."""
    nx, ny, nz = p.nx, p.ny, p.nz
    nzh = nz // 2 + 1
    dx, dy, dz = (p.xmax - p.xmin) / nx, (p.ymax - p.ymin) / ny, (p.zmax - p.zmin) / nz
    sx, sy, sz = 1.0 / (float(nx) * dx), 1.0 / (float(ny) * dy), 1.0 / (float(nz) * dz)
    i = np.arange(nx, dtype=np.float64)
    j = np.arange(ny, dtype=np.float64)
    k = np.arange(nzh, dtype=np.float64)
    fx = np.where(np.arange(nx) <= nx // 2, i, i - nx) * sx
    fy = np.where(np.arange(ny) <= ny // 2, j, j - ny) * sy
    fz = k * sz
    FX, FY, FZ = np.meshgrid(fx, fy, fz, indexing="ij")
    fk = 2.0 * np.pi * np.sqrt((FX * FX + FY * FY) + FZ * FZ)
    bw = (p.maxk - p.mink) / p.nk
    kbin = np.trunc((fk - p.mink).astype(np.float32).astype(np.float64) / bw).astype(np.int64)
    mny = min(np.pi / dx, np.pi / dy, np.pi / dz)
    valid = (kbin >= 0) & (kbin < p.nk) & (fk <= mny)
    with np.errstate(divide="ignore", invalid="ignore"):
        mu = np.where(fk > 0, 2.0 * np.pi * FZ / fk, 0.0)
    NM = np.full(fk.shape, 2.0)
    I, J = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    one = np.where(J == 0, (I != 0) & (I != nx // 2), I != nx // 2)
    NM[:, :, 0] = np.where(one, 1.0, 2.0)
    ax, ay, az = np.pi * FX * dx, np.pi * FY * dy, np.pi * FZ * dz
    with np.errstate(divide="ignore", invalid="ignore"):
        snx = np.where(FX != 0, np.sin(ax) / ax, 1.0)
        sny = np.where(FY != 0, np.sin(ay) / ay, 1.0)
        snz = np.where(FZ != 0, np.sin(az) / az, 1.0)
    cor = (1.0 / (snx * sny * snz)) ** (2 * p.interp_order)
    if p.do_interlacing:
        cor = cor * 0.25
    kh = ax + ay + az
    return kbin, valid, mu, NM, cor, np.cos(kh), np.sin(kh)


def forward_kgrids(p, pos, vel, w):
    """The oracle's deposited, mean-subtracted grids (real [NX][NY][NZ]) and
    their r2c transforms; stats."""
    rc, grids, st = O.assign(p, pos, vel, w)
    assert rc == 0, rc
    alpha, _, _ = O.scalars(p, st)
    O.mean_subtract(p, alpha, grids)
    real = [None if g is None else O.real_view(p, g)[:p.nx, :, :p.nz].copy() for g in grids]
    kg = [None if r is None else np.fft.rfftn(r) for r in real]
    return real, kg, st


def raw_bins(p, pos, vel, w, extra):
    """Raw binned sums [n_spectra][5][nk] from the oracle's own binning
    (orc_compute_power), one spectrum code at a time as the reference."""
    rc, grids, st = O.assign(p, pos, vel, w)
    assert rc == 0
    alpha, _, _ = O.scalars(p, st)
    O.mean_subtract(p, alpha, grids)
    for g in grids:
        if g is not None:
            O.fft_r2c_3d(g, p.nx, p.ny, p.nz)
    out = []
    m0 = p.momentum
    for code, kind in spectra(m0, extra):
        q = O.params(**{f: getattr(p, f) for f, _ in p._fields_})
        q.momentum = code
        if code == m0:
            g = grids
        elif kind == AUTO_A:
            g = [grids[0], grids[1], None, None]
        else:
            g = [grids[2], grids[3], None, None]
        b = O.compute_power(q, g)
        out.append(np.stack([b.pk[l] for l in range(5)]))
    return np.stack(out)


def gradient(p, pos, vel, w, G, extra):
    """dL/d(x, y, z, v_los) (n, 4) and dL/dw (n,) for L = sum G raw."""
    pos = np.asarray(pos, dtype=np.float64)
    vel = np.asarray(vel, dtype=np.float64)
    n = len(pos)
    w = np.ones(n) if w is None else np.asarray(w, dtype=np.float64)
    v = vel[:, 2]                                   # LOS = 0: v_los = vz (src/read_data.c:65)
    fa, fb = _FA[p.momentum], _FB[p.momentum]
    real, kg, st = forward_kgrids(p, pos, vel, w)
    kbin, valid, mu, NM, cor, cph, sph = mode_geometry(p)
    inter = bool(p.do_interlacing)
    A = kg[0]
    Ai = kg[1] if inter else 0.0
    B = kg[2] if kg[2] is not None else 0.0
    Bi = kg[3] if (kg[3] is not None) else 0.0
    phi = cph + 1j * sph
    delta = A + phi * Ai if inter else A
    beta = (B + phi * Bi) if inter else B
    mu2 = mu * mu
    Lw = [np.ones_like(mu), mu, 1.5 * mu2 - 0.5, 2.5 * mu2 * mu - 1.5 * mu,
          4.375 * mu2 * mu2 - 3.75 * mu2 + 0.375]
    kb = np.clip(kbin, 0, p.nk - 1)
    ls = (0, 2, 4, 1, 3) if p.odd_multipoles else (0, 2, 4)
    dd = np.zeros(A.shape, dtype=np.complex128)
    db = np.zeros(A.shape, dtype=np.complex128)
    for s, (code, kind) in enumerate(spectra(p.momentum, extra)):
        hs = np.zeros(A.shape)
        for l in ls:
            hs += G[s][l][kb] * Lw[l]
        h = np.where(valid, NM * cor * hs, 0.0)
        if kind == AUTO_A:
            dd += 2.0 * h * delta
        elif kind == AUTO_B:
            db += 2.0 * h * beta
        elif kind == CROSS_IM:
            dd += -1j * h * beta
            db += 1j * h * delta
        else:
            dd += h * beta
            db += h * delta
    gk = [dd, np.conj(phi) * dd if inter else None, db if kg[2] is not None else None,
          np.conj(phi) * db if (inter and kg[2] is not None) else None]

    def inv(g):
        nx, ny, nz = p.nx, p.ny, p.nz
        H = np.fft.ifft(np.fft.ifft(g, axis=0), axis=1) * (nx * ny)
        Hz = np.zeros((nx, ny, nz), dtype=np.complex128)
        Hz[:, :, :nz // 2 + 1] = H
        return np.real(np.fft.ifft(Hz, axis=2)) * nz

    gr = [None if g is None else inv(g) for g in gk]
    ncell = p.nx * p.ny * p.nz
    dalpha = 0.0
    if fa == 0:
        dalpha = -(gr[0].sum() + (gr[1].sum() if inter else 0.0))
    lo = np.array([p.xmin, p.ymin, p.zmin])
    Lb = np.array([p.xmax - p.xmin, p.ymax - p.ymin, p.zmax - p.zmin])
    nn = (p.nx, p.ny, p.nz)
    d = Lb / np.array(nn)
    FA = w * v ** fa
    FB = w * v ** fb if fb >= 0 else np.zeros(n)
    dfa = np.zeros(n)
    dfb = np.zeros(n)
    ds = np.zeros((3, n))
    main_ok = np.zeros(n, dtype=bool)
    for ph in range(2 if inter else 1):
        u = pos + (d / 2 if ph else 0.0)
        taps = [axis_taps(u[:, a], lo[a], d[a], Lb[a], nn[a], p.interp_order, a == 0)
                for a in range(3)]
        ok = taps[0][3] & taps[1][3] & taps[2][3]
        if ph == 0:
            main_ok = ok
        ga = gr[ph]
        gb = gr[2 + ph] if gr[2] is not None else None
        o = p.interp_order
        for tx in range(o):
            ix = (taps[0][0] + tx) % nn[0]
            for ty in range(o):
                iy = (taps[1][0] + ty) % nn[1]
                for tz in range(o):
                    iz = (taps[2][0] + tz) % nn[2]
                    a = ga[ix, iy, iz]
                    b = gb[ix, iy, iz] if gb is not None else 0.0
                    wx, wy, wz = taps[0][1][tx], taps[1][1][ty], taps[2][1][tz]
                    Wt = np.where(ok, wx * wy * wz, 0.0)
                    dfa += a * Wt
                    dfb += b * Wt
                    fv = np.where(ok, FA * a + FB * b, 0.0)
                    ds[0] += fv * taps[0][2][tx] * wy * wz
                    ds[1] += fv * wx * taps[1][2][ty] * wz
                    ds[2] += fv * wx * wy * taps[2][2][tz]
    gx = np.zeros((n, 4))
    gx[:, :3] = (ds / d[:, None]).T
    dw = dfa * v ** fa + (dfb * v ** fb if fb >= 0 else 0.0) + np.where(main_ok, dalpha / ncell, 0.0)
    dv = np.zeros(n)
    if fa > 0:
        dv += dfa * w * fa * v ** (fa - 1)
    if fb > 0:
        dv += dfb * w * fb * v ** (fb - 1)
    gx[:, 3] = dv
    return gx, dw
