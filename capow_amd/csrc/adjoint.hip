// adjoint.hip -- backward of the periodic P_l(k) path for gfx950: raw bins
// -> k-space grid gradients -> real-space grid gradients -> particles.
//
// No counterpart in the reference (it has no derivatives).  The forward
// chain it inverts is the deposit of add_to_grid (src/read_data.c:514-745),
// the mean subtraction (src/main.c:211-225), the r2c FFT and the binning of
// compute_periodic_power (src/main.c:302-468); the backward of the host-side
// normalisation (src/main.c:1014-1056) is left to the caller.
#include "binning_math.h"

namespace capow {

// Gradient of the raw bins with respect to the stored half-spectrum modes
// of the (up to) four k-space grids, written over them in place.  Per mode:
//   h_s = NM * cor * sum_l dL/draw_{s,l}(kbin) L_l(mu),
//   delta = A + phi Ai, beta = B + phi Bi (phi: interlace phase),
//   |delta|^2 -> 2 h delta,  |beta|^2 -> 2 h beta,
//   Im(conj(delta) beta) -> -i h beta, i h delta,
//   Re(conj(delta) beta) -> h beta, h delta,
//   dA = d delta, dAi = conj(phi) d delta (the same for B).
// Gradients are dL/dRe + i dL/dIm.  Modes outside the bins (and the row
// padding) get zero.  The k = 0 gradients go to g0 (Re, Im of A, Ai, B, Bi)
// for the mean-subtraction term.
template <typename T>
__global__ __launch_bounds__(256) void k_bin_backward(T* __restrict__ gA, T* __restrict__ gAi,
                                                      T* __restrict__ gB, T* __restrict__ gBi,
                                                      const double* __restrict__ tabs,
                                                      BinArgsDev a,
                                                      const double* __restrict__ G,
                                                      double* __restrict__ g0) {
  typedef typename CT<T>::t C;
  const uint32_t plane = (uint32_t)a.ny * (uint32_t)a.rowc;
  const int64_t total = (int64_t)a.nx * plane;
  C* A = reinterpret_cast<C*>(gA);
  C* Ai = reinterpret_cast<C*>(gAi);
  C* B = reinterpret_cast<C*>(gB);
  C* Bi = reinterpret_cast<C*>(gBi);
  const bool il = a.interlace != 0, hb = a.hasB != 0;
  const int NK = a.nk;
  // (x plane, offset in the plane) of the thread's mode, advanced by the
  // grid stride with a carry: one 64-bit division per thread, not per mode
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t step_i = stride / plane;
  const uint32_t step_r = (uint32_t)(stride - step_i * plane);
  int64_t i64 = m / plane;
  uint32_t r = (uint32_t)(m - i64 * plane);
  for (; m < total; m += stride) {
    const int i = (int)i64;
    i64 += step_i;
    const uint32_t rr = r;
    r += step_r;
    if (r >= plane) {
      r -= plane;
      i64++;
    }
    const uint32_t j = a.fd_rowc.div(rr);
    const int k = (int)(rr - j * (uint32_t)a.rowc);
    double dr = 0, di = 0, er = 0, ei = 0;   // d delta, d beta
    double cr = 1, ci = 0;
    if (k < a.nzh) {
      const RowVals R = row_vals(i, (int)j, tabs, a);
      const ModeGeom M = mode_geom(R, k, tabs[a.ofz + k], tabs[a.ocz + k], tabs[a.opz + 2 * k],
                                   tabs[a.opz + 2 * k + 1], a);
      cr = M.cr;
      ci = M.ci;
      if (M.ok) {
        const C va = A[m];
        double ar = (double)va.x, ai = (double)va.y, br = 0, bi = 0;
        if (il) {
          const C v = Ai[m];
          ar += (double)v.x * cr - (double)v.y * ci;
          ai += (double)v.x * ci + (double)v.y * cr;
        }
        if (hb) {
          const C v = B[m];
          br = (double)v.x;
          bi = (double)v.y;
          if (il) {
            const C w = Bi[m];
            br += (double)w.x * cr - (double)w.y * ci;
            bi += (double)w.x * ci + (double)w.y * cr;
          }
        }
        const double mu = M.mu, mu2 = mu * mu;
        const double L2 = 1.5 * mu2 - 0.5;
        const double L4 = 4.375 * mu2 * mu2 - 3.75 * mu2 + 0.375;
        const double L3 = 2.5 * mu2 * mu - 1.5 * mu;
        const double wc = M.NM * M.cor;
        for (int s = 0; s < a.nspec; s++) {
          const double* g = G + (int64_t)s * 5 * NK + M.kbin;   // [l][NK], l = 0..4
          double hs = g[0] + L2 * g[2 * NK] + L4 * g[4 * NK];
          if (a.odd) hs += mu * g[NK] + L3 * g[3 * NK];
          const double h = wc * hs;
          switch (a.kind[s]) {
            case SPEC_AUTO_A: dr += 2.0 * h * ar; di += 2.0 * h * ai; break;
            case SPEC_AUTO_B: er += 2.0 * h * br; ei += 2.0 * h * bi; break;
            case SPEC_CROSS_IM:
              dr += h * bi; di -= h * br;
              er -= h * ai; ei += h * ar;
              break;
            default:
              dr += h * br; di += h * bi;
              er += h * ar; ei += h * ai;
              break;
          }
        }
      }
    }
    // conj(phi) * gradient for the interlaced grids
    const double dir_ = cr * dr + ci * di, dii = cr * di - ci * dr;
    const double eir = cr * er + ci * ei, eii = cr * ei - ci * er;
    A[m] = C{(T)dr, (T)di};
    if (il) Ai[m] = C{(T)dir_, (T)dii};
    if (hb) {
      B[m] = C{(T)er, (T)ei};
      if (il) Bi[m] = C{(T)eir, (T)eii};
    }
    if (m == 0) {
      g0[0] = dr; g0[1] = di; g0[2] = dir_; g0[3] = dii;
      g0[4] = er; g0[5] = ei; g0[6] = eir; g0[7] = eii;
    }
  }
}

template <typename T>
__device__ __forceinline__ T vpow_b(T v, int k) {
  switch (k) {
    case 0: return T(1);
    case 1: return v;
    case 2: return v * v;
    case 3: return v * v * v;
    default: return v * v * v * v;
  }
}

// Transposed deposit: for each particle (input order), the taps of the main
// and interlaced deposits read the real-space gradient grids.  With the
// fields A = w v^a, B = w v^b deposited with weights W = wx wy wz:
//   dL/dA_p = sum_taps gA W,  dL/dx = sum_taps (A_p gA + B_p gB) dW/dx,
// then dL/dw, dL/dv through the fields, dL/dv through the RSD shift of the
// line-of-sight coordinate, and dL/dw through alpha = N / Ncells for
// mean-subtracted density fields (dL/dalpha = -Ncells Re(gA(0) + gAi(0))).
// Outputs (x, y, z, v_los) gradients and the weight gradient.
template <typename T, int ORDER>
__global__ __launch_bounds__(256) void k_gather(const T* __restrict__ parts,
                                                const T* __restrict__ wts, int64_t n, Geom g,
                                                const T* __restrict__ gA, const T* __restrict__ gAi,
                                                const T* __restrict__ gB, const T* __restrict__ gBi,
                                                const double* __restrict__ g0,
                                                T* __restrict__ gout, T* __restrict__ gw) {
  const bool hasB = g.fb >= 0;
  const double dalpha_w = g.mean_sub_a ? -(g0[0] + (g.interlace ? g0[2] : 0.0)) : 0.0;
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n;
       p += (int64_t)gridDim.x * blockDim.x) {
    const T x0 = parts[4 * p], y0 = parts[4 * p + 1], z0 = parts[4 * p + 2], v = parts[4 * p + 3];
    const T w = g.has_w ? wts[p] : T(1);
    T pos[3] = {x0, y0, z0};
    if (g.rsd_axis >= 0) {
      const T sh = (v * (T)g.rsd_f1) / (T)g.rsd_den;   // as in the forward
      pos[g.rsd_axis] += sh;
    }
    const T fa = w * vpow_b<T>(v, g.fa);
    const T fb = hasB ? w * vpow_b<T>(v, g.fb) : T(0);
    double dfa = 0, dfb = 0, ds[3] = {0, 0, 0};
    bool main_ok = false;
    for (int ph = 0; ph < (g.interlace ? 2 : 1); ph++) {
      const T* ga = ph ? gAi : gA;
      const T* gb = ph ? gBi : gB;
      const AxisOut<T> ax = axis_taps<T, ORDER>(pos[0], ph ? g.d[0] / 2.0 : 0.0, g, 0, g.los, g.edge);
      const AxisOut<T> ay = axis_taps<T, ORDER>(pos[1], ph ? g.d[1] / 2.0 : 0.0, g, 1, g.los, g.edge);
      const AxisOut<T> az = axis_taps<T, ORDER>(pos[2], ph ? g.d[2] / 2.0 : 0.0, g, 2, g.los, g.edge);
      if (!(ax.ok && ay.ok && az.ok)) continue;
      if (ph == 0) main_ok = true;
      for (int tx = 0; tx < ORDER; tx++) {
        int gx = ax.first + tx;
        if (g.periodic) gx += gx < 0 ? g.n[0] : (gx >= g.n[0] ? -g.n[0] : 0);
        else if (gx < 0 || gx >= g.n[0]) continue;
        for (int ty = 0; ty < ORDER; ty++) {
          int gy = ay.first + ty;
          if (g.periodic) gy += gy < 0 ? g.n[1] : (gy >= g.n[1] ? -g.n[1] : 0);
          else if (gy < 0 || gy >= g.n[1]) continue;
          const int64_t rowb = ((int64_t)gx * g.n[1] + gy) * g.rowlen;
          for (int tz = 0; tz < ORDER; tz++) {
            int gz = az.first + tz;
            if (g.periodic) gz += gz < 0 ? g.n[2] : (gz >= g.n[2] ? -g.n[2] : 0);
            else if (gz < 0 || gz >= g.n[2]) continue;
            const double a = (double)ga[rowb + gz];
            const double b = hasB ? (double)gb[rowb + gz] : 0.0;
            const double W = (double)ax.w[tx] * (double)ay.w[ty] * (double)az.w[tz];
            dfa += a * W;
            dfb += b * W;
            const double fv = (double)fa * a + (double)fb * b;
            ds[0] += fv * (double)ax.dw[tx] * (double)ay.w[ty] * (double)az.w[tz];
            ds[1] += fv * (double)ax.w[tx] * (double)ay.dw[ty] * (double)az.w[tz];
            ds[2] += fv * (double)ax.w[tx] * (double)ay.w[ty] * (double)az.dw[tz];
          }
        }
      }
    }
    // scaled coordinate s = (u - MIN)/d: d/du = 1/d
    double gx = ds[0] / g.d[0], gy = ds[1] / g.d[1], gz = ds[2] / g.d[2];
    // fields: A = w v^a, B = w v^b
    const int ea = g.fa, eb = g.fb;
    double dw = dfa * (double)vpow_b<T>(v, ea) + (hasB ? dfb * (double)vpow_b<T>(v, eb) : 0.0);
    double dv = 0.0;
    if (ea > 0) dv += dfa * (double)w * ea * (double)vpow_b<T>(v, ea - 1);
    if (hasB && eb > 0) dv += dfb * (double)w * eb * (double)vpow_b<T>(v, eb - 1);
    if (g.rsd_axis >= 0) {
      const double gl = g.rsd_axis == 0 ? gx : (g.rsd_axis == 1 ? gy : gz);
      dv += gl * g.rsd_f1 / g.rsd_den;
    }
    if (main_ok) dw += dalpha_w;
    gout[4 * p] = (T)gx;
    gout[4 * p + 1] = (T)gy;
    gout[4 * p + 2] = (T)gz;
    gout[4 * p + 3] = (T)dv;
    if (gw) gw[p] = (T)dw;
  }
}

// Tile gather: the same transposed deposit, one workgroup per tile of the
// forward's counting sort.  The sort leaves a tile's own particles (those
// whose main cell lies in the tile: one row per particle) contiguous, with
// every tap of either deposit within one cell below and two above the tile
// per axis (the sort flags any other).  So per phase the workgroup stages
// the tile's gradient region -- (T + 3) cells per axis, the A and B values
// of a cell side by side -- in LDS with coalesced loads, and its particles
// read their taps there: the grids are read once per tile (plus the halo)
// instead of once per tap.  A tap row's three z values are summed with the
// weights and their derivatives first (separable weights), so per row
//   dL/dA += wx wy S_a,  dL/dx += dwx wy F,  dL/dy += wx dwy F,
//   dL/dz += wx wy F'    (S_a = sum_z wz g_A, F = fa S_a + fb S_b, F' with dwz)
// all in FP64 (a derivative is a difference of neighbouring cells' values).  Particles the
// sort dropped (outside every deposit) keep the zero gradient written
// before.  Rows are written back at their input index.
#define GT_T 1024  // threads per tile workgroup (16 waves: at most 128 VGPRs)
#define GT_K (sizeof(T) == 8 ? 1 : 2)   // own particles per thread per pass over the region
#define GT_U 4     // region cells per thread whose loads are in flight together
template <typename T, int ORDER, bool HASB>
__global__ __launch_bounds__(GT_T) void k_gather_tiles(
    const T* __restrict__ sorted, const T* __restrict__ sorted_w,
    const uint32_t* __restrict__ sidx, const uint32_t* __restrict__ off, Geom g,
    const T* __restrict__ gA, const T* __restrict__ gAi, const T* __restrict__ gB,
    const T* __restrict__ gBi, const double* __restrict__ g0, T* __restrict__ gout,
    T* __restrict__ gw, unsigned int* __restrict__ errs, FastDiv divz, FastDiv divy) {
  typedef typename CT<T>::t C2;
  extern __shared__ __align__(16) unsigned char smem[];
  C2* regAB = reinterpret_cast<C2*>(smem);   // HASB: (A, B) per cell
  T* regA = reinterpret_cast<T*>(smem);      // otherwise A only
  const int TX = g.tdim[0], TY = g.tdim[1], TZ = g.tdim[2];
  const int RX = TX + 3, RY = TY + 3, RZ = TZ + 3;
  const int ncell = RX * RY * RZ;
  const int64_t tile = blockIdx.x;
  const int ttx = (int)(tile % g.ntile[0]);
  const int tty = (int)((tile / g.ntile[0]) % g.ntile[1]);
  const int ttz = (int)(tile / ((int64_t)g.ntile[0] * g.ntile[1]));
  const int t0x = ttx * TX, t0y = tty * TY, t0z = ttz * TZ;
  const uint32_t s0 = off[tile * 27], own = off[tile * 27 + 27] - s0;
  if (own == 0) return;   // the whole workgroup: no barrier reached yet
  const bool per = g.periodic != 0;
  const double dalpha_w = g.mean_sub_a ? -(g0[0] + (g.interlace ? g0[2] : 0.0)) : 0.0;
  // region index of a tile-relative first tap (the nearest periodic image)
  auto loc = [&](int first, int t0, int T_, int n) {
    int o = first - t0;
    if (per) {
      if (o > T_ + 2) o -= n;
      else if (o < -1) o += n;
    }
    return o + 1;
  };
  // grid index of region cell l along an axis (-1: outside a non-periodic grid)
  auto gidx = [&](int l, int t0, int n) {
    int q = t0 - 1 + l;
    if (q < 0 || q >= n) {
      if (!per) return -1;
      q += q < 0 ? n : -n;
    }
    return q;
  };
  for (uint32_t base = 0; base < own; base += GT_T * GT_K) {
    double dfa[GT_K], dfb[GT_K], ds0[GT_K], ds1[GT_K], ds2[GT_K];
    bool mok[GT_K];
    // the chunk's particles, loaded once for both phases (every slot's load
    // in flight before the first use)
    T px[GT_K], py[GT_K], pz[GT_K], pv[GT_K], pw[GT_K];
#pragma unroll
    for (int k = 0; k < GT_K; k++) {
      dfa[k] = dfb[k] = ds0[k] = ds1[k] = ds2[k] = 0.0;
      mok[k] = false;
      const uint32_t f = base + (uint32_t)(k * GT_T) + threadIdx.x;
      T x = T(0), y = T(0), z = T(0), v = T(0), w = T(1);
      if (f < own) {
        load4<T>(sorted, s0 + f, x, y, z, v);
        if (g.has_w) w = sorted_w[s0 + f];
      }
      px[k] = x; py[k] = y; pz[k] = z; pv[k] = v; pw[k] = w;
    }
    for (int ph = 0; ph < (g.interlace ? 2 : 1); ph++) {
      const T* ga = ph ? gAi : gA;
      const T* gb = ph ? gBi : gB;
      __syncthreads();   // the region's previous readers are done
      // GT_U cells per thread per round, their loads issued together
      for (int c0 = threadIdx.x; c0 < ncell; c0 += GT_T * GT_U) {
        T av[GT_U], bv[GT_U];
#pragma unroll
        for (int u = 0; u < GT_U; u++) {
          const int c = c0 + u * GT_T;
          const uint32_t r = divz.div((uint32_t)c);
          const int lz = c - (int)r * RZ;
          const int lx = (int)divy.div(r), ly = (int)r - lx * RY;
          const int qx = gidx(lx, t0x, g.n[0]), qy = gidx(ly, t0y, g.n[1]), qz = gidx(lz, t0z, g.n[2]);
          av[u] = bv[u] = T(0);
          if (c < ncell && qx >= 0 && qy >= 0 && qz >= 0) {
            const int64_t e = ((int64_t)qx * g.n[1] + qy) * g.rowlen + qz;
            av[u] = ga[e];
            if (HASB) bv[u] = gb[e];
          }
        }
#pragma unroll
        for (int u = 0; u < GT_U; u++) {
          const int c = c0 + u * GT_T;
          if (c < ncell) {
            if (HASB) regAB[c] = C2{av[u], bv[u]};
            else regA[c] = av[u];
          }
        }
      }
      __syncthreads();
#pragma unroll
      for (int k = 0; k < GT_K; k++) {
        const uint32_t f = base + (uint32_t)(k * GT_T) + threadIdx.x;
        if (f >= own) continue;
        const T v = pv[k], w = pw[k];
        const AxisOut<T> ax = axis_taps<T, ORDER>(px[k], ph ? g.d[0] / 2.0 : 0.0, g, 0, g.los, g.edge);
        const AxisOut<T> ay = axis_taps<T, ORDER>(py[k], ph ? g.d[1] / 2.0 : 0.0, g, 1, g.los, g.edge);
        const AxisOut<T> az = axis_taps<T, ORDER>(pz[k], ph ? g.d[2] / 2.0 : 0.0, g, 2, g.los, g.edge);
        if (!(ax.ok && ay.ok && az.ok)) continue;
        if (ph == 0) mok[k] = true;
        const int lx0 = loc(ax.first, t0x, TX, g.n[0]);
        const int ly0 = loc(ay.first, t0y, TY, g.n[1]);
        const int lz0 = loc(az.first, t0z, TZ, g.n[2]);
        // Taps outside the region can only be taps the edge rules drop
        // (zero weight; e.g. TSC x in [MAX, MAX + dx) with a line of sight:
        // its first tap is the last grid slice): they read zero.  A live
        // tap outside means the sort's reach check failed.
        bool vx[ORDER], vy[ORDER], vz[ORDER];
        bool bad = false;
#pragma unroll
        for (int t = 0; t < ORDER; t++) {
          vx[t] = (unsigned)(lx0 + t) < (unsigned)RX;
          vy[t] = (unsigned)(ly0 + t) < (unsigned)RY;
          vz[t] = (unsigned)(lz0 + t) < (unsigned)RZ;
          bad = bad || (!vx[t] && (ax.w[t] != T(0) || ax.dw[t] != T(0))) ||
                (!vy[t] && (ay.w[t] != T(0) || ay.dw[t] != T(0))) ||
                (!vz[t] && (az.w[t] != T(0) || az.dw[t] != T(0)));
        }
        if (bad) {
          atomicOr(&errs[0], 1u);
          continue;
        }
        const double fa = (double)(w * vpow_b<T>(v, g.fa));
        const double fb = HASB ? (double)(w * vpow_b<T>(v, g.fb)) : 0.0;
        // weights and the sums in FP64: a derivative is a difference of
        // neighbouring cells' values, which cancels on fine grids
        double wzd[ORDER], dwzd[ORDER];
#pragma unroll
        for (int t = 0; t < ORDER; t++) {
          wzd[t] = (double)az.w[t];
          dwzd[t] = (double)az.dw[t];
        }
#pragma unroll
        for (int tx = 0; tx < ORDER; tx++) {
#pragma unroll
          for (int ty = 0; ty < ORDER; ty++) {
            if (!(vx[tx] && vy[ty])) continue;
            const int rb = ((lx0 + tx) * RY + (ly0 + ty)) * RZ + lz0;
            double sa = 0.0, sb = 0.0, da = 0.0, db = 0.0;
#pragma unroll
            for (int tz = 0; tz < ORDER; tz++) {
              T a = T(0), b = T(0);
              if (vz[tz]) {
                if (HASB) {
                  const C2 ab = regAB[rb + tz];
                  a = ab.x;
                  b = ab.y;
                } else {
                  a = regA[rb + tz];
                }
              }
              sa += wzd[tz] * (double)a;
              da += dwzd[tz] * (double)a;
              if (HASB) {
                sb += wzd[tz] * (double)b;
                db += dwzd[tz] * (double)b;
              }
            }
            const double wx = (double)ax.w[tx], wy = (double)ay.w[ty];
            const double wxy = wx * wy;
            const double fw = fa * sa + fb * sb, fd = fa * da + fb * db;
            dfa[k] += wxy * sa;
            if (HASB) dfb[k] += wxy * sb;
            ds0[k] += (double)ax.dw[tx] * wy * fw;
            ds1[k] += wx * (double)ay.dw[ty] * fw;
            ds2[k] += wxy * fd;
          }
        }
      }
    }
    // the chain through the fields, the RSD shift and alpha, as k_gather
#pragma unroll
    for (int k = 0; k < GT_K; k++) {
      const uint32_t f = base + (uint32_t)(k * GT_T) + threadIdx.x;
      if (f >= own) continue;
      const T v = pv[k], w = pw[k];
      const double gx = ds0[k] / g.d[0], gy = ds1[k] / g.d[1], gz = ds2[k] / g.d[2];
      const int ea = g.fa, eb = g.fb;
      double dw = dfa[k] * (double)vpow_b<T>(v, ea) + (HASB ? dfb[k] * (double)vpow_b<T>(v, eb) : 0.0);
      double dv = 0.0;
      if (ea > 0) dv += dfa[k] * (double)w * ea * (double)vpow_b<T>(v, ea - 1);
      if (HASB && eb > 0) dv += dfb[k] * (double)w * eb * (double)vpow_b<T>(v, eb - 1);
      if (g.rsd_axis >= 0) {
        const double gl = g.rsd_axis == 0 ? gx : (g.rsd_axis == 1 ? gy : gz);
        dv += gl * g.rsd_f1 / g.rsd_den;
      }
      if (mok[k]) dw += dalpha_w;
      const int64_t p = (int64_t)sidx[s0 + f];
      store4<T>(gout, p, (T)gx, (T)gy, (T)gz, (T)dv);
      if (gw) gw[p] = (T)dw;
    }
  }
}

// LDS bytes of the tile gather's region for the context's tiles
size_t gather_tiles_lds(const Geom& g, bool fp64) {
  const size_t cells = (size_t)(g.tdim[0] + 3) * (g.tdim[1] + 3) * (g.tdim[2] + 3);
  return cells * (fp64 ? 8 : 4) * (g.fb >= 0 ? 2 : 1);
}

int run_bin_backward(const BinLaunch& L, const double* G, double* g0, bool fp64) {
  const int blocks = 2048;
  if (fp64)
    TIMED(L.timer, "bin_backward",
          (k_bin_backward<double><<<blocks, 256, 0, L.stream>>>(
              (double*)L.grid[0], (double*)L.grid[1], (double*)L.grid[2], (double*)L.grid[3],
              L.tabs, L.dev, G, g0)));
  else
    TIMED(L.timer, "bin_backward",
          (k_bin_backward<float><<<blocks, 256, 0, L.stream>>>(
              (float*)L.grid[0], (float*)L.grid[1], (float*)L.grid[2], (float*)L.grid[3],
              L.tabs, L.dev, G, g0)));
  return hipGetLastError() == hipSuccess ? CAPOW_OK : CAPOW_ERR_HIP;
}

template <typename T, int ORDER>
static void launch_gather(const GatherArgs& a) {
  const int64_t want = (a.n + 255) / 256;
  const int blocks = (int)(want < 4096 ? (want < 1 ? 1 : want) : 4096);
  TIMED(a.timer, "gather",
        (k_gather<T, ORDER><<<blocks, 256, 0, a.stream>>>(
            (const T*)a.parts, (const T*)a.wts, a.n, *a.g, (const T*)a.grid[0], (const T*)a.grid[1],
            (const T*)a.grid[2], (const T*)a.grid[3], a.g0, (T*)a.gout, (T*)a.gw)));
}

template <typename T, int ORDER>
static int launch_gather_tiles(const GatherArgs& a) {
  const Geom& g = *a.g;
  const size_t lds = gather_tiles_lds(g, sizeof(T) == 8);
  const int64_t ntiles = (int64_t)g.ntile[0] * g.ntile[1] * g.ntile[2];
  // particles outside every deposit are not in the sort: zero gradients
  const int64_t words = a.n * (int64_t)(sizeof(T) / 4);
  if (zero_u32((uint32_t*)a.gout, 4 * words, a.stream) != CAPOW_OK) return CAPOW_ERR_HIP;
  if (a.gw && zero_u32((uint32_t*)a.gw, words, a.stream) != CAPOW_OK) return CAPOW_ERR_HIP;
  FastDiv divz, divy;   // region cell -> (row, z), row -> (x, y)
  divz.init((uint32_t)(g.tdim[2] + 3));
  divy.init((uint32_t)(g.tdim[1] + 3));
  const bool hb = g.fb >= 0;
  const void* fn = hb ? (const void*)k_gather_tiles<T, ORDER, true> : (const void*)k_gather_tiles<T, ORDER, false>;
  if (lds > 64 * 1024 &&
      hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) != hipSuccess)
    return CAPOW_ERR_HIP;
#define GT_LAUNCH(HB)                                                                           \
  TIMED(a.timer, "gather",                                                                      \
        (k_gather_tiles<T, ORDER, HB><<<(unsigned)ntiles, GT_T, lds, a.stream>>>(                \
            (const T*)a.sorted, (const T*)a.sorted_w, a.sorted_idx, a.offsets, g,                \
            (const T*)a.grid[0], (const T*)a.grid[1], (const T*)a.grid[2], (const T*)a.grid[3],  \
            a.g0, (T*)a.gout, (T*)a.gw, a.errs, divz, divy)))
  if (hb) GT_LAUNCH(true);
  else GT_LAUNCH(false);
#undef GT_LAUNCH
  return hipGetLastError() == hipSuccess ? CAPOW_OK : CAPOW_ERR_HIP;
}

int run_gather(const GatherArgs& a, bool fp64) {
  if (a.sorted && a.sorted_idx && a.offsets && a.errs &&
      gather_tiles_lds(*a.g, fp64) <= 160 * 1024) {
    switch (a.g->order * 2 + (fp64 ? 1 : 0)) {
      case 2: return launch_gather_tiles<float, 1>(a);
      case 3: return launch_gather_tiles<double, 1>(a);
      case 4: return launch_gather_tiles<float, 2>(a);
      case 5: return launch_gather_tiles<double, 2>(a);
      case 6: return launch_gather_tiles<float, 3>(a);
      case 7: return launch_gather_tiles<double, 3>(a);
      default: return CAPOW_ERR_ARG;
    }
  }
  switch (a.g->order * 2 + (fp64 ? 1 : 0)) {
    case 2: launch_gather<float, 1>(a); break;
    case 3: launch_gather<double, 1>(a); break;
    case 4: launch_gather<float, 2>(a); break;
    case 5: launch_gather<double, 2>(a); break;
    case 6: launch_gather<float, 3>(a); break;
    case 7: launch_gather<double, 3>(a); break;
    default: return CAPOW_ERR_ARG;
  }
  return hipGetLastError() == hipSuccess ? CAPOW_OK : CAPOW_ERR_HIP;
}

}  // namespace capow
