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
  for (int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; m < total;
       m += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(m / plane);
    const uint32_t r = (uint32_t)(m - (int64_t)i * plane);
    const uint32_t j = a.fd_rowc.div(r);
    const int k = (int)(r - j * (uint32_t)a.rowc);
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
  const T d0 = (T)g.d[0], d1 = (T)g.d[1], d2 = (T)g.d[2];
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
      const T x = pos[0] + (ph ? d0 / T(2) : T(0));
      const T y = pos[1] + (ph ? d1 / T(2) : T(0));
      const T z = pos[2] + (ph ? d2 / T(2) : T(0));
      const AxisOut<T> ax = axis_taps<T, ORDER>(x, (T)g.lo[0], (T)g.hi[0], d0, (T)g.L[0], g.n[0], true, g.los, g.edge, (T)g.invd[0]);
      const AxisOut<T> ay = axis_taps<T, ORDER>(y, (T)g.lo[1], (T)g.hi[1], d1, (T)g.L[1], g.n[1], false, g.los, g.edge, (T)g.invd[1]);
      const AxisOut<T> az = axis_taps<T, ORDER>(z, (T)g.lo[2], (T)g.hi[2], d2, (T)g.L[2], g.n[2], false, g.los, g.edge, (T)g.invd[2]);
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

int run_gather(const GatherArgs& a, bool fp64) {
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
