// binning.hip -- fused deconvolution, interlace combine, power and Legendre
// binning (K4) for gfx950.
//
// Replaces compute_periodic_power (reference src/main.c:302-468).  One pass
// over the complex half-grid [NX][NY][NZ/2+1] (rows padded to rowc complex)
// of every deposited grid; per mode:
//   f = per-axis frequency with the reference's float casts (tables built on
//       the host exactly as src/main.c:329-352 computes them),
//   fktot = 2 pi sqrt(fx^2+fy^2+fz^2), mu, kbin = (int)((float)(fktot-Mink)
//       /binwidth) -- all in double with no contraction, so bin indices and
//       mode counts are bit-identical to the reference's x86 build,
//   grid_cor = prod_axis (1/sinc)^(2 InterpOrder) (x 0.25 with interlacing),
//   delta += delta_int * e^{i(ax+ay+az)} (per-axis phase tables),
//   power of each requested spectrum, NM Hermitian weight, L_l(mu).
// Accumulation: every wavefront owns a private copy of the bins in LDS when
// they fit (adds in program order => bit-reproducible), else the workgroup
// shares one copy; workgroup partials are then summed in a fixed order.
#include "common.h"
#include "kernels.h"

namespace capow {

template <typename T>
struct CT;
template <>
struct CT<float> { typedef float2 t; };
template <>
struct CT<double> { typedef double2 t; };

__device__ __forceinline__ double dmul(double a, double b) { return __dmul_rn(a, b); }
__device__ __forceinline__ double dadd(double a, double b) { return __dadd_rn(a, b); }
__device__ __forceinline__ double ddiv(double a, double b) { return __ddiv_rn(a, b); }

// Bin index of one mode, bit-identical to (int)((float)(fktot - Mink)/binwidth)
// (src/main.c:368-373): multiply by the reciprocal and fall back to the exact
// division only when the quotient lies within 1e-9 of an integer, where the
// two could truncate differently.
__device__ __forceinline__ int kbin_of(double fdk, const BinArgsDev& a) {
  const double q = fdk * a.inv_binwidth;
  const double tq = trunc(q);
  const double fr = q - tq;
  if (fabs(fr) < 1e-9 || fabs(fr) > 1.0 - 1e-9) return (int)ddiv(fdk, a.binwidth);
  return (int)tq;
}

// Per-mode binning: frequencies, |k|, mu, bin, NM, deconvolution, interlace
// phase, power of each spectrum, Legendre weights, accumulation.
template <typename T>
__device__ __forceinline__ void bin_one(int i, uint32_t r, typename CT<T>::t va,
                                        typename CT<T>::t vai, typename CT<T>::t vb,
                                        typename CT<T>::t vbi, const BinArgsDev& a,
                                        double* accS, double* acc2, double* cnt, double* cnt2,
                                        int nl, int n2d) {
  const int NK = a.nk;
  const double twopi = 2.0 * 3.14159265358979323846;  // the reference's 2.0*M_PI
  const uint32_t j = a.fd_rowc.div(r);
  const int k = (int)(r - j * (uint32_t)a.rowc);
  if (k >= a.nzh) return;   // row padding
  const double fx = a.fx[i], fy = a.fy[j], fz = a.fz[k];
  const double fktot = dmul(twopi, sqrt(dadd(dadd(dmul(fx, fx), dmul(fy, fy)), dmul(fz, fz))));
  int kbin;
  if (a.logbins) {
    if (!(fktot > 0)) return;
    kbin = (int)ddiv((double)__double2float_rn(log10(fktot) - a.mink), a.binwidth);
  } else {
    kbin = kbin_of((double)__double2float_rn(fktot - a.mink), a);
  }
  if (!((kbin >= 0) && (kbin < NK) && (fktot <= a.min_nyq))) return;
  double mu = 0.0;
  if (fktot > 0) {
    const double fl = a.los == 1 ? fx : (a.los == 2 ? fy : fz);
    if (a.out2d) {
      mu = ddiv(dmul(twopi, fl), fktot);   // exact: feeds the mu-bin index
    } else {
      double rk = __builtin_amdgcn_rcp(fktot);
      rk = rk * (2.0 - fktot * rk);
      rk = rk * (2.0 - fktot * rk);
      mu = (twopi * fl) * rk;
    }
  }
  // Hermitian weight (src/main.c:422-429)
  double NM = 2.0;
  if (k == 0) {
    if (j == 0) {
      if ((i != 0) && (i != a.nx / 2)) NM = 1.0;
    } else {
      if (i != a.nx / 2) NM = 1.0;
    }
  }
  const double cor = a.corx[i] * a.cory[j] * a.corz[k];  // x0.25 folded in corz
  double ar = (double)va.x, ai = (double)va.y, br = 0, bi = 0;
  double cr = 1, ci = 0;
  if (a.interlace) {
    const double2 px = a.phx[i], py = a.phy[j], pz = a.phz[k];
    const double qr = px.x * py.x - px.y * py.y, qi = px.x * py.y + px.y * py.x;
    cr = qr * pz.x - qi * pz.y;
    ci = qr * pz.y + qi * pz.x;
    ar += (double)vai.x * cr - (double)vai.y * ci;
    ai += (double)vai.x * ci + (double)vai.y * cr;
  }
  if (a.hasB) {
    br = (double)vb.x;
    bi = (double)vb.y;
    if (a.interlace) {
      br += (double)vbi.x * cr - (double)vbi.y * ci;
      bi += (double)vbi.x * ci + (double)vbi.y * cr;
    }
  }
  const double mu2 = mu * mu;
  const double L2 = 1.5 * mu2 - 0.5;
  const double L4 = 4.375 * mu2 * mu2 - 3.75 * mu2 + 0.375;
  const double L3 = 2.5 * mu2 * mu - 1.5 * mu;
  int b2 = 0;
  bool in2d = false;
  if (a.out2d) {
    int mubin = (int)(((mu + 1.0) / 2.0) * (double)(float)a.nmu);
    if (mubin >= a.nmu) mubin = a.nmu - 1;
    b2 = kbin * a.nmu + mubin;
    in2d = (b2 >= 0) && (b2 < NK * a.nmu);
  }
  const double wc = NM * cor;
#pragma unroll 3
  for (int s = 0; s < a.nspec; s++) {
    double p;
    switch (a.kind[s]) {
      case SPEC_AUTO_A: p = ar * ar + ai * ai; break;
      case SPEC_AUTO_B: p = br * br + bi * bi; break;
      case SPEC_CROSS_IM: p = ar * bi - ai * br; break;
      default: p = ar * br + ai * bi; break;
    }
    const double np = wc * p;
    double* b = accS + s * nl * NK + kbin;
    atomicAdd(&b[0], np);
    atomicAdd(&b[NK], L2 * np);
    atomicAdd(&b[2 * NK], L4 * np);
    if (a.odd) {
      atomicAdd(&b[3 * NK], mu * np);
      atomicAdd(&b[4 * NK], L3 * np);
    }
    if (in2d) atomicAdd(&acc2[s * n2d + b2], np);
  }
  atomicAdd(&cnt[kbin], NM);
  if (in2d) atomicAdd(&cnt2[b2], NM);
}

#define BIN_U 4   // consecutive modes per lane per iteration (loads issued together)

template <typename T>
__global__ __launch_bounds__(256) void k_bin(const T* __restrict__ gA, const T* __restrict__ gAi,
                                             const T* __restrict__ gB, const T* __restrict__ gBi,
                                             BinArgsDev a, double* __restrict__ partials) {
  typedef typename CT<T>::t C;
  extern __shared__ double sm[];
  const int NK = a.nk;
  const int nl = a.odd ? 5 : 3;
  const int F = a.nspec * nl;                       // sums per bin
  const int n2d = a.out2d ? a.nk * a.nmu : 0;
  const int per_copy = F * NK + a.nspec * n2d + NK + n2d;  // doubles (counts as double)
  const int wave = threadIdx.x >> 6;
  double* my = sm + (wave % a.copies) * per_copy;
  for (int q = threadIdx.x; q < per_copy * a.copies; q += blockDim.x) sm[q] = 0.0;
  __syncthreads();
  double* accS = my;                       // [s][l'][NK]
  double* acc2 = my + F * NK;              // [s][NK*NMU]
  double* cnt = acc2 + a.nspec * n2d;      // [NK]
  double* cnt2 = cnt + NK;                 // [NK*NMU]

  const uint32_t plane = (uint32_t)a.ny * (uint32_t)a.rowc;   // padded rows
  const int64_t total = (int64_t)a.nx * plane;
  const C* A = reinterpret_cast<const C*>(gA);
  const C* Ai = reinterpret_cast<const C*>(gAi);
  const C* B = reinterpret_cast<const C*>(gB);
  const C* Bi = reinterpret_cast<const C*>(gBi);
  const C zero = {T(0), T(0)};
  // lane handles BIN_U consecutive modes; groups are strided over the grid
  const int64_t gstride = (int64_t)gridDim.x * blockDim.x * BIN_U;
  for (int64_t m0 = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) * BIN_U; m0 < total;
       m0 += gstride) {
    C va[BIN_U], vai[BIN_U], vb[BIN_U], vbi[BIN_U];
#pragma unroll
    for (int u = 0; u < BIN_U; u++) {
      const int64_t m = m0 + u;
      const bool ok = m < total;
      va[u] = ok ? A[m] : zero;
      vai[u] = (ok && a.interlace) ? Ai[m] : zero;
      vb[u] = (ok && a.hasB) ? B[m] : zero;
      vbi[u] = (ok && a.hasB && a.interlace) ? Bi[m] : zero;
    }
    int i = (int)(m0 / plane);
    uint32_t r = (uint32_t)(m0 - (int64_t)i * plane);
#pragma unroll
    for (int u = 0; u < BIN_U; u++) {
      if (m0 + u < total) bin_one<T>(i, r, va[u], vai[u], vb[u], vbi[u], a, accS, acc2, cnt, cnt2, nl, n2d);
      if (++r == plane) { r = 0; i++; }
    }
  }
  __syncthreads();
  // sum the copies in a fixed order and write this workgroup's partial
  double* out = partials + (int64_t)blockIdx.x * per_copy;
  for (int q = threadIdx.x; q < per_copy; q += blockDim.x) {
    double t = 0.0;
    for (int c = 0; c < a.copies; c++) t += sm[c * per_copy + q];
    out[q] = t;
  }
}

// Fixed-order sum over workgroup partials.
__global__ void k_bin_final(const double* __restrict__ partials, int nwg, int per_copy,
                            double* __restrict__ out) {
  int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= per_copy) return;
  double t = 0.0;
  for (int w = 0; w < nwg; w++) t += partials[(int64_t)w * per_copy + q];
  out[q] = t;
}

template <typename T>
static int run_bin_t(BinLaunch& L) {
  const int per_copy = L.per_copy;
  const size_t lds = (size_t)per_copy * L.dev.copies * sizeof(double);
  if (lds > 64 * 1024)
    hipFuncSetAttribute((const void*)k_bin<T>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  TIMED(L.timer, "bin",
        (k_bin<T><<<L.nwg, 256, lds, L.stream>>>((const T*)L.grid[0], (const T*)L.grid[1],
                                                 (const T*)L.grid[2], (const T*)L.grid[3], L.dev,
                                                 L.partials)));
  TIMED(L.timer, "bin_final",
        (k_bin_final<<<(per_copy + 255) / 256, 256, 0, L.stream>>>(L.partials, L.nwg, per_copy,
                                                                    L.out)));
  return hipGetLastError() == hipSuccess ? CAPOW_OK : CAPOW_ERR_HIP;
}

int run_bin(BinLaunch& L, bool fp64) { return fp64 ? run_bin_t<double>(L) : run_bin_t<float>(L); }

}  // namespace capow
