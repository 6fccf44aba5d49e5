// survey.hip -- the survey (Yamamoto) multipole estimator for gfx950.
//
// Reference: compute_survey_power (src/main.c:472-776) and
// assign_survey_power (src/main.c:781-967).  For each multipole l the field
// is weighted in real space by the products of the unit vector
// r^ = (cell - origin)/|cell - origin| that span degree l, transformed, and
// each transform is paired with the monopole's k-space field times the
// matching k^ product:
//   l = 0: F0 = FFT(field); P0 += |F0|^2, P2 -= 1/2 of it, P4 += 3/8 of it
//   l = 1: 3 fields r_i;           P1 += k_i/k Im-part, P3 -= 3/2 of it
//   l = 2: 6 fields r_i r_j;       P2 += 3/2 m k_i k_j/k^2 Re-part, P4 -= 15/4
//   l = 3: 10 fields r_i r_j r_k;  P3 += 5/2 m k_i k_j k_k/k^3 Im-part
//   l = 4: 15 fields r_i^2 r_j r_k; P4 += 35/8 m k_i^2 k_j k_k/k^4 Re-part
// (m: the multiplicity of the index combination).  With a B grid
// (momentum cross spectra) the products are the symmetrised cross terms of
// src/main.c:884-898.  The transforms are the periodic path's FFT; the
// geometry (frequencies with the reference's float casts, k bins, window
// deconvolution, interlace phase, Hermitian weights) is the periodic
// binning's, so the mode counts agree with it bit for bit.
//
// Layout: every pass reads the real-space sources and writes the weighted
// copies once (HBM-bound, 2 G N_g s bytes), the FFT runs its three passes,
// and the binning reads the transformed copies plus F0 (and, for l = 0,
// writes F0).  Per-wave FP64 bin copies in LDS, reduced in a fixed order,
// keep the sums reproducible.
#include "binning_math.h"

namespace capow {

// Weight of one cell: vec[i0]...vec[i_{n-1}] / r^n (idx[0] twice for
// n = 4), the reference's r2 == 0 guard included.
__device__ __forceinline__ double survey_w(double vx, double vy, double vz, int n, int i0, int i1,
                                           int i2) {
  const double vec[3] = {vx, vy, vz};
  double r2 = vx * vx + vy * vy + vz * vz;
  if (r2 == 0.0) r2 = 1.0;   // src/main.c:557
  if (n == 1) return vec[i0] / sqrt(r2);
  if (n == 2) return vec[i0] * vec[i1] / r2;
  if (n == 3) return vec[i0] * vec[i1] * vec[i2] / (r2 * sqrt(r2));
  return vec[i0] * vec[i0] * vec[i1] * vec[i2] / (r2 * r2);
}

// One workgroup per (ix, iy) row; a thread weights V consecutive cells of
// every field with one 16-byte load and store each (rows start on 128-byte
// boundaries).  n = 0 copies.
template <typename T>
__global__ __launch_bounds__(256) void k_survey_weight(const T* __restrict__ s0,
                                                       const T* __restrict__ s1,
                                                       const T* __restrict__ s2,
                                                       const T* __restrict__ s3,
                                                       T* __restrict__ w0, T* __restrict__ w1,
                                                       T* __restrict__ w2, T* __restrict__ w3,
                                                       Geom g, double ox, double oy, double oz,
                                                       int n, int i0, int i1, int i2) {
  constexpr int V = 16 / sizeof(T);
  typedef T vec_t __attribute__((ext_vector_type(V)));
  const int64_t row = blockIdx.x;                  // ix * NY + iy
  const int ix = (int)(row / g.n[1]);
  const int iy = (int)(row - (int64_t)ix * g.n[1]);
  const int64_t base = row * g.rowlen;
  const double vx = (double)ix * g.d[0] + g.lo[0] - ox;
  const double vy = (double)iy * g.d[1] + g.lo[1] - oy;
  const int nz = g.n[2];
  const T* src[4] = {s0, s1, s2, s3};
  T* dst[4] = {w0, w1, w2, w3};
  for (int z0 = threadIdx.x * V; z0 < nz; z0 += blockDim.x * V) {
    double w[V];
#pragma unroll
    for (int u = 0; u < V; u++)
      w[u] = n == 0 ? 1.0 : survey_w(vx, vy, (double)(z0 + u) * g.d[2] + g.lo[2] - oz, n, i0, i1, i2);
    const int64_t q = base + z0;
    if (z0 + V <= nz) {
#pragma unroll
      for (int f = 0; f < 4; f++) {
        if (!src[f]) continue;
        vec_t v = *reinterpret_cast<const vec_t*>(src[f] + q);
        if (n != 0) {
#pragma unroll
          for (int u = 0; u < V; u++) v[u] = (T)((double)v[u] * w[u]);
        }
        *reinterpret_cast<vec_t*>(dst[f] + q) = v;
      }
    } else {
      for (int u = 0; z0 + u < nz; u++)
        for (int f = 0; f < 4; f++)
          if (src[f]) dst[f][q + u] = n == 0 ? src[f][q + u] : (T)((double)src[f][q + u] * w[u]);
    }
  }
}

// Sums one multipole pass touches, in the order of its LDS slots (indices
// into the accumulated [P0, P1, P2, P3, P4, counts][NK]).
__host__ __device__ __forceinline__ int survey_slots(int multipole) {
  return multipole == 0 ? 4 : (multipole <= 2 ? 2 : 1);
}
__host__ __device__ __forceinline__ int survey_slot_target(int multipole, int slot) {
  // l = 0: P0 P2 P4 counts; l = 1: P1 P3; l = 2: P2 P4; l = 3: P3; l = 4: P4
  return multipole == 0 ? (slot == 3 ? 5 : 2 * slot)
                        : (multipole == 1 ? 1 + 2 * slot : (multipole == 2 ? 2 + 2 * slot : multipole));
}

// One multipole pass.  A wave owns one LDS copy of the pass's sums
// (survey_slots(l) x NK doubles) and walks whole (i, j) rows.
template <typename T>
__global__ __launch_bounds__(256) void k_survey_bin(const T* __restrict__ gA,
                                                    const T* __restrict__ gAi,
                                                    const T* __restrict__ gB,
                                                    const T* __restrict__ gBi,
                                                    T* __restrict__ F0, T* __restrict__ F0m,
                                                    const double* __restrict__ tabs,
                                                    BinArgsDev a, int mom, int multipole, int ii,
                                                    int jj, int kk, double* __restrict__ partials) {
  typedef typename CT<T>::t C;
  extern __shared__ double sm[];
  const int NK = a.nk;
  const int per_copy = survey_slots(multipole) * NK;
  const int nwave = blockDim.x >> 6;
  for (int q = threadIdx.x; q < per_copy * nwave; q += blockDim.x) sm[q] = 0.0;
  __syncthreads();
  double* acc = sm + (threadIdx.x >> 6) * per_copy;
  const C* A = reinterpret_cast<const C*>(gA);
  const C* Ai = reinterpret_cast<const C*>(gAi);
  const C* B = reinterpret_cast<const C*>(gB);
  const C* Bi = reinterpret_cast<const C*>(gBi);
  C* f0 = reinterpret_cast<C*>(F0);
  C* f0m = reinterpret_cast<C*>(F0m);
  const double twopi = 2.0 * 3.14159265358979323846;
  // one wave per (i, j) row: the row's x and y quantities are wave-uniform,
  // lanes take kz = lane, lane + 64, ...
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int lane = threadIdx.x & 63;
  const int64_t nrows = (int64_t)a.nx * a.ny;
  for (int64_t row = (int64_t)blockIdx.x * nwave + wave; row < nrows;
       row += (int64_t)gridDim.x * nwave) {
    const int i = (int)(row / a.ny);
    const int j = (int)(row - (int64_t)i * a.ny);
    const RowVals R = row_vals(i, j, tabs, a);
    for (int k = lane; k < a.nzh; k += 64) {
      const int64_t m = row * a.rowc + k;
      const double fz = tabs[a.ofz + k];
      const ModeGeom G = mode_geom(R, k, fz, tabs[a.ocz + k], tabs[a.opz + 2 * k],
                                   tabs[a.opz + 2 * k + 1], a);
      if (!G.ok) continue;
      double dkr = (double)A[m].x, dki = (double)A[m].y;
      if (a.interlace) {
        const C v = Ai[m];
        dkr += (double)v.x * G.cr - (double)v.y * G.ci;
        dki += (double)v.x * G.ci + (double)v.y * G.cr;
      }
      double dkr_m = 0.0, dki_m = 0.0;
      if (mom) {
        dkr_m = (double)B[m].x;
        dki_m = (double)B[m].y;
        if (a.interlace) {
          const C v = Bi[m];
          dkr_m += (double)v.x * G.cr - (double)v.y * G.ci;
          dki_m += (double)v.x * G.ci + (double)v.y * G.cr;
        }
      }
      const double cor = G.cor;
      double power;
      if (multipole == 0) {
        C o;
        o.x = (T)dkr;
        o.y = (T)dki;
        f0[m] = o;
        if (mom) {
          o.x = (T)dkr_m;
          o.y = (T)dki_m;
          f0m[m] = o;
          power = (dkr * dki_m - dki * dkr_m) * cor;
        } else {
          power = (dkr * dkr + dki * dki) * cor;
        }
      } else {
        const C z = f0[m];
        const double zr = (double)z.x, zi = (double)z.y;
        if (!mom) {
          power = (multipole & 1) ? (dkr * zi - dki * zr) * cor : (dkr * zr + dki * zi) * cor;
        } else {
          const C zm = f0m[m];
          const double mr = (double)zm.x, mi = (double)zm.y;
          power = (multipole & 1)
                      ? 0.5 * (dkr * mr + dki * mi + dkr_m * zr + dki_m * zi) * cor
                      : 0.5 * (dkr * mi - dki * mr + dkr_m * zi - dki_m * zr) * cor;
        }
      }
      power *= G.NM;
      const int kb = G.kbin;
      if (multipole == 0) {   // slots P0 P2 P4 counts
        atomicAdd(&acc[kb], power);
        atomicAdd(&acc[NK + kb], -0.5 * power);
        atomicAdd(&acc[2 * NK + kb], 0.375 * power);
        atomicAdd(&acc[3 * NK + kb], G.NM);
        continue;
      }
      const double fk = G.fktot;
      if (!(fk > 0.0)) continue;
      const double kv[3] = {twopi * R.fx, twopi * R.fy, twopi * fz};
      if (multipole == 1) {   // slots P1 P3
        power *= kv[ii] / fk;
        atomicAdd(&acc[kb], power);
        atomicAdd(&acc[NK + kb], -1.5 * power);
      } else if (multipole == 2) {   // slots P2 P4
        const double pre = ii != jj ? 2.0 : 1.0;
        power *= pre * kv[ii] * kv[jj] / (fk * fk);
        atomicAdd(&acc[kb], 1.5 * power);
        atomicAdd(&acc[NK + kb], -3.75 * power);
      } else if (multipole == 3) {   // slot P3
        const double pre = ii != kk ? (ii != jj ? 6.0 : 3.0) : 1.0;
        atomicAdd(&acc[kb], 2.5 * pre * kv[ii] * kv[jj] * kv[kk] / (fk * fk * fk) * power);
      } else {   // slot P4
        const double pre = ii == jj ? (ii != kk ? 4.0 : 1.0) : (jj == kk ? 6.0 : 12.0);
        atomicAdd(&acc[kb],
                  4.375 * pre * kv[ii] * kv[ii] * kv[jj] * kv[kk] / (fk * fk * fk * fk) * power);
      }
    }
  }
  __syncthreads();
  double* out = partials + (int64_t)blockIdx.x * per_copy;
  for (int q = threadIdx.x; q < per_copy; q += blockDim.x) {
    double t = 0.0;
    for (int c = 0; c < nwave; c++) t += sm[c * per_copy + q];
    out[q] = t;
  }
}

// The same pass over mirror groups (symmetric x and y tables): the modes
// (i, j, k), (NX - i, j, k), (i, NY - j, k), (NX - i, NY - j, k) of a
// canonical (i, j) (0 <= i <= NX/2, 0 <= j <= NY/2) share |k|, the bin, NM
// and the window deconvolution, and their k^ components differ only in the
// signs of x and y.  One wave per canonical (i, j), lanes over k: the
// geometry and the pass's k^ factor F are computed once per group, each
// live mirror m contributes s_m NM cor (its power) with its own grid values
// and interlace phase (s_m: the sign the flipped axes give the k^
// product), and the group's sum times F goes to the wave's LDS copy in one
// set of adds.
template <typename T>
__global__ __launch_bounds__(256) void k_survey_bin_mirror(const T* __restrict__ gA,
                                                           const T* __restrict__ gAi,
                                                           const T* __restrict__ gB,
                                                           const T* __restrict__ gBi,
                                                           T* __restrict__ F0, T* __restrict__ F0m,
                                                           const double* __restrict__ tabs,
                                                           BinArgsDev a, int mom, int multipole,
                                                           int ii, int jj, int kk,
                                                           double* __restrict__ partials) {
  typedef typename CT<T>::t C;
  extern __shared__ double sm[];
  const int NK = a.nk;
  const int per_copy = survey_slots(multipole) * NK;
  const int nwave = blockDim.x >> 6;
  for (int q = threadIdx.x; q < per_copy * nwave; q += blockDim.x) sm[q] = 0.0;
  __syncthreads();
  double* acc = sm + (threadIdx.x >> 6) * per_copy;
  const C* A = reinterpret_cast<const C*>(gA);
  const C* Ai = reinterpret_cast<const C*>(gAi);
  const C* B = reinterpret_cast<const C*>(gB);
  const C* Bi = reinterpret_cast<const C*>(gBi);
  C* f0 = reinterpret_cast<C*>(F0);
  C* f0m = reinterpret_cast<C*>(F0m);
  const double twopi = 2.0 * 3.14159265358979323846;
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int lane = threadIdx.x & 63;
  const int hy = a.ny / 2 + 1;
  const int64_t ngroups = (int64_t)(a.nx / 2 + 1) * hy;
  // the sign a flipped axis gives the pass's k^ product, per mirror
  // (bit 0: x flipped, bit 1: y flipped): ax = axes of the product, x or y
  // counted once per odd power
  int odd_x = 0, odd_y = 0;
  {
    int e[3] = {0, 0, 0};
    if (multipole >= 1) e[ii]++;
    if (multipole >= 2) e[jj]++;
    if (multipole == 3) e[kk]++;
    if (multipole == 4) { e[ii]++; e[kk]++; }
    odd_x = e[0] & 1;
    odd_y = e[1] & 1;
  }
  for (int64_t grp = (int64_t)blockIdx.x * nwave + wave; grp < ngroups;
       grp += (int64_t)gridDim.x * nwave) {
    const int i = (int)(grp / hy);
    const int j = (int)(grp - (int64_t)i * hy);
    const bool mx = i > 0 && 2 * i < a.nx, my = j > 0 && 2 * j < a.ny;
    const int i2 = mx ? a.nx - i : i, j2 = my ? a.ny - j : j;
    const RowVals R = row_vals(i, j, tabs, a);
    // each mirror's row and interlace phase (x and y part)
    const int mi[4] = {i, i2, i, i2}, mj[4] = {j, j, j2, j2};
    const bool live[4] = {true, mx, my, mx && my};
    int64_t rowb[4];
    double qr[4], qi[4], sgn[4];
#pragma unroll
    for (int m = 0; m < 4; m++) {
      rowb[m] = ((int64_t)mi[m] * a.ny + mj[m]) * a.rowc;
      qr[m] = 1.0;
      qi[m] = 0.0;
      if (a.interlace) {
        const double pxr = tabs[a.opx + 2 * mi[m]], pxi = tabs[a.opx + 2 * mi[m] + 1];
        const double pyr = tabs[a.opy + 2 * mj[m]], pyi = tabs[a.opy + 2 * mj[m] + 1];
        qr[m] = pxr * pyr - pxi * pyi;
        qi[m] = pxr * pyi + pxi * pyr;
      }
      sgn[m] = ((odd_x && (m & 1)) != (odd_y && (m & 2))) ? -1.0 : 1.0;
    }
    for (int k = lane; k < a.nzh; k += 64) {
      const double fz = tabs[a.ofz + k];
      const double pzr = tabs[a.opz + 2 * k], pzi = tabs[a.opz + 2 * k + 1];
      const ModeGeom G = mode_geom(R, k, fz, tabs[a.ocz + k], pzr, pzi, a);
      if (!G.ok) continue;
      const double cor = G.cor;
      double sum = 0.0, nmodes = 0.0;
#pragma unroll
      for (int m = 0; m < 4; m++) {
        if (!live[m]) continue;
        const double cr = qr[m] * pzr - qi[m] * pzi, ci = qr[m] * pzi + qi[m] * pzr;
        const int64_t q = rowb[m] + k;
        double dkr = (double)A[q].x, dki = (double)A[q].y;
        if (a.interlace) {
          const C w = Ai[q];
          dkr += (double)w.x * cr - (double)w.y * ci;
          dki += (double)w.x * ci + (double)w.y * cr;
        }
        double dkr_m = 0.0, dki_m = 0.0;
        if (mom) {
          dkr_m = (double)B[q].x;
          dki_m = (double)B[q].y;
          if (a.interlace) {
            const C w = Bi[q];
            dkr_m += (double)w.x * cr - (double)w.y * ci;
            dki_m += (double)w.x * ci + (double)w.y * cr;
          }
        }
        double power;
        if (multipole == 0) {
          C o;
          o.x = (T)dkr;
          o.y = (T)dki;
          f0[q] = o;
          if (mom) {
            o.x = (T)dkr_m;
            o.y = (T)dki_m;
            f0m[q] = o;
            power = (dkr * dki_m - dki * dkr_m) * cor;
          } else {
            power = (dkr * dkr + dki * dki) * cor;
          }
        } else {
          const C z = f0[q];
          const double zr = (double)z.x, zi = (double)z.y;
          if (!mom) {
            power = (multipole & 1) ? (dkr * zi - dki * zr) * cor : (dkr * zr + dki * zi) * cor;
          } else {
            const C zm = f0m[q];
            const double mr = (double)zm.x, mi2 = (double)zm.y;
            power = (multipole & 1)
                        ? 0.5 * (dkr * mr + dki * mi2 + dkr_m * zr + dki_m * zi) * cor
                        : 0.5 * (dkr * mi2 - dki * mr + dkr_m * zi - dki_m * zr) * cor;
          }
        }
        sum += sgn[m] * (power * G.NM);
        nmodes += G.NM;
      }
      const int kb = G.kbin;
      if (multipole == 0) {   // slots P0 P2 P4 counts
        atomicAdd(&acc[kb], sum);
        atomicAdd(&acc[NK + kb], -0.5 * sum);
        atomicAdd(&acc[2 * NK + kb], 0.375 * sum);
        atomicAdd(&acc[3 * NK + kb], nmodes);
        continue;
      }
      const double fk = G.fktot;
      if (!(fk > 0.0)) continue;
      // the canonical mode's k^ factor (the mirrors' signs are in sum)
      const double kv[3] = {twopi * R.fx, twopi * R.fy, twopi * fz};
      if (multipole == 1) {   // slots P1 P3
        const double p1 = sum * (kv[ii] / fk);
        atomicAdd(&acc[kb], p1);
        atomicAdd(&acc[NK + kb], -1.5 * p1);
      } else if (multipole == 2) {   // slots P2 P4
        const double pre = ii != jj ? 2.0 : 1.0;
        const double p2 = sum * (pre * kv[ii] * kv[jj] / (fk * fk));
        atomicAdd(&acc[kb], 1.5 * p2);
        atomicAdd(&acc[NK + kb], -3.75 * p2);
      } else if (multipole == 3) {   // slot P3
        const double pre = ii != kk ? (ii != jj ? 6.0 : 3.0) : 1.0;
        atomicAdd(&acc[kb], 2.5 * pre * kv[ii] * kv[jj] * kv[kk] / (fk * fk * fk) * sum);
      } else {   // slot P4
        const double pre = ii == jj ? (ii != kk ? 4.0 : 1.0) : (jj == kk ? 6.0 : 12.0);
        atomicAdd(&acc[kb],
                  4.375 * pre * kv[ii] * kv[ii] * kv[jj] * kv[kk] / (fk * fk * fk * fk) * sum);
      }
    }
  }
  __syncthreads();
  double* out = partials + (int64_t)blockIdx.x * per_copy;
  for (int q = threadIdx.x; q < per_copy; q += blockDim.x) {
    double t = 0.0;
    for (int c = 0; c < nwave; c++) t += sm[c * per_copy + q];
    out[q] = t;
  }
}

// acc[target] += sum over the workgroup partials of one pass: one block per
// sum, each thread a fixed stride of workgroups, then a fixed-shape tree, so
// the order of the additions never changes.
__global__ __launch_bounds__(256) void k_survey_final(const double* __restrict__ partials, int nwg,
                                                      int len, int nk, int multipole,
                                                      double* __restrict__ acc) {
  __shared__ double red[256];
  const int q = blockIdx.x;   // < len
  double t = 0.0;
  for (int w = threadIdx.x; w < nwg; w += 256) t += partials[(int64_t)w * len + q];
  red[threadIdx.x] = t;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const int slot = q / nk, k = q - slot * nk;
    acc[survey_slot_target(multipole, slot) * nk + k] += red[0];
  }
}

template <typename T>
static int weight_t(SurveyLaunch& L, int n, const int idx[3]) {
  const Geom& g = *L.g;
  const int64_t rows = (int64_t)g.n[0] * g.n[1];
  TIMED(L.timer, "survey_weight",
        (k_survey_weight<T><<<(unsigned)rows, 256, 0, L.stream>>>(
            (const T*)L.src[0], (const T*)L.src[1], (const T*)L.src[2], (const T*)L.src[3],
            (T*)L.work[0], (T*)L.work[1], (T*)L.work[2], (T*)L.work[3], g, L.origin[0],
            L.origin[1], L.origin[2], n, idx[0], idx[1], idx[2])));
  return hipGetLastError() == hipSuccess ? CAPOW_OK : CAPOW_ERR_HIP;
}

int run_survey_weight(SurveyLaunch& L, bool fp64, int n, const int idx[3]) {
  return fp64 ? weight_t<double>(L, n, idx) : weight_t<float>(L, n, idx);
}

template <typename T>
static int bin_t(SurveyLaunch& L, int multipole, const int idx[3]) {
  const int NK = L.dev.nk;
  const int per_copy = survey_slots(multipole) * NK;
  const size_t lds = (size_t)4 * per_copy * sizeof(double);
  const void* fn = L.mirror ? (const void*)k_survey_bin_mirror<T> : (const void*)k_survey_bin<T>;
  if (lds > 64 * 1024) hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  const T* w1 = (const T*)(L.work[1] ? L.work[1] : L.work[0]);
  const T* w2 = (const T*)(L.work[2] ? L.work[2] : L.work[0]);
  const T* w3 = (const T*)(L.work[3] ? L.work[3] : L.work[0]);
  if (L.mirror) {
    TIMED(L.timer, "survey_bin",
          (k_survey_bin_mirror<T><<<L.nwg, 256, lds, L.stream>>>(
              (const T*)L.work[0], w1, w2, w3, (T*)L.F0, (T*)(L.F0m ? L.F0m : L.F0), L.tabs, L.dev,
              L.mom, multipole, idx[0], idx[1], idx[2], L.partials)));
  } else {
    TIMED(L.timer, "survey_bin",
          (k_survey_bin<T><<<L.nwg, 256, lds, L.stream>>>(
              (const T*)L.work[0], w1, w2, w3, (T*)L.F0, (T*)(L.F0m ? L.F0m : L.F0), L.tabs, L.dev,
              L.mom, multipole, idx[0], idx[1], idx[2], L.partials)));
  }
  TIMED(L.timer, "survey_final",
        (k_survey_final<<<per_copy, 256, 0, L.stream>>>(L.partials, L.nwg, per_copy, NK, multipole,
                                                       L.acc)));
  return hipGetLastError() == hipSuccess ? CAPOW_OK : CAPOW_ERR_HIP;
}

int run_survey_bin(SurveyLaunch& L, bool fp64, int multipole, const int idx[3]) {
  return fp64 ? bin_t<double>(L, multipole, idx) : bin_t<float>(L, multipole, idx);
}

}  // namespace capow
