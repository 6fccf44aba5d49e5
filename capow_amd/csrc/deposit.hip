// deposit.hip -- mass assignment (K1) + mean subtraction (K2) for gfx950.
//
// Replaces the per-particle scatter of add_to_grid (reference
// src/read_data.c:514-745, called from src/read_data.c:91-130), the halo add
// (src/read_data.c:152-175) and the mean subtraction (src/main.c:211-225).
//
// Design (one GPU holds the whole grid):
//   1. k_bucket<COUNT>: per particle, apply the RSD shift, decide acceptance
//      of the main and interlaced deposits, compute the bucket
//      (deposit tile x 27 edge regions) and count it; block partials of the
//      global sums (N gridded, sum w v^n) and the max |field| per field.
//   2. k_stats_final: fixed-order reduction of the block partials, alpha,
//      per-field fixed-point scales.
//   3. exclusive scan of the bucket counts.
//   4. k_bucket<SCATTER>: counting-sort scatter of the shifted particles.
//   5. k_deposit: one workgroup per output tile gathers the particles of its
//      own tile and of the edge regions of its 26 neighbours, accumulates
//      every tap that lands in the tile into LDS, then writes the tile once
//      (value - alpha for density fields).  No grid memset, no global
//      atomics on the grid.
// Determinism: every tap contribution is rounded to an integer multiple of
// a per-field quantum 2^-e (e from the field's max |value|, so one tap is
// < 2^41 quanta) and accumulated in LDS as an integer-valued double.  Sums
// of such values below 2^53 quanta are exact, so the grid does not depend
// on the order in which particles or taps are added.
#include "common.h"
#include "kernels.h"

namespace capow {

template <typename T>
__device__ __forceinline__ T vpow(T v, int k) {
  // same multiplication order as the reference's tvr*tvr*tvr
  switch (k) {
    case 0: return T(1);
    case 1: return v;
    case 2: return v * v;
    case 3: return v * v * v;
    default: return v * v * v * v;
  }
}

template <typename T>
struct Vec4;
template <>
struct Vec4<float> { typedef float4 type; };
template <>
struct Vec4<double> { typedef double4 type; };

template <typename T>
__device__ __forceinline__ void load4(const T* p, int64_t i, T& x, T& y, T& z, T& v) {
  typedef typename Vec4<T>::type V;
  V q = reinterpret_cast<const V*>(p)[i];
  x = q.x; y = q.y; z = q.z; v = q.w;
}

template <typename T>
__device__ __forceinline__ void store4(T* p, int64_t i, T x, T y, T z, T v) {
  typedef typename Vec4<T>::type V;
  V q;
  q.x = x; q.y = y; q.z = z; q.w = v;
  reinterpret_cast<V*>(p)[i] = q;
}

// Per-particle classification shared by the count and scatter passes.
template <typename T, int ORDER>
struct PClass {
  T x, y, z;       // position after the RSD shift
  bool acc_main, acc_int;
  int64_t bucket;  // -1: rejected by every deposit
  bool reach_ok;
};

template <typename T, int ORDER>
__device__ __forceinline__ PClass<T, ORDER> classify(T x, T y, T z, T v, const Geom& g) {
  PClass<T, ORDER> c;
  if (g.rsd_axis >= 0) {
    // tx += tvr*(1+z)/(100 sqrt(Om (1+z)^3 + 1 - Om))   src/read_data.c:67-76
    T sh = (v * (T)g.rsd_f1) / (T)g.rsd_den;
    if (g.rsd_axis == 0) x += sh;
    else if (g.rsd_axis == 1) y += sh;
    else z += sh;
  }
  c.x = x; c.y = y; c.z = z;
  const T lo0 = (T)g.lo[0], lo1 = (T)g.lo[1], lo2 = (T)g.lo[2];
  const T hi0 = (T)g.hi[0], hi1 = (T)g.hi[1], hi2 = (T)g.hi[2];
  const T d0 = (T)g.d[0], d1 = (T)g.d[1], d2 = (T)g.d[2];
  const T L0 = (T)g.L[0], L1 = (T)g.L[1], L2 = (T)g.L[2];
  AxisOut<T> mx = axis_taps<T, ORDER>(x, lo0, hi0, d0, L0, g.n[0], true, g.los, g.edge);
  AxisOut<T> my = axis_taps<T, ORDER>(y, lo1, hi1, d1, L1, g.n[1], false, g.los, g.edge);
  AxisOut<T> mz = axis_taps<T, ORDER>(z, lo2, hi2, d2, L2, g.n[2], false, g.los, g.edge);
  c.acc_main = mx.ok && my.ok && mz.ok;
  AxisOut<T> ix, iy, iz;
  c.acc_int = false;
  if (g.interlace) {
    ix = axis_taps<T, ORDER>(x + d0 / T(2), lo0, hi0, d0, L0, g.n[0], true, g.los, g.edge);
    iy = axis_taps<T, ORDER>(y + d1 / T(2), lo1, hi1, d1, L1, g.n[1], false, g.los, g.edge);
    iz = axis_taps<T, ORDER>(z + d2 / T(2), lo2, hi2, d2, L2, g.n[2], false, g.los, g.edge);
    c.acc_int = ix.ok && iy.ok && iz.ok;
  }
  c.reach_ok = true;
  if (!c.acc_main && !c.acc_int) {
    c.bucket = -1;
    return c;
  }
  int key[3];
  key[0] = axis_key<T, ORDER>(x, lo0, hi0, d0, L0, g.n[0], true, g.los, g.edge);
  key[1] = axis_key<T, ORDER>(y, lo1, hi1, d1, L1, g.n[1], false, g.los, g.edge);
  key[2] = axis_key<T, ORDER>(z, lo2, hi2, d2, L2, g.n[2], false, g.los, g.edge);
  // Self-check: every kept tap must lie within [key+reach_lo, key+reach_hi]
  // (mod N on periodic axes), otherwise a tile that owns it would not read
  // this particle.
  const AxisOut<T>* ax[2][3] = {{&mx, &my, &mz}, {&ix, &iy, &iz}};
  for (int gi = 0; gi < 2; gi++) {
    if (gi == 0 ? !c.acc_main : !c.acc_int) continue;
    for (int a = 0; a < 3; a++) {
      const AxisOut<T>* o = ax[gi][a];
      for (int t = 0; t < ORDER; t++) {
        if (o->w[t] == T(0)) continue;
        int idx = o->first + t;
        int off = idx - key[a];
        if (g.periodic) {
          int n = g.n[a];
          off %= n;
          if (off < 0) off += n;
          if (off > n / 2) off -= n;
        }
        if (off < g.reach_lo || off > g.reach_hi) c.reach_ok = false;
      }
    }
  }
  int tile[3], reg[3];
  const int lw = -g.reach_lo, hw = g.reach_hi;
  for (int a = 0; a < 3; a++) {
    tile[a] = key[a] / g.tdim[a];
    int local = key[a] - tile[a] * g.tdim[a];
    reg[a] = g.ntile[a] == 1 ? 1 : region_of(local, g.tdim[a], lw, hw);
  }
  int64_t t = ((int64_t)tile[2] * g.ntile[1] + tile[1]) * g.ntile[0] + tile[0];
  c.bucket = t * 27 + reg[0] + 3 * reg[1] + 9 * reg[2];
  return c;
}

// Block partial layout (doubles): see STAT_* in kernels.h.
template <int NT>
__device__ __forceinline__ double block_sum(double v, double* sh) {
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int s = NT / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  double r = sh[0];
  __syncthreads();
  return r;
}

template <typename T, int ORDER, bool SCATTER>
__global__ __launch_bounds__(256) void k_bucket(const T* __restrict__ parts,
                                                const T* __restrict__ wts, int64_t n,
                                                Geom g, uint32_t* __restrict__ counts,
                                                uint32_t* __restrict__ cursor,
                                                T* __restrict__ sorted,
                                                T* __restrict__ sorted_w,
                                                double* __restrict__ partials,
                                                unsigned long long* __restrict__ gmax,
                                                unsigned int* __restrict__ errs) {
  __shared__ double sh[256];
  double s_ngrid = 0, s_w = 0, s_v1 = 0, s_v2 = 0, s_v3 = 0, s_v4 = 0, s_out = 0;
  double m_a = 0, m_b = 0;
  unsigned int oob = 0, reach_bad = 0;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    T x, y, z, v;
    load4<T>(parts, i, x, y, z, v);
    T w = g.has_w ? wts[i] : T(1);
    PClass<T, ORDER> c = classify<T, ORDER>(x, y, z, v, g);
    if (SCATTER) {
      if (c.bucket >= 0) {
        uint32_t pos = atomicAdd(&cursor[c.bucket], 1u);
        if ((int64_t)pos < n) {
          store4<T>(sorted, pos, c.x, c.y, c.z, v);
          if (g.has_w) sorted_w[pos] = w;
        }
      }
    } else {
      if (c.bucket >= 0) atomicAdd(&counts[c.bucket], 1u);
      if (!c.reach_ok) reach_bad++;
      // src/read_data.c:81-88 (on the shifted position)
      if ((c.x < (T)g.lo[0]) || (c.x > (T)g.hi[0]) || (c.y < (T)g.lo[1]) ||
          (c.y > (T)g.hi[1]) || (c.z < (T)g.lo[2]) || (c.z > (T)g.hi[2])) {
        if (g.los == 0) oob++;
        else s_out += 1.0;
      }
      if (c.acc_main) s_ngrid += (double)w;
      // moments over every particle read (src/read_data.c:132-138)
      double wd = (double)w, vd = (double)v;
      s_w += wd;
      s_v1 += (double)(w * v);
      s_v2 += (double)(w * v * v);
      s_v3 += (double)(w * v * v * v);
      s_v4 += (double)(w * v * v * v * v);
      if (c.acc_main || c.acc_int) {
        double fa = fabs((double)(w * vpow<T>(v, g.fa)));
        m_a = fmax(m_a, fa);
        if (g.fb >= 0) m_b = fmax(m_b, fabs((double)(w * vpow<T>(v, g.fb))));
      }
      (void)vd;
    }
  }
  if (!SCATTER) {
    double r[7] = {s_ngrid, s_w, s_v1, s_v2, s_v3, s_v4, s_out};
    for (int q = 0; q < 7; q++) {
      double t = block_sum<256>(r[q], sh);
      if (threadIdx.x == 0) partials[(int64_t)blockIdx.x * STAT_N + q] = t;
    }
    // max |field| (non-negative doubles order as their bit patterns)
    atomicMax(&gmax[0], (unsigned long long)__double_as_longlong(m_a));
    atomicMax(&gmax[1], (unsigned long long)__double_as_longlong(m_b));
    if (oob) atomicAdd(&errs[0], oob);
    if (reach_bad) atomicAdd(&errs[1], reach_bad);
  }
}

// Fixed-order reduction of block partials -> stats[STAT_N], alpha, scales.
__global__ void k_stats_final(const double* __restrict__ partials, int nblocks,
                              const unsigned long long* __restrict__ gmax, double ncells,
                              double* __restrict__ stats) {
  __shared__ double sh[256];
  for (int q = 0; q < 7; q++) {
    double acc = 0.0;
    // each thread sums a strided subset in fixed order, then a fixed tree
    for (int b = threadIdx.x; b < nblocks; b += 256) acc += partials[(int64_t)b * STAT_N + q];
    double t = block_sum<256>(acc, sh);
    if (threadIdx.x == 0) stats[q] = t;
  }
  if (threadIdx.x == 0) {
    stats[STAT_ALPHA] = stats[STAT_NGRID] / ncells;  // src/main.c:117
    for (int f = 0; f < 2; f++) {
      double m = __longlong_as_double((long long)gmax[f]);
      int e = 0;
      if (m > 0.0) {
        int ex;
        frexp(m, &ex);           // m < 2^ex
        e = 40 - ex;             // |tap| <= m < 2^(ex) -> scaled < 2^40
      }
      stats[STAT_SCALE_A + f] = ldexp(1.0, e);
      stats[STAT_INVSCALE_A + f] = ldexp(1.0, -e);
    }
  }
}

// ---------------------------------------------------------------------------
// Exclusive scan of uint32 counts (3-phase, 4096 elements per block).
// ---------------------------------------------------------------------------
#define SCAN_T 256
#define SCAN_PER 16
#define SCAN_CHUNK (SCAN_T * SCAN_PER)

__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* sh, uint32_t* total) {
  // Hillis-Steele inclusive scan over SCAN_T threads
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int o = 1; o < SCAN_T; o <<= 1) {
    uint32_t t = (int)threadIdx.x >= o ? sh[threadIdx.x - o] : 0u;
    __syncthreads();
    sh[threadIdx.x] += t;
    __syncthreads();
  }
  uint32_t incl = sh[threadIdx.x];
  *total = sh[SCAN_T - 1];
  __syncthreads();
  return incl - v;
}

__global__ __launch_bounds__(SCAN_T) void k_scan_reduce(const uint32_t* __restrict__ in, int64_t n,
                                                        uint32_t* __restrict__ bsum) {
  __shared__ uint32_t sh[SCAN_T];
  int64_t base = (int64_t)blockIdx.x * SCAN_CHUNK;
  uint32_t s = 0;
  for (int q = 0; q < SCAN_PER; q++) {
    int64_t i = base + (int64_t)q * SCAN_T + threadIdx.x;
    if (i < n) s += in[i];
  }
  uint32_t tot;
  block_excl_scan(s, sh, &tot);
  if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

__global__ __launch_bounds__(SCAN_T) void k_scan_apply(const uint32_t* __restrict__ in, int64_t n,
                                                       const uint32_t* __restrict__ boff,
                                                       uint32_t* __restrict__ out,
                                                       uint32_t* __restrict__ out2) {
  __shared__ uint32_t sh[SCAN_T];
  int64_t base = (int64_t)blockIdx.x * SCAN_CHUNK + (int64_t)threadIdx.x * SCAN_PER;
  uint32_t v[SCAN_PER];
  uint32_t s = 0;
  for (int q = 0; q < SCAN_PER; q++) {
    int64_t i = base + q;
    v[q] = i < n ? in[i] : 0u;
    s += v[q];
  }
  uint32_t tot;
  uint32_t ex = block_excl_scan(s, sh, &tot) + (boff ? boff[blockIdx.x] : 0u);
  for (int q = 0; q < SCAN_PER; q++) {
    int64_t i = base + q;
    if (i < n) {
      out[i] = ex;
      if (out2) out2[i] = ex;
    }
    ex += v[q];
  }
  if (blockIdx.x == gridDim.x - 1 && threadIdx.x == SCAN_T - 1) {
    out[n] = ex;  // grand total at out[n]
  }
}

// ---------------------------------------------------------------------------
// Tile deposit.
// ---------------------------------------------------------------------------
#define DEP_T 256
#define MAX_BK 128

template <typename T, int ORDER>
__global__ __launch_bounds__(DEP_T) void k_deposit(const T* __restrict__ sorted,
                                                   const T* __restrict__ sorted_w,
                                                   const uint32_t* __restrict__ off,
                                                   Geom g, T* __restrict__ gA,
                                                   T* __restrict__ gAi, T* __restrict__ gB,
                                                   T* __restrict__ gBi,
                                                   const double* __restrict__ stats) {
  extern __shared__ double acc[];  // [ngr][ncell]
  __shared__ uint32_t s_bk_start[MAX_BK];
  __shared__ uint32_t s_pref[MAX_BK + 1];
  __shared__ uint32_t s_nbk;

  const int TX = g.tdim[0], TY = g.tdim[1], TZ = g.tdim[2];
  const int ncell = TX * TY * TZ;
  const bool hasB = g.fb >= 0;
  const int ngr = (hasB ? 2 : 1) * (g.interlace ? 2 : 1);
  // slot layout in LDS: 0 A, 1 A_int (if interlace), then B, B_int
  const int slotAi = 1, slotB = g.interlace ? 2 : 1, slotBi = 3;

  // tile of this workgroup
  int64_t tileid = blockIdx.x;
  const int ttx = (int)(tileid % g.ntile[0]);
  const int tty = (int)((tileid / g.ntile[0]) % g.ntile[1]);
  const int ttz = (int)(tileid / ((int64_t)g.ntile[0] * g.ntile[1]));
  const int t0[3] = {ttx * TX, tty * TY, ttz * TZ};

  for (int i = threadIdx.x; i < ngr * ncell; i += DEP_T) acc[i] = 0.0;
  if (threadIdx.x == 0) s_nbk = 0;
  __syncthreads();

  // Enumerate the (neighbour tile, region) buckets whose particles can touch
  // this tile: offset -1 needs the neighbour's high region, 0 all regions,
  // +1 the low region.
  const int tt[3] = {ttx, tty, ttz};
  for (int cand = threadIdx.x; cand < 729; cand += DEP_T) {
    int nc = cand / 27, r = cand % 27;
    int o[3] = {nc % 3 - 1, (nc / 3) % 3 - 1, nc / 9 - 1};
    int rr[3] = {r % 3, (r / 3) % 3, r / 9};
    bool ok = true;
    int nt[3];
    for (int a = 0; a < 3 && ok; a++) {
      if (g.ntile[a] == 1) {
        ok = (o[a] == 0) && (rr[a] == 1);
        nt[a] = 0;
        continue;
      }
      if (o[a] == -1 && rr[a] != 2) ok = false;
      if (o[a] == 1 && rr[a] != 0) ok = false;
      int q = tt[a] + o[a];
      if (q < 0 || q >= g.ntile[a]) {
        if (g.periodic) q = (q + g.ntile[a]) % g.ntile[a];
        else ok = false;
      }
      nt[a] = q;
    }
    if (!ok) continue;
    int64_t b = (((int64_t)nt[2] * g.ntile[1] + nt[1]) * g.ntile[0] + nt[0]) * 27 + r;
    uint32_t s0 = off[b], s1 = off[b + 1];
    if (s1 > s0) {
      uint32_t slot = atomicAdd(&s_nbk, 1u);
      if (slot < MAX_BK) {
        s_bk_start[slot] = s0;
        s_pref[slot + 1] = s1 - s0;
      }
    }
  }
  __syncthreads();
  const int nbk = (int)min(s_nbk, (uint32_t)MAX_BK);
  if (threadIdx.x == 0) {
    s_pref[0] = 0;
    for (int q = 0; q < nbk; q++) s_pref[q + 1] += s_pref[q];
  }
  __syncthreads();
  const uint32_t total = s_pref[nbk];

  const double scA = stats[STAT_SCALE_A], scB = stats[STAT_SCALE_B];
  const T d0 = (T)g.d[0], d1 = (T)g.d[1], d2 = (T)g.d[2];

  for (uint32_t i = threadIdx.x; i < total; i += DEP_T) {
    // locate the bucket holding flat index i
    int lo = 0, hi = nbk - 1;
    while (lo < hi) {
      int mid = (lo + hi + 1) >> 1;
      if (s_pref[mid] <= i) lo = mid;
      else hi = mid - 1;
    }
    int64_t src = (int64_t)s_bk_start[lo] + (i - s_pref[lo]);
    T x, y, z, v;
    load4<T>(sorted, src, x, y, z, v);
    T w = g.has_w ? sorted_w[src] : T(1);
    const double va = (double)(w * vpow<T>(v, g.fa)) * scA;
    const double vb = hasB ? (double)(w * vpow<T>(v, g.fb)) * scB : 0.0;
    for (int gi = 0; gi < (g.interlace ? 2 : 1); gi++) {
      T ux = gi ? x + d0 / T(2) : x;
      T uy = gi ? y + d1 / T(2) : y;
      T uz = gi ? z + d2 / T(2) : z;
      AxisOut<T> a[3];
      a[0] = axis_taps<T, ORDER>(ux, (T)g.lo[0], (T)g.hi[0], d0, (T)g.L[0], g.n[0], true, g.los, g.edge);
      a[1] = axis_taps<T, ORDER>(uy, (T)g.lo[1], (T)g.hi[1], d1, (T)g.L[1], g.n[1], false, g.los, g.edge);
      a[2] = axis_taps<T, ORDER>(uz, (T)g.lo[2], (T)g.hi[2], d2, (T)g.L[2], g.n[2], false, g.los, g.edge);
      if (!(a[0].ok && a[1].ok && a[2].ok)) continue;
      // local tile coordinates of each tap (or -1)
      int loc[3][3];
      for (int ax = 0; ax < 3; ax++) {
        const int n = g.n[ax], td = g.tdim[ax];
#pragma unroll
        for (int t = 0; t < ORDER; t++) {
          int gidx = a[ax].first + t;
          if (g.periodic) {
            if (gidx < 0) gidx += n;
            else if (gidx >= n) gidx -= n;
          }
          int l = gidx - t0[ax];
          loc[ax][t] = (a[ax].w[t] != T(0) && l >= 0 && l < td) ? l : -1;
        }
      }
      double* accA = acc + (gi ? slotAi : 0) * ncell;
      double* accB = acc + (gi ? slotBi : slotB) * ncell;
#pragma unroll
      for (int tx = 0; tx < ORDER; tx++) {
        if (loc[0][tx] < 0) continue;
#pragma unroll
        for (int ty = 0; ty < ORDER; ty++) {
          if (loc[1][ty] < 0) continue;
          const T wxy = a[0].w[tx] * a[1].w[ty];
          const int rowc = (loc[0][tx] * TY + loc[1][ty]) * TZ;
#pragma unroll
          for (int tz = 0; tz < ORDER; tz++) {
            if (loc[2][tz] < 0) continue;
            const double wt = (double)(wxy * a[2].w[tz]);
            const int cell = rowc + loc[2][tz];
            atomicAdd(&accA[cell], rint(va * wt));
            if (hasB) atomicAdd(&accB[cell], rint(vb * wt));
          }
        }
      }
    }
  }
  __syncthreads();

  // write the tile once: value/scale - alpha for density fields
  const double alpha = stats[STAT_ALPHA];
  const double iA = stats[STAT_INVSCALE_A], iB = stats[STAT_INVSCALE_B];
  T* outs[4] = {gA, g.interlace ? gAi : gB, gB, gBi};
  // map LDS slot -> (grid pointer, inv scale, mean)
  for (int s = 0; s < ngr; s++) {
    T* dst;
    double isc, mean;
    if (s == 0) { dst = gA; isc = iA; mean = g.mean_sub_a ? alpha : 0.0; }
    else if (g.interlace && s == slotAi) { dst = gAi; isc = iA; mean = g.mean_sub_a ? alpha : 0.0; }
    else if (s == slotB) { dst = gB; isc = iB; mean = 0.0; }
    else { dst = gBi; isc = iB; mean = 0.0; }
    const double* src = acc + s * ncell;
    for (int c = threadIdx.x; c < ncell; c += DEP_T) {
      int lz = c % TZ, ly = (c / TZ) % TY, lx = c / (TZ * TY);
      int64_t gi = ((int64_t)(t0[0] + lx) * g.n[1] + (t0[1] + ly)) * g.rowlen + (t0[2] + lz);
      dst[gi] = (T)(src[c] * isc - mean);
    }
  }
  (void)outs;
}

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------
template <typename T, int ORDER>
static int run_deposit_t(DepositArgs& a) {
  const Geom& g = *a.g;
  const int nb = a.count_blocks;
  hipStream_t st = a.stream;
  TIMED(a.timer, "bucket_count",
        (k_bucket<T, ORDER, false><<<nb, 256, 0, st>>>((const T*)a.parts, (const T*)a.wts, a.n, g,
                                                        a.counts, nullptr, nullptr, nullptr,
                                                        a.partials, a.gmax, a.errs)));
  TIMED(a.timer, "stats_final",
        (k_stats_final<<<1, 256, 0, st>>>(a.partials, nb, a.gmax,
                                         (double)g.n[0] * (double)g.n[1] * (double)g.n[2],
                                         a.stats)));
  // scan counts -> offsets (and cursor copy)
  int64_t nbk = g.nbuckets;
  int64_t nblk = (nbk + SCAN_CHUNK - 1) / SCAN_CHUNK;
  if (nblk > SCAN_CHUNK) return CAPOW_ERR_UNSUPPORTED;
  TIMED(a.timer, "scan", (k_scan_reduce<<<(int)nblk, SCAN_T, 0, st>>>(a.counts, nbk, a.bsum)));
  TIMED(a.timer, "scan", (k_scan_apply<<<1, SCAN_T, 0, st>>>(a.bsum, nblk, nullptr, a.boff, nullptr)));
  TIMED(a.timer, "scan", (k_scan_apply<<<(int)nblk, SCAN_T, 0, st>>>(a.counts, nbk, a.boff, a.offsets,
                                                                 a.cursor)));
  TIMED(a.timer, "bucket_scatter",
        (k_bucket<T, ORDER, true><<<nb, 256, 0, st>>>((const T*)a.parts, (const T*)a.wts, a.n, g,
                                                       nullptr, a.cursor, (T*)a.sorted,
                                                       (T*)a.sorted_w, nullptr, nullptr, nullptr)));
  const int ngr = (g.fb >= 0 ? 2 : 1) * (g.interlace ? 2 : 1);
  const size_t lds = (size_t)ngr * g.tdim[0] * g.tdim[1] * g.tdim[2] * sizeof(double);
  const int64_t ntiles = (int64_t)g.ntile[0] * g.ntile[1] * g.ntile[2];
  TIMED(a.timer, "deposit",
        (k_deposit<T, ORDER><<<(unsigned)ntiles, DEP_T, lds, st>>>(
            (const T*)a.sorted, (const T*)a.sorted_w, a.offsets, g, (T*)a.grid[0],
            (T*)a.grid[1], (T*)a.grid[2], (T*)a.grid[3], a.stats)));
  return hipGetLastError() == hipSuccess ? CAPOW_OK : CAPOW_ERR_HIP;
}

int run_deposit(DepositArgs& a, bool fp64) {
  switch (a.g->order * 2 + (fp64 ? 1 : 0)) {
    case 2: return run_deposit_t<float, 1>(a);
    case 3: return run_deposit_t<double, 1>(a);
    case 4: return run_deposit_t<float, 2>(a);
    case 5: return run_deposit_t<double, 2>(a);
    case 6: return run_deposit_t<float, 3>(a);
    case 7: return run_deposit_t<double, 3>(a);
  }
  return CAPOW_ERR_ARG;
}

}  // namespace capow
