// deposit.hip -- mass assignment (K1) + mean subtraction (K2) for gfx950.
//
// Replaces the per-particle scatter of add_to_grid (reference
// src/read_data.c:514-745, called from src/read_data.c:91-130), the halo add
// (src/read_data.c:152-175) and the mean subtraction (src/main.c:211-225).
//
// Design (one GPU holds the whole grid):
//   1. k_bucket<COUNT>: per particle, apply the RSD shift, decide acceptance
//      of the main and interlaced deposits, find the deposit tiles its taps
//      can reach (its own tile, plus the neighbour tile along each axis when
//      the particle sits within the tap reach of that tile's edge: at most 8
//      tiles) and count one copy per tile; block partials of the global sums
//      (N gridded, sum w v^n) and the max |field| per field.
//   2. k_stats_final: fixed-order reduction of the block partials, alpha,
//      per-field fixed-point scales.
//   3. exclusive scan of the per-tile copy counts.
//   4. k_bucket<SCATTER>: counting-sort scatter of the shifted particles, one
//      copy per reached tile, so every tile's particles are one contiguous
//      range.
//   5. k_deposit: one workgroup per output tile reads its range, accumulates
//      every tap that lands in the tile into LDS, then writes the tile once
//      (value - alpha for density fields).  No grid memset, no global
//      atomics on the grid.
// Determinism: every tap contribution is rounded to an integer multiple of
// a per-field quantum 2^-e (e from the field's max |value|, so one tap is
// < 2^41 quanta) and accumulated in LDS as an integer-valued double.  Sums
// of such values below 2^53 quanta are exact, so the grid does not depend
// on the order in which particles or taps are added.
#include <type_traits>

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

template <typename T>
__device__ __forceinline__ void store2(T* p, T a, T b) {
  if constexpr (sizeof(T) == 4) {
    float2 q;
    q.x = a; q.y = b;
    *reinterpret_cast<float2*>(p) = q;
  } else {
    double2 q;
    q.x = a; q.y = b;
    *reinterpret_cast<double2*>(p) = q;
  }
}

// Per-particle classification shared by the count and scatter passes.
template <typename T, int ORDER>
struct PClass {
  T x, y, z;       // position after the RSD shift
  bool acc_main, acc_int;
  int64_t bucket;  // -1: rejected by every deposit
  bool reach_ok;
  int tile[3][2];  // reached tiles per axis
  int nt[3];       // how many (1 or 2)
};

// Every kept tap must lie within [key+reach_lo, key+reach_hi] (mod N on
// periodic axes), otherwise a tile that owns it would not read the particle.
template <typename T, int ORDER>
__device__ __forceinline__ bool reach_ok(const AxisOut<T>& o, int key, int n, const Geom& g) {
  bool ok = true;
#pragma unroll
  for (int t = 0; t < ORDER; t++) {
    if (o.w[t] == T(0)) continue;
    int off = o.first + t - key;
    if (g.periodic) {
      off %= n;
      if (off < 0) off += n;
      if (off > n / 2) off -= n;
    }
    if (off < g.reach_lo || off > g.reach_hi) ok = false;
  }
  return ok;
}

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
  AxisOut<T> mx = axis_taps<T, ORDER>(x, lo0, hi0, d0, L0, g.n[0], true, g.los, g.edge, (T)g.invd[0]);
  AxisOut<T> my = axis_taps<T, ORDER>(y, lo1, hi1, d1, L1, g.n[1], false, g.los, g.edge, (T)g.invd[1]);
  AxisOut<T> mz = axis_taps<T, ORDER>(z, lo2, hi2, d2, L2, g.n[2], false, g.los, g.edge, (T)g.invd[2]);
  c.acc_main = mx.ok && my.ok && mz.ok;
  AxisOut<T> ix = mx, iy = my, iz = mz;
  c.acc_int = false;
  if (g.interlace) {
    ix = axis_taps<T, ORDER>(x + d0 / T(2), lo0, hi0, d0, L0, g.n[0], true, g.los, g.edge, (T)g.invd[0]);
    iy = axis_taps<T, ORDER>(y + d1 / T(2), lo1, hi1, d1, L1, g.n[1], false, g.los, g.edge, (T)g.invd[1]);
    iz = axis_taps<T, ORDER>(z + d2 / T(2), lo2, hi2, d2, L2, g.n[2], false, g.los, g.edge, (T)g.invd[2]);
    c.acc_int = ix.ok && iy.ok && iz.ok;
  }
  c.reach_ok = true;
  if (!c.acc_main && !c.acc_int) {
    c.bucket = -1;
    return c;
  }
  int key[3];
  key[0] = axis_key<T, ORDER>(x, lo0, hi0, d0, L0, g.n[0], true, g.los, g.edge, (T)g.invd[0]);
  key[1] = axis_key<T, ORDER>(y, lo1, hi1, d1, L1, g.n[1], false, g.los, g.edge, (T)g.invd[1]);
  key[2] = axis_key<T, ORDER>(z, lo2, hi2, d2, L2, g.n[2], false, g.los, g.edge, (T)g.invd[2]);
  if (c.acc_main) {
    c.reach_ok &= reach_ok<T, ORDER>(mx, key[0], g.n[0], g);
    c.reach_ok &= reach_ok<T, ORDER>(my, key[1], g.n[1], g);
    c.reach_ok &= reach_ok<T, ORDER>(mz, key[2], g.n[2], g);
  }
  if (c.acc_int) {
    c.reach_ok &= reach_ok<T, ORDER>(ix, key[0], g.n[0], g);
    c.reach_ok &= reach_ok<T, ORDER>(iy, key[1], g.n[1], g);
    c.reach_ok &= reach_ok<T, ORDER>(iz, key[2], g.n[2], g);
  }
  // Deposit tiles reached along each axis: the key's tile, plus the tile
  // above when the key lies within reach_hi of its upper edge, plus the tile
  // below when it lies within -reach_lo of its lower edge.
  const int lw = -g.reach_lo, hw = g.reach_hi;
#pragma unroll
  for (int a = 0; a < 3; a++) {
    const int td = g.tdim[a], nt = g.ntile[a];
    const int t0 = key[a] / td;
    const int local = key[a] - t0 * td;
    int second = -1;
    if (nt > 1) {
      if (local >= td - hw) {
        second = t0 + 1;
        if (second >= nt) second = g.periodic ? second - nt : -1;
      } else if (local < lw) {
        second = t0 - 1;
        if (second < 0) second = g.periodic ? second + nt : -1;
      }
    }
    c.tile[a][0] = t0;
    c.tile[a][1] = second;
    c.nt[a] = second >= 0 ? 2 : 1;
  }
  c.bucket = 0;
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

// Adds one to ctr[key] for every active lane with has = true.  Lanes of a
// wavefront that share a key are merged into one atomic (the leader adds the
// group size); after 4 groups the remaining lanes fall back to one atomic
// each.  RET: returns each lane's slot (old value + rank within its group).
template <bool RET>
__device__ __forceinline__ uint32_t wave_inc(uint32_t* __restrict__ ctr, int key, bool has) {
  const int lane = __lane_id();
  const unsigned long long lt = (1ull << lane) - 1ull;
  unsigned long long pending = __ballot(has);
  uint32_t slot = 0;
  for (int it = 0; pending; it++) {
    if (it == 4) {
      if ((pending >> lane) & 1ull) {
        if (RET) slot = atomicAdd(&ctr[key], 1u);
        else atomicAdd(&ctr[key], 1u);
      }
      break;
    }
    const int leader = __ffsll((long long)pending) - 1;
    const int lk = __builtin_amdgcn_readlane(key, leader);
    const unsigned long long m = __ballot(has && key == lk) & pending;
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(&ctr[lk], (uint32_t)__popcll(m));
    if (RET) {
      base = (uint32_t)__builtin_amdgcn_readlane((int)base, leader);
      if ((m >> lane) & 1ull) slot = base + (uint32_t)__popcll(m & lt);
    }
    pending &= ~m;
  }
  return slot;
}

template <typename T, int ORDER, bool SCATTER>
__global__ __launch_bounds__(256) void k_bucket(const T* __restrict__ parts,
                                                const T* __restrict__ wts, int64_t n,
                                                Geom g, uint32_t* __restrict__ counts,
                                                uint32_t* __restrict__ cursor,
                                                const uint32_t* __restrict__ off,
                                                T* __restrict__ sorted,
                                                T* __restrict__ sorted_w, int64_t cap,
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
    const bool take = c.bucket >= 0;
    // one copy per reached tile, in a fixed order (iz, iy, ix)
    for (int q = 0; q < 8; q++) {
      const int qx = q & 1, qy = (q >> 1) & 1, qz = q >> 2;
      const bool has = take && qx < c.nt[0] && qy < c.nt[1] && qz < c.nt[2];
      if (!__any(has)) continue;
      int tile = -1;
      if (has) {
        const int tx = qx ? c.tile[0][1] : c.tile[0][0];
        const int ty = qy ? c.tile[1][1] : c.tile[1][0];
        const int tz = qz ? c.tile[2][1] : c.tile[2][0];
        tile = (tz * g.ntile[1] + ty) * g.ntile[0] + tx;
      }
      if (SCATTER) {
        // own copies fill each tile's range from the front, ghost copies
        // from the back, so deposit waves are mostly homogeneous
        uint32_t pos;
        if (q == 0) {
          const uint32_t r = wave_inc<true>(cursor, tile, has);
          pos = has ? off[tile] + r : 0u;
        } else {
          const uint32_t r = wave_inc<true>(cursor + g.nbuckets, tile, has);
          pos = has ? off[tile + 1] - 1u - r : 0u;
        }
        if (has && (int64_t)pos < cap) {
          store4<T>(sorted, pos, c.x, c.y, c.z, v);
          if (g.has_w) sorted_w[pos] = w;
        }
      } else {
        wave_inc<false>(counts, tile, has);
      }
    }
    if (!SCATTER) {
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
#pragma unroll
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
      stats[STAT_GMAX_A + f] = m;
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
#define DEP_T 512

// Accumulator policy.  int (FP32 form): each tap is rounded to an int32
// multiple of a per-tile quantum q chosen so that (copies in the tile) x
// max|field| x (max tap weight) / q <= 2^30 -- no cell can overflow, and
// integer sums are exact.  double (FP64 form): each tap is rounded to a
// multiple of 2^-e, e from the field's max |value| so that one tap is below
// 2^40 quanta; sums below 2^53 quanta are exact in ds_add_f64.
template <typename ACC>
struct Quant;
template <>
struct Quant<int> {
  double scale, inv;
  __device__ void init(uint32_t count, double gmax, double wmax, double /*gscale*/) {
    const double bound = (double)(count ? count : 1) * (gmax > 0 ? gmax : 1.0) * wmax;
    int ex;
    frexp(bound, &ex);                 // bound < 2^ex
    scale = ldexp(1.0, 30 - ex);       // bound * scale < 2^30
    inv = ldexp(1.0, ex - 30);
  }
  template <typename T>
  __device__ __forceinline__ int q(T vs, T w) const { return __float2int_rn((float)(vs * w)); }
};
template <>
struct Quant<double> {
  double scale, inv;
  __device__ void init(uint32_t, double, double, double gscale) {
    scale = gscale;
    inv = 1.0 / gscale;   // power of two: exact
  }
  template <typename T>
  __device__ __forceinline__ double q(T vs, T w) const { return rint((double)vs * (double)w); }
};

// z-index swizzle of LDS tile row `rowi`: an even XOR mask below TZ when TZ
// is a power of two, so rows that differ only in (x, y) land in different
// LDS banks (without it every row starts on one of two bank groups); z pairs
// (2j, 2j+1) stay adjacent.
__device__ __forceinline__ int zswz(int rowi, int TZ) {
  return ((TZ & (TZ - 1)) == 0 && TZ >= 4) ? ((rowi << 1) & (TZ - 2)) : 0;
}

// One workgroup per TX x TY x TZ tile.  The tile's particle copies (own
// first, then ghosts) are one contiguous range of the sorted array.  Phase 0
// deposits the main grids (A, B), phase 1 the interlaced grids: LDS holds
// one phase, i.e. nf x ncell accumulators.
template <typename T, typename ACC, int ORDER>
__global__ __launch_bounds__(DEP_T) void k_deposit(const T* __restrict__ sorted,
                                                   const T* __restrict__ sorted_w,
                                                   const uint32_t* __restrict__ off,
                                                   Geom g, T* __restrict__ gA,
                                                   T* __restrict__ gAi, T* __restrict__ gB,
                                                   T* __restrict__ gBi,
                                                   const double* __restrict__ stats) {
  extern __shared__ unsigned char smem[];
  ACC* acc = reinterpret_cast<ACC*>(smem);  // [nf][ncell]
  const int TX = g.tdim[0], TY = g.tdim[1], TZ = g.tdim[2];
  const int ncell = TX * TY * TZ;
  const bool hasB = g.fb >= 0;
  const int nf = hasB ? 2 : 1;

  const int64_t tileid = blockIdx.x;
  const int ttx = (int)(tileid % g.ntile[0]);
  const int tty = (int)((tileid / g.ntile[0]) % g.ntile[1]);
  const int ttz = (int)(tileid / ((int64_t)g.ntile[0] * g.ntile[1]));
  const int t0x = ttx * TX, t0y = tty * TY, t0z = ttz * TZ;
  const uint32_t s0 = off[tileid], s1 = off[tileid + 1];

  const double wmax = ORDER == 3 ? 0.421875 : 1.0;
  Quant<ACC> qa, qb;
  qa.init(s1 - s0, stats[STAT_GMAX_A], wmax, stats[STAT_SCALE_A]);
  qb.init(s1 - s0, stats[STAT_GMAX_B], wmax, stats[STAT_SCALE_B]);
  const T scA = (T)qa.scale, scB = (T)qb.scale;
  const T d0 = (T)g.d[0], d1 = (T)g.d[1], d2 = (T)g.d[2];
  const double alpha = stats[STAT_ALPHA];

  for (int ph = 0; ph < (g.interlace ? 2 : 1); ph++) {
    for (int c = threadIdx.x; c < nf * ncell; c += DEP_T) acc[c] = ACC(0);
    __syncthreads();
    for (uint32_t i = s0 + threadIdx.x; i < s1; i += DEP_T) {
      T x, y, z, v;
      load4<T>(sorted, i, x, y, z, v);
      const T w = g.has_w ? sorted_w[i] : T(1);
      const T va = w * vpow<T>(v, g.fa) * scA;
      const T vb = hasB ? w * vpow<T>(v, g.fb) * scB : T(0);
      if (ph) {
        x += d0 / T(2);
        y += d1 / T(2);
        z += d2 / T(2);
      }
      const AxisOut<T> ax = axis_taps<T, ORDER>(x, (T)g.lo[0], (T)g.hi[0], d0, (T)g.L[0], g.n[0], true, g.los, g.edge, (T)g.invd[0]);
      const AxisOut<T> ay = axis_taps<T, ORDER>(y, (T)g.lo[1], (T)g.hi[1], d1, (T)g.L[1], g.n[1], false, g.los, g.edge, (T)g.invd[1]);
      const AxisOut<T> az = axis_taps<T, ORDER>(z, (T)g.lo[2], (T)g.hi[2], d2, (T)g.L[2], g.n[2], false, g.los, g.edge, (T)g.invd[2]);
      if (!(ax.ok && ay.ok && az.ok)) continue;
      // tile-local tap coordinates; weight 0 marks taps outside the tile
      int lx[ORDER], ly[ORDER], lz[ORDER];
      T wx[ORDER], wy[ORDER], wz[ORDER];
#pragma unroll
      for (int t = 0; t < ORDER; t++) {
        int gx = ax.first + t, gy = ay.first + t, gz = az.first + t;
        if (g.periodic) {
          gx += gx < 0 ? g.n[0] : (gx >= g.n[0] ? -g.n[0] : 0);
          gy += gy < 0 ? g.n[1] : (gy >= g.n[1] ? -g.n[1] : 0);
          gz += gz < 0 ? g.n[2] : (gz >= g.n[2] ? -g.n[2] : 0);
        }
        lx[t] = gx - t0x;
        ly[t] = gy - t0y;
        lz[t] = gz - t0z;
        wx[t] = ((unsigned)lx[t] < (unsigned)TX) ? ax.w[t] : T(0);
        wy[t] = ((unsigned)ly[t] < (unsigned)TY) ? ay.w[t] : T(0);
        wz[t] = ((unsigned)lz[t] < (unsigned)TZ) ? az.w[t] : T(0);
      }
#pragma unroll
      for (int tx = 0; tx < ORDER; tx++) {
#pragma unroll
        for (int ty = 0; ty < ORDER; ty++) {
          const T wxy = wx[tx] * wy[ty];
          if (wxy == T(0)) continue;
          const int rowi = lx[tx] * TY + ly[ty];
          const int row = rowi * TZ;
          const int swz = zswz(rowi, TZ);
          const T pa = va * wxy, pb = vb * wxy;
#pragma unroll
          for (int tz = 0; tz < ORDER; tz++) {
            if (wz[tz] == T(0)) continue;
            const int cell = row + (lz[tz] ^ swz);
            atomicAdd(&acc[cell], qa.q(pa, wz[tz]));
            if (hasB) atomicAdd(&acc[ncell + cell], qb.q(pb, wz[tz]));
          }
        }
      }
    }
    __syncthreads();
    // write this phase's grids once: value * q - alpha for density fields
    T* dA = ph ? gAi : gA;
    T* dB = ph ? gBi : gB;
    const double meanA = g.mean_sub_a ? alpha : 0.0;
    for (int c2 = threadIdx.x; c2 < ncell / 2; c2 += DEP_T) {
      const int c = 2 * c2;
      const int lz = c % TZ, rowi = c / TZ;
      const int ly = rowi % TY, lxx = rowi / TY;
      const int64_t gi = ((int64_t)(t0x + lxx) * g.n[1] + (t0y + ly)) * g.rowlen + (t0z + lz);
      const int cs = rowi * TZ + (lz ^ zswz(rowi, TZ));   // lz even, swizzle even: pair stays adjacent
      T a0 = (T)((double)acc[cs] * qa.inv - meanA), a1 = (T)((double)acc[cs + 1] * qa.inv - meanA);
      store2<T>(dA + gi, a0, a1);
      if (hasB) {
        T b0 = (T)((double)acc[ncell + cs] * qb.inv), b1 = (T)((double)acc[ncell + cs + 1] * qb.inv);
        store2<T>(dB + gi, b0, b1);
      }
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------
template <typename T, int ORDER>
static int run_count_t(DepositArgs& a) {
  const Geom& g = *a.g;
  const int nb = a.count_blocks;
  hipStream_t st = a.stream;
  TIMED(a.timer, "bucket_count",
        (k_bucket<T, ORDER, false><<<nb, 256, 0, st>>>((const T*)a.parts, (const T*)a.wts, a.n, g,
                                                        a.counts, nullptr, nullptr, nullptr, nullptr, 0,
                                                        a.partials, a.gmax, a.errs)));
  TIMED(a.timer, "stats_final",
        (k_stats_final<<<1, 256, 0, st>>>(a.partials, nb, a.gmax,
                                         (double)g.n[0] * (double)g.n[1] * (double)g.n[2],
                                         a.stats)));
  // scan counts -> offsets; zero the own/ghost cursors
  int64_t nbk = g.nbuckets;
  int64_t nblk = (nbk + SCAN_CHUNK - 1) / SCAN_CHUNK;
  if (nblk > SCAN_CHUNK) return CAPOW_ERR_UNSUPPORTED;
  TIMED(a.timer, "scan", (k_scan_reduce<<<(int)nblk, SCAN_T, 0, st>>>(a.counts, nbk, a.bsum)));
  TIMED(a.timer, "scan", (k_scan_apply<<<1, SCAN_T, 0, st>>>(a.bsum, nblk, nullptr, a.boff, nullptr)));
  TIMED(a.timer, "scan", (k_scan_apply<<<(int)nblk, SCAN_T, 0, st>>>(a.counts, nbk, a.boff, a.offsets,
                                                                 nullptr)));
  if (hipMemsetAsync(a.cursor, 0, (size_t)nbk * 2 * sizeof(uint32_t), st) != hipSuccess)
    return CAPOW_ERR_HIP;
  return hipGetLastError() == hipSuccess ? CAPOW_OK : CAPOW_ERR_HIP;
}

template <typename T, int ORDER>
static int run_scatter_t(DepositArgs& a) {
  const Geom& g = *a.g;
  const int nb = a.count_blocks;
  hipStream_t st = a.stream;
  TIMED(a.timer, "bucket_scatter",
        (k_bucket<T, ORDER, true><<<nb, 256, 0, st>>>((const T*)a.parts, (const T*)a.wts, a.n, g,
                                                       nullptr, a.cursor, a.offsets, (T*)a.sorted,
                                                       (T*)a.sorted_w, a.sorted_cap, nullptr,
                                                       nullptr, nullptr)));
  typedef typename std::conditional<sizeof(T) == 4, int, double>::type ACC;
  const int nf = g.fb >= 0 ? 2 : 1;
  const size_t lds = (size_t)nf * g.tdim[0] * g.tdim[1] * g.tdim[2] * sizeof(ACC);
  const int64_t ntiles = (int64_t)g.ntile[0] * g.ntile[1] * g.ntile[2];
  if (lds > 64 * 1024)
    hipFuncSetAttribute((const void*)k_deposit<T, ACC, ORDER>,
                        hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  TIMED(a.timer, "deposit",
        (k_deposit<T, ACC, ORDER><<<(unsigned)ntiles, DEP_T, lds, st>>>(
            (const T*)a.sorted, (const T*)a.sorted_w, a.offsets, g, (T*)a.grid[0],
            (T*)a.grid[1], (T*)a.grid[2], (T*)a.grid[3], a.stats)));
  return hipGetLastError() == hipSuccess ? CAPOW_OK : CAPOW_ERR_HIP;
}

#define DISPATCH(fn)                                \
  switch (a.g->order * 2 + (fp64 ? 1 : 0)) {        \
    case 2: return fn<float, 1>(a);                 \
    case 3: return fn<double, 1>(a);                \
    case 4: return fn<float, 2>(a);                 \
    case 5: return fn<double, 2>(a);                \
    case 6: return fn<float, 3>(a);                 \
    case 7: return fn<double, 3>(a);                \
  }                                                 \
  return CAPOW_ERR_ARG;

int run_count(DepositArgs& a, bool fp64) { DISPATCH(run_count_t) }
int run_scatter_deposit(DepositArgs& a, bool fp64) { DISPATCH(run_scatter_t) }

}  // namespace capow
