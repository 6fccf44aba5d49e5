// deposit.hip -- mass assignment (K1) + mean subtraction (K2) for gfx950.
//
// Replaces the per-particle scatter of add_to_grid (reference
// src/read_data.c:514-745, called from src/read_data.c:91-130), the halo add
// (src/read_data.c:152-175) and the mean subtraction (src/main.c:211-225).
//
// Design (one GPU holds the whole grid):
//   1. k_bucket<COUNT>: per particle, apply the RSD shift, decide acceptance
//      of the main and interlaced deposits, find its key: its tile and its
//      border class (per axis: whether its taps reach the tile below, the
//      tile above, or neither; 27 classes) and count it; block partials of
//      the global sums (N gridded, sum w v^n) and the max |field| per field.
//   2. k_stats_final: fixed-order reduction of the block partials, alpha,
//      per-field fixed-point scales.
//   3. exclusive scan of the per-(tile, class) counts.
//   4. k_bucket<SCATTER>: counting-sort scatter of the shifted particles, one
//      copy each, so every (tile, class) is one contiguous range and every
//      tile's own particles are the contiguous range of its 27 classes.
//   5. k_deposit: one workgroup per output tile reads its own range and, from
//      each of its 26 neighbours, the class ranges whose taps reach it (63
//      ranges in all), accumulates every tap that lands in the tile into
//      LDS, then writes the tile once (value - alpha for density fields).
//      No grid memset, no global atomics on the grid.
// Determinism: every tap contribution is rounded to an integer multiple of
// a quantum (FP32 form: per tile and field, int32 taps in one int64
// accumulator per cell; FP64 form: per field, integer-valued doubles), and
// integer sums do not depend on the order in which particles or taps are
// added, so the grid does not depend on the sort order.
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

// FP32: a non-temporal store (the grid is read back only by the FFT's first
// pass, long after the write-out; streaming it past the caches measured
// about 3% faster for the whole deposit).  FP64: a plain store (no gain).
template <typename T>
__device__ __forceinline__ void store2(T* p, T a, T b) {
  if constexpr (sizeof(T) == 4) {
    typedef float f2 __attribute__((ext_vector_type(2)));
    f2 q;
    q.x = a; q.y = b;
    __builtin_nontemporal_store(q, reinterpret_cast<f2*>(p));
  } else {
    double2 q;
    q.x = a; q.y = b;
    *reinterpret_cast<double2*>(p) = q;
  }
}

// Per-particle classification shared by the count and scatter passes.
// The sort key is tile * 27 + border class: per axis, class 0 when the
// particle's taps reach the tile below (key cell within -reach_lo of the
// tile's lower edge), 2 when they reach the tile above (within reach_hi of
// the upper edge), 1 otherwise (and always 1 along an axis with a single
// tile); class = cx*9 + cy*3 + cz.  Every particle has exactly one sorted
// copy; the deposit of a tile reads its own 27 classes and, from each
// neighbour, the classes that reach it.
template <typename T, int ORDER>
struct PClass {
  T x, y, z;       // position after the RSD shift
  bool acc_main, acc_int;
  bool take;       // accepted by at least one deposit
  bool reach_ok;
  uint32_t key;    // tile * 27 + class
};

// Canonical periodic offset in [n/2 - n + 1, n/2] of an offset within one
// period either side.
__device__ __forceinline__ int canon_off(int off, int n) {
  const int h = n / 2;
  off = off > h ? off - n : off;
  off = off <= h - n ? off + n : off;
  return off;
}

// Taps of one deposit (main or interlaced) along an axis span
// [first, first + ORDER - 1] (dropped taps only narrow this, and
// non-periodic axes drop taps outside [0, N)); all of them must lie within
// [key + reach_lo, key + reach_hi], otherwise a tile that owns a tap would
// not hold a copy of the particle.
template <int ORDER>
__device__ __forceinline__ bool reach_ok(int first, int key, int n, const Geom& g, bool per) {
  int a = first, b = first + ORDER - 1;
  if (!per) {
    a = a < 0 ? 0 : a;
    b = b > n - 1 ? n - 1 : b;
    if (a > b) return true;   // every tap dropped
    return a - key >= g.reach_lo && b - key <= g.reach_hi;
  }
  return canon_off(a - key, n) >= g.reach_lo && canon_off(b - key, n) <= g.reach_hi;
}

template <typename T, int ORDER, int EK>
__device__ __forceinline__ PClass<T, ORDER> classify(T x, T y, T z, T v, const Geom& g) {
  const int edge = ek_edge<EK>(g), los = ek_los<EK>(g);
  PClass<T, ORDER> c;
  if (g.rsd_axis >= 0) {
    // tx += tvr*(1+z)/(100 sqrt(Om (1+z)^3 + 1 - Om))   src/read_data.c:67-76
    T sh = (v * (T)g.rsd_f1) / (T)g.rsd_den;
    if (g.rsd_axis == 0) x += sh;
    else if (g.rsd_axis == 1) y += sh;
    else z += sh;
  }
  c.x = x; c.y = y; c.z = z;
  const T u[3] = {x, y, z};
  bool okm = true, oki = (g.interlace != 0);
  int cm[3], ci[3];
#pragma unroll
  for (int a = 0; a < 3; a++) {
    const AxisCell<T> m = axis_cell<T, ORDER>(u[a], 0.0, g.lo[a], g.hi[a], g.d[a], g.L[a], g.n[a],
                                              a == 0, los, edge, g.invd[a]);
    okm = okm && m.ok;
    cm[a] = m.c;
    ci[a] = m.c;
    if (g.interlace) {
      const AxisCell<T> i = axis_cell<T, ORDER>(u[a], g.d[a] / 2.0, g.lo[a], g.hi[a], g.d[a], g.L[a],
                                                g.n[a], a == 0, los, edge, g.invd[a]);
      oki = oki && i.ok;
      ci[a] = i.c;
    }
  }
  c.acc_main = okm;
  c.acc_int = oki;
  c.take = okm || oki;
  c.reach_ok = true;
  // the key's tile, and per axis whether the taps reach the tile above
  // (the key lies within reach_hi of its upper edge) or below (within
  // -reach_lo of its lower edge)
  const int lw = -g.reach_lo, hw = g.reach_hi;
  const bool per = ek_periodic<EK>(g);
  int tc[3], cl[3];
#pragma unroll
  for (int a = 0; a < 3; a++) {
    const int n = g.n[a], td = g.tdim[a], nt = g.ntile[a];
    const int key = cell_key(cm[a], n, per);
    const int t0 = (int)g.tdiv[a].div((uint32_t)key);
    const int local = key - t0 * td;
    cl[a] = 1;
    if (nt > 1) {
      const int first_m = ORDER == 3 ? cm[a] - 1 : cm[a];
      const int first_i = ORDER == 3 ? ci[a] - 1 : ci[a];
      const bool r = (!okm || reach_ok<ORDER>(first_m, key, n, g, per)) &&
                     (!oki || reach_ok<ORDER>(first_i, key, n, g, per));
      c.reach_ok = c.reach_ok && r;
      cl[a] = local >= td - hw ? 2 : (local < lw ? 0 : 1);
    }
    tc[a] = t0;
  }
  const uint32_t tile = ((uint32_t)tc[2] * (uint32_t)g.ntile[1] + (uint32_t)tc[1]) *
                            (uint32_t)g.ntile[0] + (uint32_t)tc[0];
  c.key = tile * 27u + (uint32_t)(cl[0] * 9 + cl[1] * 3 + cl[2]);
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

// Block-local aggregation of the counting sort.  A block takes a chunk of
// BK_CHUNK consecutive particles; every particle's key (tile * 27 + border
// class) is counted in an LDS hash table (open addressing, linear probing).
// Runs of lanes with the same key are merged first (the run head adds the
// run length), so a slot sees one LDS atomic per run.  At the end of the
// chunk each occupied slot issues ONE global atomic for all its copies:
// with catalogues stored in spatial order (simulation snapshots, lattices)
// a chunk reaches a few dozen tiles, so the device-scope atomics drop to one
// per (chunk, key).  A key that finds no free slot within BK_PROBE probes
// falls back to the global counter.
#define BK_T 256
#define BK_K 8
#define BK_HS_BITS 11
#define BK_HS (1 << BK_HS_BITS)
#define BK_PROBE 32
#define BK_EMPTY 0xffffffffu
static_assert(BK_T * BK_K == BK_CHUNK, "chunk size");
static_assert(BK_T * BK_K <= (1 << (32 - BK_HS_BITS - 1)), "rank field");

__device__ __forceinline__ int hs_find(uint32_t* __restrict__ hkey, uint32_t key, int probes) {
  uint32_t h = (key * 2654435761u) >> (32 - BK_HS_BITS);
#pragma unroll 1
  for (int p = 0; p < probes; p++) {
    const uint32_t cur = hkey[h];
    if (cur == key) return (int)h;
    if (cur == BK_EMPTY) {
      const uint32_t old = atomicCAS(&hkey[h], BK_EMPTY, key);
      if (old == BK_EMPTY || old == key) return (int)h;
    }
    h = (h + 1) & (BK_HS - 1);
  }
  return -1;
}

// Counts one copy of `key` for every lane with has = true.  Lanes are
// merged into runs: a run is a maximal block of consecutive lanes with has
// and the same key (catalogues stored in spatial order give long runs).
// Only the run heads touch the LDS table, each adding its run length; a
// head whose key finds no free slot adds the run to the global counter
// `gctr` (the key's own counter) instead.  For RET the other lanes take the
// head's slot and base with one permute and add their offset in the run.
// Returns the slot, or -1 when the run went to gctr (for RET, `rank` is
// then the lane's rank among gctr's copies); rank: the lane's rank among
// the slot's copies in this chunk.
template <bool RET>
__device__ __forceinline__ int hs_add(uint32_t* __restrict__ hkey, uint32_t* __restrict__ hcnt,
                                      uint32_t key, bool has, uint32_t* gctr, int probes,
                                      uint32_t& rank) {
  const int lane = __lane_id();
  const uint32_t pkey = (uint32_t)__shfl_up((int)key, 1);
  const unsigned long long hm = __ballot(has);
  const bool phas = lane > 0 && ((hm >> (lane - 1)) & 1ull);
  const bool head = has && !(phas && pkey == key);
  const unsigned long long heads = __ballot(head);
  // run end: the next lane that starts a run or has no copy
  const unsigned long long stops = heads | ~hm;
  int slot = -1;
  uint32_t base = 0;
  if (head) {
    const unsigned long long after = lane == 63 ? 0ull : (stops >> (lane + 1));
    const int len = after ? __ffsll((long long)after) : 64 - lane;
    slot = hs_find(hkey, key, probes);
    if (slot >= 0) {
      if (RET) base = atomicAdd(&hcnt[slot], (uint32_t)len);
      else atomicAdd(&hcnt[slot], (uint32_t)len);
    } else {
      if (RET) base = atomicAdd(gctr, (uint32_t)len);
      else atomicAdd(gctr, (uint32_t)len);
    }
  }
  rank = 0;
  if (RET) {
    // the lane's head: the highest head at or below it
    const unsigned long long upto = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);
    const unsigned long long below = heads & upto;
    const int h = below ? 63 - __clzll((long long)below) : lane;
    const int hs = __shfl(slot, h);
    const uint32_t hb = (uint32_t)__shfl((int)base, h);
    if (has) {
      slot = hs;
      rank = hb + (uint32_t)(lane - h);
    }
  }
  return slot;
}

// IDX (scatter only): also write each sorted row's input index (the
// backward's tile gather needs it; the forward does not pay for it).
template <typename T, int ORDER, bool SCATTER, int EK, bool IDX = false>
__global__ __launch_bounds__(BK_T) void k_bucket(const T* __restrict__ parts,
                                                 const T* __restrict__ wts, int64_t n,
                                                 Geom g, uint32_t* __restrict__ counts,
                                                 uint32_t* __restrict__ cursor,
                                                 const uint32_t* __restrict__ off,
                                                 T* __restrict__ sorted,
                                                 T* __restrict__ sorted_w, int64_t cap,
                                                 double* __restrict__ partials,
                                                 unsigned int* __restrict__ errs,
                                                 uint32_t* __restrict__ sorted_idx,
                                                 const MockDev* __restrict__ mocks) {
  __shared__ uint32_t hkey[BK_HS];
  __shared__ uint32_t hcnt[BK_HS];   // copies per slot; in the scatter pass, then the slot's base
  __shared__ double sh[BK_T];
  // stacked batch: catalogue blockIdx.y from the table, its buckets after
  // those of the catalogues before it
  const int64_t mock = blockIdx.y;
  if (mocks) {
    const MockDev md = mocks[mock];
    parts = (const T*)md.parts;
    wts = (const T*)md.w;
    n = md.n;
  }
  const uint32_t kbase = (uint32_t)(mock * g.ntiles_mock * 27);
  double s_ngrid = 0, s_w = 0, s_v1 = 0, s_v2 = 0, s_v3 = 0, s_v4 = 0, s_out = 0;
  unsigned int oob = 0, reach_bad = 0;
  const int64_t nchunks = (n + BK_CHUNK - 1) / BK_CHUNK;
  // blocks of this catalogue: one per chunk, at most the launch's width
  // (the count a launch for this catalogue alone has)
  const int64_t nbm = max((int64_t)1, min(nchunks, (int64_t)gridDim.x));
  if ((int64_t)blockIdx.x >= nbm) return;
  for (int64_t ch = blockIdx.x; ch < nchunks; ch += nbm) {
    for (int s = threadIdx.x; s < BK_HS; s += BK_T) {
      hkey[s] = BK_EMPTY;
      hcnt[s] = 0;
    }
    __syncthreads();
    // scatter pass: the chunk's particles wait in registers for their slot
    // bases
    T px[BK_K], py[BK_K], pz[BK_K], pv[BK_K], pw[BK_K];
    uint32_t pr[BK_K];   // slot | rank << BK_HS_BITS; BK_EMPTY: none / written
#pragma unroll
    for (int k = 0; k < BK_K; k++) {
      const int64_t i = ch * BK_CHUNK + (int64_t)k * BK_T + threadIdx.x;
      const bool valid = i < n;
      T x = T(0), y = T(0), z = T(0), v = T(0), w = T(1);
      if (valid) {
        load4<T>(parts, i, x, y, z, v);
        if (g.has_w && wts) w = wts[i];
      }
      PClass<T, ORDER> c = classify<T, ORDER, EK>(x, y, z, v, g);
      const bool take = valid && c.take;
      px[k] = c.x; py[k] = c.y; pz[k] = c.z; pv[k] = v; pw[k] = w;
      pr[k] = BK_EMPTY;
      const uint32_t key = take ? kbase + c.key : 0u;
      uint32_t rank;
      uint32_t* gctr = SCATTER ? cursor + key : counts + key;
      const int slot = hs_add<SCATTER>(hkey, hcnt, key, take, gctr, g.sort_probes, rank);
      if (SCATTER && take) {
        if (slot < 0) {
          // table full: the run took its ranks from the global cursor
          const uint32_t pos = off[key] + rank;
          if ((int64_t)pos < cap) {
            store4<T>(sorted, pos, c.x, c.y, c.z, v);
            if (g.has_w) sorted_w[pos] = w;
            if (IDX) sorted_idx[pos] = (uint32_t)i;
          }
        } else {
          pr[k] = (uint32_t)slot | (rank << BK_HS_BITS);
        }
      }
      if (!SCATTER && valid) {
        if (!c.reach_ok) reach_bad++;
        // src/read_data.c:81-88 (on the shifted position)
        if ((c.x < (T)g.lo[0]) || (c.x > (T)g.hi[0]) || (c.y < (T)g.lo[1]) ||
            (c.y > (T)g.hi[1]) || (c.z < (T)g.lo[2]) || (c.z > (T)g.hi[2])) {
          if (ek_los<EK>(g) == 0) oob++;
          else s_out += 1.0;
        }
        if (c.acc_main) s_ngrid += (double)w;
        // moments over every particle read (src/read_data.c:132-138)
        s_w += (double)w;
        s_v1 += (double)(w * v);
        s_v2 += (double)(w * v * v);
        s_v3 += (double)(w * v * v * v);
        s_v4 += (double)(w * v * v * v * v);
      }
    }
    __syncthreads();
    // one global atomic per occupied slot
    for (int s = threadIdx.x; s < BK_HS; s += BK_T) {
      const uint32_t key = hkey[s];
      if (key == BK_EMPTY) continue;
      if (SCATTER) hcnt[s] = atomicAdd(&cursor[key], hcnt[s]);
      else atomicAdd(&counts[key], hcnt[s]);
    }
    if (SCATTER) {
      __syncthreads();
#pragma unroll
      for (int k = 0; k < BK_K; k++) {
        const uint32_t e = pr[k];
        if (e == BK_EMPTY) continue;
        const uint32_t slot = e & (BK_HS - 1);
        const uint32_t key = hkey[slot];
        const uint32_t pos = off[key] + hcnt[slot] + (e >> BK_HS_BITS);
        if ((int64_t)pos < cap) {
          store4<T>(sorted, pos, px[k], py[k], pz[k], pv[k]);
          if (g.has_w) sorted_w[pos] = pw[k];
          if (IDX) sorted_idx[pos] = (uint32_t)(ch * BK_CHUNK + (int64_t)k * BK_T + threadIdx.x);
        }
      }
    }
    __syncthreads();
  }
  if (!SCATTER) {
    double r[7] = {s_ngrid, s_w, s_v1, s_v2, s_v3, s_v4, s_out};
#pragma unroll
    for (int q = 0; q < 7; q++) {
      double t = block_sum<BK_T>(r[q], sh);
      if (threadIdx.x == 0) partials[(mock * gridDim.x + blockIdx.x) * STAT_N + q] = t;
    }
    if (oob) atomicAdd(&errs[mock * 4 + 0], oob);
    if (reach_bad) atomicAdd(&errs[mock * 4 + 1], reach_bad);
  }
}

// Fixed-order reduction of block partials -> stats[STAT_N] and alpha.
// Stacked batch (mocks set): catalogue blockIdx.x, whose partials are the
// first blocks (as many as its k_bucket used) of its nblocks rows.
__global__ void k_stats_final(const double* __restrict__ partials, int nblocks, double ncells,
                              double* __restrict__ stats, const MockDev* __restrict__ mocks) {
  __shared__ double sh[256];
  partials += (int64_t)blockIdx.x * nblocks * STAT_N;
  stats += (int64_t)blockIdx.x * STAT_TOTAL;
  if (mocks) {
    const int64_t nch = (mocks[blockIdx.x].n + BK_CHUNK - 1) / BK_CHUNK;
    nblocks = (int)max((int64_t)1, min(nch, (int64_t)nblocks));
  }
  for (int q = 0; q < 7; q++) {
    double acc = 0.0;
    // each thread sums a strided subset in fixed order, then a fixed tree
    for (int b = threadIdx.x; b < nblocks; b += 256) acc += partials[(int64_t)b * STAT_N + q];
    double t = block_sum<256>(acc, sh);
    if (threadIdx.x == 0) stats[q] = t;
  }
  if (threadIdx.x == 0) {
    stats[STAT_ALPHA] = stats[STAT_NGRID] / ncells;  // src/main.c:117
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

// Exclusive scan of n block sums by one workgroup, chunk after chunk with a
// running carry (any n).
__global__ __launch_bounds__(SCAN_T) void k_scan_carry(const uint32_t* __restrict__ in, int64_t n,
                                                       uint32_t* __restrict__ out) {
  __shared__ uint32_t sh[SCAN_T];
  uint32_t carry = 0;
  for (int64_t c0 = 0; c0 < n; c0 += SCAN_CHUNK) {
    const int64_t base = c0 + (int64_t)threadIdx.x * SCAN_PER;
    uint32_t v[SCAN_PER];
    uint32_t sum = 0;
    for (int q = 0; q < SCAN_PER; q++) {
      const int64_t i = base + q;
      v[q] = i < n ? in[i] : 0u;
      sum += v[q];
    }
    uint32_t tot;
    uint32_t ex = block_excl_scan(sum, sh, &tot) + carry;
    for (int q = 0; q < SCAN_PER; q++) {
      const int64_t i = base + q;
      if (i < n) out[i] = ex;
      ex += v[q];
    }
    carry += tot;
  }
}

// ---------------------------------------------------------------------------
// Tile deposit.
// ---------------------------------------------------------------------------

// The particle ranges a tile's deposit reads: its own 27 classes, then for
// each neighbour offset (dx, dy, dz) the classes cx*9 + cy*3 + cz whose taps
// reach back into this tile (per axis: any class for d = 0, class 2 for the
// tile below (d = -1), class 0 for the tile above (d = +1)), merged into
// runs of consecutive classes: 1 + 2 x-faces + 2*3 y-faces + 2*9 z-faces +
// 4 xy-edges + 4*3 xz-edges + 4*3 yz-edges + 8 corners = 63 runs.
struct DepRange {
  signed char d[3];
  unsigned char lo, hi;   // classes [lo, hi) of the neighbour
};
struct DepRanges {
  DepRange r[64];
  int n;
};
constexpr DepRanges make_dep_ranges() {
  DepRanges t{};
  t.n = 0;
  t.r[t.n++] = DepRange{{0, 0, 0}, 0, 27};
  for (int dx = -1; dx <= 1; dx++)
    for (int dy = -1; dy <= 1; dy++)
      for (int dz = -1; dz <= 1; dz++) {
        if (dx == 0 && dy == 0 && dz == 0) continue;
        const int first = t.n;
        for (int cx = 0; cx < 3; cx++) {
          if (dx != 0 && cx != (dx < 0 ? 2 : 0)) continue;
          for (int cy = 0; cy < 3; cy++) {
            if (dy != 0 && cy != (dy < 0 ? 2 : 0)) continue;
            const int lo = cx * 9 + cy * 3 + (dz == 0 ? 0 : (dz < 0 ? 2 : 0));
            const int hi = dz == 0 ? lo + 3 : lo + 1;
            if (t.n > first && t.r[t.n - 1].hi == lo) {
              t.r[t.n - 1].hi = (unsigned char)hi;   // contiguous: extend the run
            } else {
              t.r[t.n++] = DepRange{{(signed char)dx, (signed char)dy, (signed char)dz},
                                    (unsigned char)lo, (unsigned char)hi};
            }
          }
        }
      }
  return t;
}
constexpr DepRanges kDepRanges = make_dep_ranges();
static_assert(kDepRanges.n == 63, "63 ranges per tile");
__constant__ DepRanges c_dep_ranges = kDepRanges;

#define DEP_T 512
#define DEP_K 4       // particle copies per thread kept in registers for both phases
#define DEP_TRASH 64  // per-lane sink cells for taps outside the tile
// LDS bytes of the largest tile: FP32 16x16x32 (int64 per cell), FP64
// 16x16x16 (two int64 per cell), each with the sink cells
#define DEP_LDS_MAX (2 * (16 * 16 * 16 + DEP_TRASH) * 8)
static_assert((16 * 16 * 32 + DEP_TRASH) * 8 <= DEP_LDS_MAX, "FP32 tile");

// Exact accumulation with a per-tile quantum q (a power of two) chosen so
// that (copies in the tile) x (the tile's largest |field| value) x (max tap
// weight) / q < 2^BITS:
// a cell takes at most one tap per copy, so no partial sum of a field
// leaves (-2^BITS - copies/2, 2^BITS + copies/2) whatever the order.
// FP32 form (BITS 30): fields A and B of a cell share one int64
// accumulator, A * 2^32 + B, updated with one 64-bit LDS atomic per tap
// (the low word read as int32 is B, the rest is A); halo-row tiles without
// a B field hold one int32 per cell (half the LDS, more workgroups per CU).  FP64 form (BITS 62):
// one int64 accumulator per field; a tile of c copies rounds each tap to
// about c x (largest |field|) x 2^-63, below the FP64 rounding of the
// tile's largest values for every tile of fewer than 2^10 copies, no bound
// on how many taps one cell may take.  Integer sums do not depend on the
// order of the additions, so the grids are deterministic.
__device__ __forceinline__ double tile_scale(uint32_t count, double gmax, double wmax, int bits) {
  const double bound = (double)(count ? count : 1) * (gmax > 0 ? gmax : 1.0) * wmax;
  int ex;
  frexp(bound, &ex);               // bound < 2^ex
  return ldexp(1.0, bits - ex);    // bound * scale < 2^bits
}

// floor(x + 0.5) as int32 in one instruction (ties round up); the rounding
// rule only has to be the same for every tap.
__device__ __forceinline__ int round_rpi(float x) {
  int r;
  asm("v_cvt_rpi_i32_f32 %0, %1" : "=v"(r) : "v"(x));
  return r;
}

// XOR mask on the z index of LDS row `rowi` (hash of the row, below TZ when
// TZ is a power of two): taps of neighbouring particles in different rows
// land in different LDS banks.  An aligned cell pair (2j, 2j+1) maps to an
// aligned pair, swapped when the mask is odd.
__device__ __forceinline__ int zmask(int rowi, int TZ) {
  return ((TZ & (TZ - 1)) == 0) ? (int)((__umul24((uint32_t)rowi, 0x9E3779u) >> 12) & (uint32_t)(TZ - 1)) : 0;
}

// One workgroup per TX x TY x TZ tile, reading its own particles first, then
// the neighbours' particles whose taps reach it.  Phase 0 deposits the main
// grids (A, B), phase 1 the interlaced grids: LDS holds one phase.  The tap
// loop has no branches: taps outside the tile go to the lane's sink cell,
// or, with HALO (FP32 16x16x32 tiles, two or more tiles along z), every LDS
// row has DEP_ZLO cells below and DEP_ZHI above the tile's z range, so that
// all z taps of a particle land in its row at consecutive addresses (halo
// cells are not written out), and rows outside the tile in x or y add zeros
// to a row of the tile.  APOS (field A = w v^0 with no weights: A >= 0):
// the int64 accumulator is B * 2^32 + A, so the int32 taps need no sign
// borrow.
//
// BANK (HALO only): before each phase the tile's own copies are reordered
// in LDS so that the 64 lanes of a wave start their taps in evenly spread
// LDS banks, and tap rows outside the tile are masked off instead of adding
// zeros.  A tap is the lane's base cell plus an offset common to all lanes
// ((tx TY + ty) RZ + tz cells), so lanes whose base cells are spread evenly
// over the 16 8-byte bank pairs stay spread over every tap: 64-bit LDS
// atomics from such lanes take about half the time of random ones.  The
// base cell's bank pair is (lx0 TY RZ + ly0 RZ + lz0) mod 16 = (6 ly0 + lz0)
// mod 16 up to a constant (TY RZ = 608 = 0 mod 16, RZ = 38 = 6 mod 16).  A
// counting sort by that key with slots in (rank within key, key) order makes
// consecutive slots cycle through the keys.  The accumulation is exact
// integer arithmetic, so the order changes no result.
#define DEP_ZLO 4   // halo cells below the tile's z range (taps reach -3)
#define DEP_ZHI 2   // above (taps reach TZ + 2: the next row's first cells)
#define DEP_NKEY 16 // bank-pair keys of the BANK order
template <typename T, int ORDER, bool HASB, int TYC, int TZC, bool HALO, bool APOS, int EK,
          bool BANK>
__global__ __launch_bounds__(DEP_T) void k_deposit(const T* __restrict__ sorted,
                                                   const T* __restrict__ sorted_w,
                                                   const uint32_t* __restrict__ off,
                                                   Geom g, T* __restrict__ gA,
                                                   T* __restrict__ gAi, T* __restrict__ gB,
                                                   T* __restrict__ gBi,
                                                   const double* __restrict__ stats,
                                                   int64_t cap, unsigned int* __restrict__ errs) {
  constexpr bool F32 = sizeof(T) == 4;
  // static LDS (its base folds into the tap addresses): the accumulators of
  // the largest tile (TX <= 16; with compile-time extents, of this shape),
  // checked on the host
  constexpr int RZ = TZC + DEP_ZLO + DEP_ZHI;   // HALO row length (cells)
  // HALO without a B field: one int32 accumulator per cell (A alone; the
  // tile's quantum keeps every partial sum inside int32), else the packed
  // int64 of A and B
  constexpr int CB = (HALO && !HASB) ? 4 : 8;   // HALO bytes per cell
  constexpr int LDSB = HALO ? ((16 * TYC * RZ + 2) * CB + 15) / 16 * 16
                            : ((TYC && TZC) ? (16 * TYC * TZC + DEP_TRASH) * 8 * (F32 ? 1 : 2)
                                            : DEP_LDS_MAX);
  static_assert(!HALO || (F32 && TYC && TZC), "halo rows: FP32, compile-time extents");
  static_assert(!BANK || HALO, "bank order: halo rows only");
  // BANK scratch (before a phase's clear): key counts, the (rank, key) slot
  // tables and the reordered copies (x, y, z, a[, b])
  static_assert(!BANK || (DEP_NKEY + 2 * DEP_T) * 4 + (HASB ? 5 : 4) * DEP_T * DEP_K * 4 <= LDSB,
                "bank scratch");
  __shared__ __align__(16) unsigned char smem[LDSB];
  // TYC, TZC: compile-time tile extents (0: the runtime ones)
  const int TX = g.tdim[0];
  const int TY = TYC ? TYC : g.tdim[1], TZ = TZC ? TZC : g.tdim[2];
  const int ncell = TX * TY * TZ;
  const int nacc = ncell + DEP_TRASH;
  // FP32: packed int64 [nacc]; FP64: int64 A [nacc], then B [nacc]
  unsigned long long* acc64 = reinterpret_cast<unsigned long long*>(smem);
  const int lane = threadIdx.x & 63;
  const int trash = ncell + lane;

  // catalogue (stacked batches; 0 otherwise) and its tile
  const int64_t mock = (int64_t)blockIdx.x / g.ntiles_mock;
  const int64_t tileid = (int64_t)blockIdx.x - mock * g.ntiles_mock;
  const int ttx = (int)(tileid % g.ntile[0]);
  const int tty = (int)((tileid / g.ntile[0]) % g.ntile[1]);
  const int ttz = (int)(tileid / ((int64_t)g.ntile[0] * g.ntile[1]));
  const int t0x = ttx * TX, t0y = tty * TY, t0z = ttz * TZ;
  // copies beyond the sorted buffer's capacity were not stored: flag the
  // overflow and deposit only what is there (the result is then discarded)
  if (blockIdx.x == 0 && threadIdx.x == 0 && (int64_t)off[(int64_t)gridDim.x * 27] > cap) errs[2] = 1u;
  // the tile's 63 ranges (wave 0): start and exclusive prefix of the lengths
  __shared__ uint32_t r_start[64], r_pre[65];
  // the tile's max |w v^a|, |w v^b| as bit patterns (non-negative doubles
  // order as their bits)
  __shared__ unsigned long long s_fmax[2];
  if (threadIdx.x < 64) {
    const int r = threadIdx.x;
    if (r == 0) s_fmax[0] = s_fmax[1] = 0ull;
    uint32_t st = 0, len = 0;
    if (r < c_dep_ranges.n) {
      const DepRange d = c_dep_ranges.r[r];
      const int tt[3] = {ttx, tty, ttz};
      int nb[3];
      bool ok = true;
#pragma unroll
      for (int a = 0; a < 3; a++) {
        int q = tt[a] + d.d[a];
        if (q < 0 || q >= g.ntile[a]) {
          if (ek_periodic<EK>(g)) q += q < 0 ? g.ntile[a] : -g.ntile[a];
          else ok = false;
        }
        nb[a] = q;
      }
      if (ok) {
        const int64_t b =
            (mock * g.ntiles_mock + ((int64_t)nb[2] * g.ntile[1] + nb[1]) * g.ntile[0] + nb[0]) * 27;
        const uint32_t s0 = (uint32_t)min((int64_t)off[b + d.lo], cap);
        const uint32_t s1 = (uint32_t)min((int64_t)off[b + d.hi], cap);
        st = s0;
        len = s1 - s0;
      }
    }
    uint32_t inc = len;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t u = (uint32_t)__shfl_up((int)inc, o);
      if (r >= o) inc += u;
    }
    r_start[r] = st;
    r_pre[r] = inc - len;
    if (r == 63) r_pre[64] = inc;
  }
  __syncthreads();
  const uint32_t total = r_pre[64], own = r_pre[1];
  // sorted index of the f-th particle of the tile's ranges (f < total): the
  // last range whose prefix is <= f holds it
  auto pidx = [&](uint32_t f) -> uint32_t {
    if (f < own) return r_start[0] + f;
    int lo = 1, hi = 62;
#pragma unroll 1
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (r_pre[mid] <= f) lo = mid;
      else hi = mid - 1;
    }
    return r_start[lo] + (f - r_pre[lo]);
  };

  const T d1 = (T)g.d[1], d2 = (T)g.d[2];   // bank keys
  const double alpha = stats[mock * STAT_TOTAL + STAT_ALPHA];
  const bool single = total <= (uint32_t)(DEP_T * DEP_K);

  // The quanta come from this tile's own largest field values, so a large
  // value elsewhere in the catalogue does not coarsen this tile's rounding.
  // A tile of at most DEP_T * DEP_K copies keeps them in registers for both
  // phases; a larger one is read once more here.
  T px[DEP_K], py[DEP_K], pz[DEP_K], pa[DEP_K], pb[DEP_K];
  bool ph_has[DEP_K];
  // phase 0's accumulators are cleared while the particles load; the
  // barrier after the maximum covers both
  auto clear_acc = [&]() {
    uint4* z4 = reinterpret_cast<uint4*>(smem);
    const int n4 = HALO ? ((TX * TY * RZ + 2) * CB + 15) / 16 : (F32 ? nacc / 2 : nacc);   // 16-byte words
    for (int c = threadIdx.x; c < n4; c += DEP_T) z4[c] = make_uint4(0u, 0u, 0u, 0u);
  };
  // (BANK: the LDS holds the reorder scratch first; each phase clears)
  if (!BANK) clear_acc();
  double ma = 0.0, mb = 0.0;
  if (single) {
    // every index first, then every load, then the first use: one memory
    // latency per tile instead of one per copy slot
    uint32_t idx[DEP_K];
#pragma unroll
    for (int k = 0; k < DEP_K; k++) {
      const uint32_t f = (uint32_t)(k * DEP_T) + threadIdx.x;
      ph_has[k] = f < total;
      idx[k] = ph_has[k] ? pidx(f) : 0u;
    }
    T pv[DEP_K], pw[DEP_K];
#pragma unroll
    for (int k = 0; k < DEP_K; k++) {
      T x = T(0), y = T(0), z = T(0), v = T(0), w = T(1);
      if (ph_has[k]) {
        load4<T>(sorted, idx[k], x, y, z, v);
        if (g.has_w) w = sorted_w[idx[k]];
      }
      px[k] = x; py[k] = y; pz[k] = z;
      pv[k] = v; pw[k] = w;
    }
#pragma unroll
    for (int k = 0; k < DEP_K; k++) {
      pa[k] = ph_has[k] ? pw[k] * vpow<T>(pv[k], g.fa) : T(0);
      pb[k] = (HASB && ph_has[k]) ? pw[k] * vpow<T>(pv[k], g.fb) : T(0);
      ma = fmax(ma, fabs((double)pa[k]));
      mb = fmax(mb, fabs((double)pb[k]));
    }
  } else {
    for (uint32_t f = threadIdx.x; f < total; f += DEP_T) {
      T x, y, z, v, w = T(1);
      const uint32_t i = pidx(f);
      load4<T>(sorted, i, x, y, z, v);
      if (g.has_w) w = sorted_w[i];
      ma = fmax(ma, fabs((double)(w * vpow<T>(v, g.fa))));
      if (HASB) mb = fmax(mb, fabs((double)(w * vpow<T>(v, g.fb))));
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    ma = fmax(ma, __shfl_xor(ma, o));
    mb = fmax(mb, __shfl_xor(mb, o));
  }
  if (lane == 0) {
    atomicMax(&s_fmax[0], (unsigned long long)__double_as_longlong(ma));
    atomicMax(&s_fmax[1], (unsigned long long)__double_as_longlong(mb));
  }
  __syncthreads();
  const double wmax = ORDER == 3 ? 0.421875 : 1.0;
  const double scA = tile_scale(total, __longlong_as_double((long long)s_fmax[0]), wmax, F32 ? 30 : 62);
  const double scB = tile_scale(total, __longlong_as_double((long long)s_fmax[1]), wmax, F32 ? 30 : 62);
  const double invA = 1.0 / scA, invB = 1.0 / scB;   // powers of two: exact
  if (single) {
#pragma unroll
    for (int k = 0; k < DEP_K; k++) {
      pa[k] *= (T)scA;
      pb[k] *= (T)scB;
    }
  }
  for (int ph = 0; ph < (g.interlace ? 2 : 1); ph++) {
    if constexpr (BANK) {
      if (single && own > 0) {
        uint32_t* s_cnt = reinterpret_cast<uint32_t*>(smem);   // [DEP_NKEY]
        uint32_t* s_T = s_cnt + DEP_NKEY;                      // [DEP_T]
        uint32_t* s_M = s_T + DEP_T;                           // [DEP_T]
        T* s_p = reinterpret_cast<T*>(s_M + DEP_T);            // [5][DEP_T DEP_K]
        constexpr int NS = DEP_T * DEP_K;
        if (threadIdx.x < DEP_NKEY) s_cnt[threadIdx.x] = 0u;
        __syncthreads();
        // key of each own copy: the bank pair of its first tap (for this
        // phase's shift); an approximate cell is enough, the key only
        // spreads the lanes
        const T sh1 = ph ? d1 / T(2) : T(0), sh2 = ph ? d2 / T(2) : T(0);
        const T rnd = ORDER == 3 ? T(0.5) : T(0);
        uint32_t key[DEP_K], rk[DEP_K];
#pragma unroll
        for (int k = 0; k < DEP_K; k++) {
          const uint32_t f = (uint32_t)(k * DEP_T) + threadIdx.x;
          key[k] = 0u;
          rk[k] = 0u;
          if (f < own) {
            const int cy = (int)floor((py[k] + sh1 - (T)g.lo[1]) * (T)g.invd[1] + rnd);
            const int cz = (int)floor((pz[k] + sh2 - (T)g.lo[2]) * (T)g.invd[2] + rnd);
            key[k] = (uint32_t)(6 * cy + cz) & (DEP_NKEY - 1);
            rk[k] = atomicAdd(&s_cnt[key[k]], 1u);
          }
        }
        __syncthreads();
        // slot of (rank r, key b): T[r] = sum_b min(cnt_b, r) copies of lower
        // rank, then the keys below b that still have a copy of rank r
        // (mask M[r]); thread r fills row r
        uint32_t tr = 0u, mr = 0u, mx = 0u;
#pragma unroll
        for (int b = 0; b < DEP_NKEY; b++) {
          const uint32_t cb = s_cnt[b];
          tr += min(cb, (uint32_t)threadIdx.x);
          mr |= (cb > (uint32_t)threadIdx.x ? 1u : 0u) << b;
          mx = max(mx, cb);
        }
        s_T[threadIdx.x] = tr;
        s_M[threadIdx.x] = mr;
        __syncthreads();
        // mx is the same in every thread; ranks beyond the table keep the
        // sorted order
        if (mx <= (uint32_t)DEP_T) {
#pragma unroll
          for (int k = 0; k < DEP_K; k++) {
            const uint32_t f = (uint32_t)(k * DEP_T) + threadIdx.x;
            if (f < own) {
              const uint32_t pos = s_T[rk[k]] + (uint32_t)__popc(s_M[rk[k]] & ((1u << key[k]) - 1u));
              s_p[pos] = px[k];
              s_p[NS + pos] = py[k];
              s_p[2 * NS + pos] = pz[k];
              s_p[3 * NS + pos] = pa[k];
              if (HASB) s_p[4 * NS + pos] = pb[k];
            }
          }
          __syncthreads();
#pragma unroll
          for (int k = 0; k < DEP_K; k++) {
            const uint32_t f = (uint32_t)(k * DEP_T) + threadIdx.x;
            if (f < own) {
              px[k] = s_p[f];
              py[k] = s_p[NS + f];
              pz[k] = s_p[2 * NS + f];
              pa[k] = s_p[3 * NS + f];
              if (HASB) pb[k] = s_p[4 * NS + f];
            }
          }
        }
        __syncthreads();
      }
    }
    if (ph || BANK) {
      clear_acc();
      __syncthreads();
    }
    for (uint32_t base = 0; base < total; base += DEP_T * DEP_K) {
      if (!single) {
#pragma unroll
        for (int k = 0; k < DEP_K; k++) {
          const uint32_t f = base + (uint32_t)(k * DEP_T) + threadIdx.x;
          ph_has[k] = f < total;
          T x = T(0), y = T(0), z = T(0), v = T(0), w = T(1);
          if (ph_has[k]) {
            const uint32_t i = pidx(f);
            load4<T>(sorted, i, x, y, z, v);
            if (g.has_w) w = sorted_w[i];
          }
          px[k] = x; py[k] = y; pz[k] = z;
          pa[k] = w * vpow<T>(v, g.fa) * (T)scA;
          pb[k] = HASB ? w * vpow<T>(v, g.fb) * (T)scB : T(0);
        }
      }
#pragma unroll
      for (int k = 0; k < DEP_K; k++) {
        if (!ph_has[k]) continue;
        const int edge = ek_edge<EK>(g), los = ek_los<EK>(g);
        const bool per = ek_periodic<EK>(g);
        // the interlaced phase's half-cell shift goes into the FP64 cell
        // arithmetic
        const AxisOut<T> ax = axis_taps<T, ORDER>(px[k], ph ? g.d[0] / 2.0 : 0.0, g, 0, los, edge);
        const AxisOut<T> ay = axis_taps<T, ORDER>(py[k], ph ? g.d[1] / 2.0 : 0.0, g, 1, los, edge);
        const AxisOut<T> az = axis_taps<T, ORDER>(pz[k], ph ? g.d[2] / 2.0 : 0.0, g, 2, los, edge);
        const bool ok = ax.ok && ay.ok && az.ok;
        const T va = ok ? pa[k] : T(0), vb = ok ? pb[k] : T(0);
        if constexpr (HALO) {
          // x, y: tile-local tap rows; rows outside the tile add zeros to row
          // `lane` (a row of the tile: TX * TY >= 64).  With two or more
          // tiles along x and y (checked on the host), the first tap's
          // offset from the tile taken as the nearest periodic image places
          // every tap: tap t is in the tile iff that offset + t is in
          // [0, T) (an image one period away lies at least N - T - 2 >= T - 2
          // cells off, outside the tile for every tap).
          // Only a tile on the box's border reads copies across the periodic
          // boundary (a uniform branch).
          int lx0 = ax.first - t0x, ly0 = ay.first - t0y;
          if (per && (ttx == 0 || ttx == g.ntile[0] - 1)) lx0 = canon_off(lx0, g.n[0]);
          if (per && (tty == 0 || tty == g.ntile[1] - 1)) ly0 = canon_off(ly0, g.n[1]);
          // z: the first tap relative to the tile (periodic: the nearest
          // image); a particle of this tile's lists has it in [-3, TZ], the
          // clamp only keeps any other in the halo
          int lz0 = az.first - t0z;
          if (per && (ttz == 0 || ttz == g.ntile[2] - 1)) lz0 = canon_off(lz0, g.n[2]);
          lz0 = lz0 < -3 ? -3 : (lz0 > TZ ? TZ : lz0);
          // LDS byte offsets stay below 2^17: 32-bit unsigned arithmetic
          const uint32_t zb8 = (uint32_t)(lz0 + DEP_ZLO) * (uint32_t)CB;
          const uint32_t lane_row = __umul24((uint32_t)lane, (uint32_t)(RZ * CB)) + zb8;
          // row (lx0 + tx, ly0 + ty) at base + a compile-time step
          int base_b = (lx0 * TY + ly0) * (RZ * CB) + (int)zb8;
          // kept in a register (not recomputed per row with a 32-bit
          // multiply): an empty asm the compiler cannot see through
          asm volatile("" : "+v"(base_b));
#pragma unroll
          for (int tx = 0; tx < ORDER; tx++) {
            const bool vx = (unsigned)(lx0 + tx) < (unsigned)TX;
#pragma unroll
            for (int ty = 0; ty < ORDER; ty++) {
              const bool vxy = vx && (unsigned)(ly0 + ty) < (unsigned)TY;
              // a row outside the tile for every lane (ghost particles of
              // one border class share the pattern): nothing to add
              if (__builtin_amdgcn_ballot_w64(vxy) == 0) continue;
              const uint32_t rowb = (uint32_t)(base_b + (tx * TY + ty) * (RZ * CB));
              // BANK: lanes outside the tile sit this row out (the row's
              // three atomics run under a narrower mask); otherwise they add
              // zeros to a row of the tile
              if (BANK && !vxy) continue;
              const T wxy = vxy ? ax.w[tx] * ay.w[ty] : T(0);
              const T qa0 = va * wxy, qb0 = vb * wxy;
              unsigned char* rb = smem + ((BANK || vxy) ? rowb : lane_row);
#pragma unroll
              for (int tz = 0; tz < ORDER; tz++) {
                const int qa = round_rpi((float)(qa0 * az.w[tz]));
                if constexpr (!HASB) {
                  // two's complement: the int32 sum wraps back into range
                  atomicAdd(reinterpret_cast<unsigned int*>(rb + tz * 4), (unsigned int)qa);
                  continue;
                }
                const int qb = HASB ? round_rpi((float)(qb0 * az.w[tz])) : 0;
                unsigned long long packed;
                if constexpr (APOS) {
                  // B * 2^32 + A, A >= 0
                  packed = ((unsigned long long)(uint32_t)qb << 32) | (uint32_t)qa;
                } else {
                  // A * 2^32 + B with B sign-extended
                  const uint32_t hi = (uint32_t)qa + (uint32_t)(qb >> 31);
                  packed = ((unsigned long long)hi << 32) | (uint32_t)qb;
                }
                atomicAdd(reinterpret_cast<unsigned long long*>(rb + tz * 8), packed);
              }
            }
          }
          continue;
        }
        // tile-local tap coordinates; taps outside the tile are invalid
        int lx[ORDER], ly[ORDER], lz[ORDER], lz8[ORDER];
        bool vx[ORDER], vy[ORDER], vz[ORDER];
#pragma unroll
        for (int t = 0; t < ORDER; t++) {
          int gx = ax.first + t, gy = ay.first + t, gz = az.first + t;
          if (per) {
            gx += gx < 0 ? g.n[0] : (gx >= g.n[0] ? -g.n[0] : 0);
            gy += gy < 0 ? g.n[1] : (gy >= g.n[1] ? -g.n[1] : 0);
            gz += gz < 0 ? g.n[2] : (gz >= g.n[2] ? -g.n[2] : 0);
          }
          lx[t] = gx - t0x;
          ly[t] = gy - t0y;
          lz[t] = gz - t0z;
          vx[t] = (unsigned)lx[t] < (unsigned)TX;
          vy[t] = (unsigned)ly[t] < (unsigned)TY;
          vz[t] = (unsigned)lz[t] < (unsigned)TZ;
          lz8[t] = lz[t] * 8;
        }
#pragma unroll
        for (int tx = 0; tx < ORDER; tx++) {
#pragma unroll
          for (int ty = 0; ty < ORDER; ty++) {
            const bool vxy = vx[tx] && vy[ty];
            if (!__any(vxy)) continue;   // outside the tile for every lane
            const int rowi = lx[tx] * TY + ly[ty];
            const int m8 = zmask(rowi, TZ) * 8;
            const int row8 = rowi * TZ * 8;
            const T wxy = ax.w[tx] * ay.w[ty];
            const T qa0 = va * wxy, qb0 = vb * wxy;
#pragma unroll
            for (int tz = 0; tz < ORDER; tz++) {
              // byte offset of the cell's accumulator (8 bytes per cell)
              const int cell8 = (vxy && vz[tz]) ? row8 + (lz8[tz] ^ m8) : trash * 8;
              if constexpr (F32) {
                const int qa = round_rpi((float)(qa0 * az.w[tz]));
                const int qb = HASB ? round_rpi((float)(qb0 * az.w[tz])) : 0;
                // A * 2^32 + B with B sign-extended
                const uint32_t hi = (uint32_t)qa + (uint32_t)(qb >> 31);
                const unsigned long long packed = ((unsigned long long)hi << 32) | (uint32_t)qb;
                atomicAdd(reinterpret_cast<unsigned long long*>(smem + cell8), packed);
              } else {
                // |tap| < 2^62 quanta: the rounded double converts exactly
                atomicAdd(reinterpret_cast<unsigned long long*>(smem + cell8),
                          (unsigned long long)(long long)rint((double)qa0 * (double)az.w[tz]));
                if (HASB)
                  atomicAdd(reinterpret_cast<unsigned long long*>(smem + (size_t)nacc * 8 + cell8),
                            (unsigned long long)(long long)rint((double)qb0 * (double)az.w[tz]));
              }
            }
          }
        }
      }
    }
    __syncthreads();
    // write this phase's grids once: value * q - alpha for density fields
    T* dA = (ph ? gAi : gA) + mock * g.mock_stride;
    T* dB = (ph ? gBi : gB) + (HASB ? mock * g.mock_stride : 0);
    const double meanA = g.mean_sub_a ? alpha : 0.0;
    // cell pair c2 = (lxx*TY + ly)*(TZ/2) + lz/2, advanced by DEP_T pairs per
    // iteration with carries instead of divisions
    const int TZ2 = TZ / 2;
    const int step_row = DEP_T / TZ2, step_z = DEP_T - step_row * TZ2;
    const int step_x = step_row / TY, step_y = step_row - step_x * TY;
    int zp = (int)threadIdx.x % TZ2, rowi0 = (int)threadIdx.x / TZ2;
    int ly = rowi0 % TY, lxx = rowi0 / TY;
    for (int c2 = threadIdx.x; c2 < ncell / 2; c2 += DEP_T) {
      const int lz = 2 * zp, rowi = lxx * TY + ly;
      const int64_t gi = ((int64_t)(t0x + lxx) * g.n[1] + (t0y + ly)) * g.rowlen + (t0z + lz);
      const int m = HALO ? 0 : zmask(rowi, TZ);
      // aligned pair holding z = lz, lz+1
      const int cs = HALO ? rowi * RZ + DEP_ZLO + lz : rowi * TZ + ((lz ^ m) & ~1);
      const bool swap = (m & 1) != 0;
      zp += step_z;
      ly += step_y;
      lxx += step_x;
      if (zp >= TZ2) { zp -= TZ2; ly += 1; }
      if (ly >= TY) { ly -= TY; lxx += 1; }
      if constexpr (F32 && HALO && !HASB) {
        // int32 cells: the pair at an even cell index (8-byte aligned)
        const uint2 q = *reinterpret_cast<const uint2*>(smem + (size_t)cs * 4);
        store2<T>(dA + gi, (T)((double)(int)q.x * invA - meanA), (T)((double)(int)q.y * invA - meanA));
      } else if constexpr (F32) {
        const ulonglong2 q = *reinterpret_cast<const ulonglong2*>(acc64 + cs);
        const long long p0 = (long long)(swap ? q.y : q.x), p1 = (long long)(swap ? q.x : q.y);
        int a0, a1, b0, b1;
        if (APOS && HALO) {
          a0 = (int)(unsigned)(p0 & 0xffffffffll);
          a1 = (int)(unsigned)(p1 & 0xffffffffll);
          b0 = (int)(p0 >> 32);
          b1 = (int)(p1 >> 32);
        } else {
          b0 = (int)(unsigned)(p0 & 0xffffffffll);
          b1 = (int)(unsigned)(p1 & 0xffffffffll);
          a0 = (int)((p0 - (long long)b0) >> 32);
          a1 = (int)((p1 - (long long)b1) >> 32);
        }
        store2<T>(dA + gi, (T)((double)a0 * invA - meanA), (T)((double)a1 * invA - meanA));
        if (HASB) store2<T>(dB + gi, (T)((double)b0 * invB), (T)((double)b1 * invB));
      } else {
        const longlong2 qa = *reinterpret_cast<const longlong2*>(acc64 + cs);
        const double a0 = (double)(swap ? qa.y : qa.x), a1 = (double)(swap ? qa.x : qa.y);
        store2<T>(dA + gi, (T)(a0 * invA - meanA), (T)(a1 * invA - meanA));
        if (HASB) {
          const longlong2 qb = *reinterpret_cast<const longlong2*>(acc64 + nacc + cs);
          const double b0 = (double)(swap ? qb.y : qb.x), b1 = (double)(swap ? qb.x : qb.y);
          store2<T>(dB + gi, (T)(b0 * invB), (T)(b1 * invB));
        }
      }
    }
    __syncthreads();
  }
}

// Zero fill by a kernel: a stream a caller captures into a graph holds no
// memset nodes.
__global__ void k_zero_u32(uint32_t* __restrict__ p, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 0u;
}

int zero_u32(uint32_t* p, int64_t n, hipStream_t st) {
  if (n <= 0) return CAPOW_OK;
  int64_t blocks = (n + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  k_zero_u32<<<(unsigned)blocks, 256, 0, st>>>(p, n);
  return hipGetLastError() == hipSuccess ? CAPOW_OK : CAPOW_ERR_HIP;
}

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------
// Calls f with the edge kind as a compile-time constant.
template <typename F>
static void launch_edge(int ek, F&& f) {
  if (ek == EK_PERIODIC) f(std::integral_constant<int, EK_PERIODIC>());
  else if (ek == EK_REF_LOS0) f(std::integral_constant<int, EK_REF_LOS0>());
  else f(std::integral_constant<int, EK_REF_LOS>());
}
template <typename T, int ORDER>
static int run_count_t(DepositArgs& a) {
  const Geom& g = *a.g;
  const int nb = a.count_blocks;
  hipStream_t st = a.stream;
  const dim3 grid((unsigned)nb, (unsigned)a.nmock);
  TIMED(a.timer, "bucket_count",
        (launch_edge(edge_kind(g), [&](auto ek) {
          k_bucket<T, ORDER, false, decltype(ek)::value><<<grid, BK_T, 0, st>>>(
              (const T*)a.parts, (const T*)a.wts, a.n, g, a.counts, nullptr, nullptr, nullptr, nullptr,
              0, a.partials, a.errs, nullptr, a.mocks);
        })));
  TIMED(a.timer, "stats_final",
        (k_stats_final<<<a.nmock, 256, 0, st>>>(a.partials, nb,
                                               (double)g.n[0] * (double)g.n[1] * (double)g.n[2],
                                               a.stats, a.mocks)));
  // scan counts -> offsets (every catalogue's buckets in one scan); zero
  // the own/ghost cursors
  int64_t nbk = g.nbuckets * a.nmock;
  int64_t nblk = (nbk + SCAN_CHUNK - 1) / SCAN_CHUNK;
  TIMED(a.timer, "scan", (k_scan_reduce<<<(int)nblk, SCAN_T, 0, st>>>(a.counts, nbk, a.bsum)));
  TIMED(a.timer, "scan", (k_scan_carry<<<1, SCAN_T, 0, st>>>(a.bsum, nblk, a.boff)));
  TIMED(a.timer, "scan", (k_scan_apply<<<(int)nblk, SCAN_T, 0, st>>>(a.counts, nbk, a.boff, a.offsets,
                                                                 nullptr)));
  if (zero_u32(a.cursor, nbk, st) != CAPOW_OK)
    return CAPOW_ERR_HIP;
  return hipGetLastError() == hipSuccess ? CAPOW_OK : CAPOW_ERR_HIP;
}

template <typename T, int ORDER>
static int run_scatter_t(DepositArgs& a) {
  const Geom& g = *a.g;
  const int nb = a.count_blocks;
  hipStream_t st = a.stream;
  const dim3 sgrid((unsigned)nb, (unsigned)a.nmock);
  TIMED(a.timer, "bucket_scatter",
        (launch_edge(edge_kind(g), [&](auto ek) {
          if (a.sorted_idx)
            k_bucket<T, ORDER, true, decltype(ek)::value, true><<<sgrid, BK_T, 0, st>>>(
                (const T*)a.parts, (const T*)a.wts, a.n, g, nullptr, a.cursor, a.offsets,
                (T*)a.sorted, (T*)a.sorted_w, a.sorted_cap, nullptr, nullptr, a.sorted_idx, a.mocks);
          else
            k_bucket<T, ORDER, true, decltype(ek)::value, false><<<sgrid, BK_T, 0, st>>>(
                (const T*)a.parts, (const T*)a.wts, a.n, g, nullptr, a.cursor, a.offsets,
                (T*)a.sorted, (T*)a.sorted_w, a.sorted_cap, nullptr, nullptr, nullptr, a.mocks);
        })));
  const int ncell = g.tdim[0] * g.tdim[1] * g.tdim[2];
  const size_t nacc = (size_t)ncell + DEP_TRASH;
  const size_t lds = sizeof(T) == 4 ? nacc * 8 : 2 * nacc * 8;
  const int64_t ntiles = (int64_t)g.ntile[0] * g.ntile[1] * g.ntile[2] * a.nmock;
  const bool hasB = g.fb >= 0;
  if (lds > DEP_LDS_MAX) return CAPOW_ERR_UNSUPPORTED;
  // the default tile shapes get compile-time extents (shifts instead of
  // integer multiplies in the tap addressing)
  const int ty = g.tdim[1], tz = g.tdim[2];
  const bool t1632 = ty == 16 && tz == 32, t1616 = ty == 16 && tz == 16;
#define DEP_LAUNCH_EK(HB, TYC, TZC, HALO, APOS, EK, BK)                                        \
  TIMED(a.timer, "deposit",                                                                    \
        (k_deposit<T, ORDER, HB, TYC, TZC, HALO, APOS, EK, BK><<<(unsigned)ntiles, DEP_T, 0, st>>>( \
            (const T*)a.sorted, (const T*)a.sorted_w, a.offsets, g, (T*)a.grid[0],              \
            (T*)a.grid[1], (T*)a.grid[2], (T*)a.grid[3], a.stats, a.sorted_cap, a.errs)))
#define DEP_LAUNCH(HB, TYC, TZC, HALO, APOS) \
  DEP_LAUNCH_EK(HB, TYC, TZC, HALO, APOS, EK_RUNTIME, false)
#define DEP_LAUNCH_HALO_B(HB, APOS, BK)                                                         \
  switch (edge_kind(g)) {                                                                      \
    case EK_PERIODIC: DEP_LAUNCH_EK(HB, 16, 32, true, APOS, EK_PERIODIC, BK); break;           \
    case EK_REF_LOS0: DEP_LAUNCH_EK(HB, 16, 32, true, APOS, EK_REF_LOS0, BK); break;           \
    default: DEP_LAUNCH_EK(HB, 16, 32, true, APOS, EK_REF_LOS, BK); break;                     \
  }
#define DEP_LAUNCH_HALO(HB, APOS)                 \
  if (a.bank) { DEP_LAUNCH_HALO_B(HB, APOS, true) } \
  else { DEP_LAUNCH_HALO_B(HB, APOS, false) }
  // halo rows: FP32 16x16x32 tiles with two or more tiles along every axis
  // (one tile along z wraps its own z taps; along x or y the first tap's
  // nearest image does not place every tap); A >= 0 when it is the plain
  // count (no weights)
  const bool halo = a.halo && sizeof(T) == 4 && t1632 && g.ntile[0] >= 2 && g.ntile[1] >= 2 &&
                    g.ntile[2] >= 2 && g.tdim[0] * 16 >= 64;
  const bool apos = !g.has_w && g.fa == 0;
  bool done = false;
  if constexpr (sizeof(T) == 4) {
    if (halo) {
      if (hasB && apos) { DEP_LAUNCH_HALO(true, true) }
      else if (hasB) { DEP_LAUNCH_HALO(true, false) }
      else if (apos) { DEP_LAUNCH_HALO(false, true) }
      else { DEP_LAUNCH_HALO(false, false) }
      done = true;
    }
  }
  if (done) {
  } else if (hasB) {
    if (sizeof(T) == 4 && t1632) DEP_LAUNCH(true, 16, 32, false, false);
    else if (sizeof(T) == 8 && t1616) DEP_LAUNCH(true, 16, 16, false, false);
    else DEP_LAUNCH(true, 0, 0, false, false);
  } else {
    if (sizeof(T) == 4 && t1632) DEP_LAUNCH(false, 16, 32, false, false);
    else if (sizeof(T) == 8 && t1616) DEP_LAUNCH(false, 16, 16, false, false);
    else DEP_LAUNCH(false, 0, 0, false, false);
  }
#undef DEP_LAUNCH
#undef DEP_LAUNCH_EK
#undef DEP_LAUNCH_HALO
#undef DEP_LAUNCH_HALO_B
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
