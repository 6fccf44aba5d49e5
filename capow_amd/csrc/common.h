// common.h -- shared device/host definitions of the gfx950 P_l(k) path.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "capow.h"

namespace capow {

// Fast unsigned 32-bit division by a runtime constant (round-up method).
struct FastDiv {
  uint32_t d, m, s;
  __host__ void init(uint32_t div) {
    d = div;
    s = 0;
    while ((1ull << s) < div) s++;
    m = (uint32_t)(((1ull << 32) * ((1ull << s) - div)) / div + 1);
  }
  __device__ __forceinline__ uint32_t div(uint32_t x) const {
    uint32_t t = __umulhi(x, m);
    return (t + ((x - t) >> (s ? 1 : 0))) >> (s ? s - 1 : 0);
  }
};

// ---------------------------------------------------------------------------
// Grid geometry and mass-assignment setup, passed by value to kernels.
// ---------------------------------------------------------------------------
struct Geom {
  int n[3];              // NX NY NZ
  int nzh;               // NZ/2+1 complex per row
  int64_t rowc;          // complex elements per row in device memory: NZ/2+1
                         // rounded up to a 128-byte multiple, so every row
                         // starts on a cache line
  int64_t rowlen;        // 2*rowc reals per padded row
  double lo[3], hi[3];   // box
  double d[3];           // cell size (MAX-MIN)/N  (src/main.c:1126-1128)
  double invd[3];        // 1/d (FP32 form)
  double L[3];           // MAX-MIN
  int order;             // 1 NGP, 2 CIC, 3 TSC
  int interlace;         // 0/1
  int los;               // 0..3
  int edge;              // CAPOW_EDGE_*
  int periodic;          // 1: taps wrap mod N (reference LOS=0 or EDGE_PERIODIC)
  int fa, fb;            // field A = w v^fa, field B = w v^fb (fb < 0: none)
  int mean_sub_a;        // A is mean-subtracted (src/main.c:214)
  int has_w;             // explicit weights
  int rsd_axis;          // -1, or 0/1/2 for LOS 1/2/3
  double rsd_f1, rsd_den;// shift = v*f1/den   (src/read_data.c:67-76)
  // deposit tiling
  int tdim[3];           // tile extent per axis (cells)
  FastDiv tdiv[3];       // division by tdim
  int ntile[3];          // tiles per axis
  int reach_lo, reach_hi;// taps of a particle lie in [key+reach_lo, key+reach_hi]
  int64_t nbuckets;      // ntiles (one bucket per deposit tile)
  int sort_probes;       // LDS hash probes per key in the sort (0: always global)
  // stacked batches (many catalogues of one parameter set on one set of
  // grids, catalogue-major): the tiles and buckets of catalogue m follow
  // those of catalogue m - 1, and its grids start mock_stride reals after
  // those of catalogue m - 1
  int64_t ntiles_mock;   // tiles per catalogue (every tile: one catalogue)
  int64_t mock_stride;   // reals per grid per catalogue (0: one catalogue)
};

// ---------------------------------------------------------------------------
// Per-axis tap rules.
//
// For one coordinate u along one axis this returns whether the reference's
// add_to_grid accepts it, the first tap index (unwrapped; taps are
// first, first+1, ..., first+order-1) and the order weights, with weight 0
// for every tap the reference drops.  Restates src/read_data.c:514-745 at
// NTask = 1 (xmin = XMIN, xmax = XMAX, nx = NX + InterpOrder) followed by the
// halo add of src/read_data.c:152-175 (LOS = 0: buffer slices wrap onto
// slices 0..InterpOrder-1; LOS != 0: buffer dropped):
//   NGP  accept [MIN, MAX); tap (int)s                          (519-541)
//   CIC  accept [MIN, MAX); taps (int)s, +1, weights 1-f, f     (544-605)
//   TSC  LOS = 0, x: u += L when u-MIN < d, accept [MIN+d, MAX+d)
//        LOS != 0, x: accept [MIN, MAX+d); y, z: accept [MIN, MAX)
//        centre c = (int)(s+0.5), f = s-c, weights 0.5(0.5-f)^2,
//        0.75-f^2, 0.5(0.5+f)^2; the low tap is kept only when its
//        index, after the LOS = 0 wrap of -1 to N-1, is > 0      (608-743)
// Periodic axes wrap taps mod N; non-periodic axes drop taps outside
// [0, N).  EDGE_PERIODIC is textbook periodic assignment: positions wrapped
// into the box, every tap kept.
// `key` is the particle's bucket cell along the axis: the main-grid centre
// (TSC) or floor cell (CIC/NGP), wrapped (periodic) or clamped.
// ---------------------------------------------------------------------------
template <typename T>
struct AxisOut {
  int first;    // first tap (unwrapped)
  T w[3];
  T dw[3];      // d w / d s (scaled coordinate), zero for dropped taps
  bool ok;
};

template <typename T>
__device__ __forceinline__ int fl2i(T v) { return (int)v; }  // trunc toward 0


// A particle row (x, y, z, v_los) as one 16- or 32-byte access.
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

// Centre (TSC) or floor (CIC, NGP) cell of one coordinate and whether the
// reference accepts it; the one place where the cell arithmetic lives, so
// the sort keys and the deposit taps agree bit for bit.  The arithmetic runs
// in the working precision T: the box geometry (FP64 in Geom) is rounded to
// T first, and the interlacing shift sh (0 or d/2) is added to u in T.
template <typename T>
struct AxisCell {
  T s;      // scaled coordinate (u + sh - MIN)/d after the TSC periodic x shift
  int c;    // unwrapped cell: TSC (int)(s + 0.5), CIC/NGP (int)s
  bool ok;
};

// Scaled coordinate (u - lo)/d.  FP64 divides exactly as the reference;
// the FP32 form multiplies by the reciprocal (cell membership within an ulp
// of a cell edge is part of the FP32 form's error).
template <typename T>
__device__ __forceinline__ T scaled(T r, T d, T invd) {
#pragma clang fp contract(off)
  if constexpr (sizeof(T) == 4) return r * invd;
  else return r / d;
}

template <typename T, int ORDER>
__device__ __forceinline__ AxisCell<T> axis_cell(T u_in, double sh_d, double lo_d, double hi_d,
                                                 double d_d, double L_d, int n, bool is_x,
                                                 int los, int edge, double invd_d) {
  // No multiply-add contraction: the sort, the deposit and the gradient
  // gather inline this into different code, and must round alike.
#pragma clang fp contract(off)
  const T lo = (T)lo_d, hi = (T)hi_d, d = (T)d_d, L = (T)L_d, invd = (T)invd_d;
  T u = u_in + (T)sh_d;
  AxisCell<T> o;
  if (edge == CAPOW_EDGE_PERIODIC) {
    // wrap into [lo, hi)
    T r = u - lo;
    if (r < T(0) || r >= L) {
      r = r - L * floor(r / L);
      if (r >= L) r = T(0);
    }
    o.s = scaled<T>(r, d, invd);
    o.ok = true;
    o.c = (ORDER == 3) ? (int)floor(o.s + T(0.5)) : (int)floor(o.s);
    return o;
  }
  const bool periodic = (los == 0);
  if (ORDER == 3) {
    if (is_x) {
      if (periodic) {
        u = (u - lo < d) ? u + L : u;
        o.ok = !((u < lo + d) || (u >= hi + d));
      } else {
        o.ok = !((u < lo) || (u >= hi + d));
      }
    } else {
      o.ok = !((u < lo) || (u >= hi));
    }
    o.s = scaled<T>(u - lo, d, invd);
    o.c = fl2i<T>(o.s + T(0.5));
  } else {
    o.ok = !((u < lo) || (u >= hi));
    o.s = scaled<T>(u - lo, d, invd);
    o.c = fl2i<T>(o.s);
    if (ORDER == 1 && !is_x && o.c == n) o.c = n - 1;   // src/read_data.c:534-535
  }
  return o;
}

// Taps of one coordinate (with the interlacing shift sh) along axis a of g.
template <typename T, int ORDER>
__device__ __forceinline__ AxisOut<T> axis_taps(T u, double sh, const Geom& g, int a, int los,
                                                int edge) {
#pragma clang fp contract(off)
  const int n = g.n[a];
  const AxisCell<T> ac = axis_cell<T, ORDER>(u, sh, g.lo[a], g.hi[a], g.d[a], g.L[a], n, a == 0,
                                             los, edge, g.invd[a]);
  AxisOut<T> o;
  o.ok = ac.ok;
  o.w[0] = o.w[1] = o.w[2] = T(0);
  o.dw[0] = o.dw[1] = o.dw[2] = T(0);
  const int c = ac.c;
  const T f = ac.s - (T)c;
  const bool periodic = (edge == CAPOW_EDGE_PERIODIC) || (los == 0);
  if (ORDER == 3) {
    o.first = c - 1;
    const T wl = T(0.5) * (T(0.5) - f) * (T(0.5) - f);
    const T wc = T(0.75) - f * f;
    const T wh = T(0.5) * (T(0.5) + f) * (T(0.5) + f);
    bool k0 = true, k1 = true, k2 = true;
    if (edge != CAPOW_EDGE_PERIODIC) {
      // The reference keeps the low tap only when its index is > 0 after the
      // LOS = 0 wrap (-1 -> N-1): periodic axes drop it exactly when c == 1
      // (x: c-1 = slice 0; y/z: iy-1 = 0), non-periodic axes whenever c <= 1.
      // Non-periodic axes also drop taps at index >= N (x: dropped buffer
      // slices; y, z: the < N guards).
      k0 = (periodic ? (c != 1) : (c - 1 > 0)) && (periodic || c - 1 < n);
      k1 = periodic || c < n;
      k2 = periodic || c + 1 < n;
    }
    o.w[0] = k0 ? wl : T(0);
    o.w[1] = k1 ? wc : T(0);
    o.w[2] = k2 ? wh : T(0);
    o.dw[0] = k0 ? -(T(0.5) - f) : T(0);
    o.dw[1] = k1 ? T(-2) * f : T(0);
    o.dw[2] = k2 ? T(0.5) + f : T(0);
  } else if (ORDER == 2) {
    o.first = c;
    const bool k0 = periodic || c < n, k1 = periodic || c + 1 < n;
    o.w[0] = k0 ? T(1) - f : T(0);
    o.w[1] = k1 ? f : T(0);
    o.dw[0] = k0 ? T(-1) : T(0);
    o.dw[1] = k1 ? T(1) : T(0);
  } else {
    o.first = c;
    o.w[0] = (periodic || c < n) ? T(1) : T(0);
  }
  return o;
}

// Edge rules fixed at compile time for the sort and deposit kernels, so
// that the per-particle rules above compile to straight-line code:
// EK_RUNTIME reads Geom::edge, ::los and ::periodic; EK_PERIODIC is
// CAPOW_EDGE_PERIODIC; EK_REF_LOS0 the reference rules with LOS = 0 (taps
// wrap, the TSC x shift); EK_REF_LOS the reference rules with LOS != 0
// (taps outside [0, N) dropped; the survey deposit too).
enum { EK_RUNTIME = 0, EK_PERIODIC = 1, EK_REF_LOS0 = 2, EK_REF_LOS = 3 };
template <int EK>
__device__ __forceinline__ int ek_edge(const Geom& g) {
  return EK == EK_RUNTIME ? g.edge : (EK == EK_PERIODIC ? CAPOW_EDGE_PERIODIC : CAPOW_EDGE_REFERENCE);
}
// only "LOS == 0 or not" matters to the reference rules (the periodic rules
// ignore it; the out-of-box check reads it)
template <int EK>
__device__ __forceinline__ int ek_los(const Geom& g) {
  return EK == EK_REF_LOS0 ? 0 : (EK == EK_REF_LOS ? 3 : g.los);
}
template <int EK>
__device__ __forceinline__ bool ek_periodic(const Geom& g) {
  return EK == EK_RUNTIME ? g.periodic != 0 : EK != EK_REF_LOS;
}
inline int edge_kind(const Geom& g) {
  return g.edge == CAPOW_EDGE_PERIODIC ? EK_PERIODIC : (g.los == 0 ? EK_REF_LOS0 : EK_REF_LOS);
}

// Bucket key cell along an axis: the cell of axis_cell, wrapped mod N on
// periodic axes (accepted coordinates give c in [-1, N + 2]), clamped to
// [0, N) otherwise.
__device__ __forceinline__ int cell_key(int c, int n, bool periodic) {
  if (periodic) {
    // one period either side by a select; further out (never for accepted
    // coordinates) by the remainder
    int r = c >= n ? c - n : (c < 0 ? c + n : c);
    if ((unsigned)r >= (unsigned)n) {
      r = c % n;
      if (r < 0) r += n;
    }
    return r;
  }
  return c < 0 ? 0 : (c >= n ? n - 1 : c);
}

}  // namespace capow
