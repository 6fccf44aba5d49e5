// fft.hip -- in-place forward r2c 3-D FFT (K3) for gfx950.
//
// Replaces fftw_execute of the in-place r2c plans (reference
// src/main.c:1200-1217, executed at src/main.c:266-271): unnormalised,
// e^{-i k.x}, real input [NX][NY][NZ] in rows padded to 2(NZ/2+1), complex
// output [NX][NY][NZ/2+1] in the same memory (FFTW's non-transposed layout).
//
// Three axis passes, each one workgroup per batch of 1-D transforms held in
// LDS (radix 2/3/4/5/8 Stockham, butterflies in registers, one LDS buffer):
//   z: rows are contiguous; a length-NZ real row is transformed as NZ/2
//      complex values followed by the even/odd split, so a row is read and
//      written once;
//   y, x: B consecutive complex columns (64 B of contiguous memory per
//      axis index) are gathered, transformed along the strided axis and
//      written back in place.
// Twiddles: exp(-2 pi i j / N) generated in double on the host, stored in the
// working precision.
#include "binning_math.h"
#include "common.h"
#include "kernels.h"
#include "twiddle_table.h"

namespace capow {

template <typename T>
struct C2;
template <>
struct C2<float> { typedef float2 t; };
template <>
struct C2<double> { typedef double2 t; };

template <typename T>
using cx = typename C2<T>::t;

template <typename T>
__device__ __forceinline__ cx<T> mk(T a, T b) { cx<T> r; r.x = a; r.y = b; return r; }
template <typename T>
__device__ __forceinline__ cx<T> add(cx<T> a, cx<T> b) { return mk<T>(a.x + b.x, a.y + b.y); }
template <typename T>
__device__ __forceinline__ cx<T> sub(cx<T> a, cx<T> b) { return mk<T>(a.x - b.x, a.y - b.y); }
template <typename T>
__device__ __forceinline__ cx<T> mul(cx<T> a, cx<T> b) {
  return mk<T>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <typename T>
__device__ __forceinline__ cx<T> mul_mi(cx<T> a) { return mk<T>(a.y, -a.x); }  // a * (-i)

// Forward DFT of size R in registers: out[q] = sum_r v[r] exp(-2 pi i r q / R)
template <typename T, int R>
__device__ __forceinline__ void butterfly(cx<T>* v) {
  if (R == 2) {
    cx<T> a = v[0], b = v[1];
    v[0] = add<T>(a, b);
    v[1] = sub<T>(a, b);
  } else if (R == 4) {
    cx<T> s02 = add<T>(v[0], v[2]), d02 = sub<T>(v[0], v[2]);
    cx<T> s13 = add<T>(v[1], v[3]), d13 = sub<T>(v[1], v[3]);
    v[0] = add<T>(s02, s13);
    v[2] = sub<T>(s02, s13);
    v[1] = add<T>(d02, mul_mi<T>(d13));
    v[3] = sub<T>(d02, mul_mi<T>(d13));
  } else if (R == 8) {
    const T h = (T)0.70710678118654752440084436210485;
    cx<T> e[4] = {v[0], v[2], v[4], v[6]};
    cx<T> o[4] = {v[1], v[3], v[5], v[7]};
    butterfly<T, 4>(e);
    butterfly<T, 4>(o);
    // o[q] *= exp(-2 pi i q / 8)
    o[1] = mk<T>(h * (o[1].x + o[1].y), h * (o[1].y - o[1].x));
    o[2] = mul_mi<T>(o[2]);
    o[3] = mk<T>(h * (o[3].y - o[3].x), -h * (o[3].x + o[3].y));
#pragma unroll
    for (int q = 0; q < 4; q++) {
      v[q] = add<T>(e[q], o[q]);
      v[q + 4] = sub<T>(e[q], o[q]);
    }
  } else if (R == 3) {
    const T s3 = (T)0.86602540378443864676372317075294;
    cx<T> t1 = add<T>(v[1], v[2]), t2 = sub<T>(v[1], v[2]);
    cx<T> a = v[0];
    v[0] = add<T>(a, t1);
    cx<T> m = mk<T>(a.x - T(0.5) * t1.x, a.y - T(0.5) * t1.y);
    cx<T> r = mk<T>(s3 * t2.y, -s3 * t2.x);  // -i s3 t2
    v[1] = add<T>(m, r);
    v[2] = sub<T>(m, r);
  } else if (R == 5) {
    const T c1 = (T)0.30901699437494742410229341718282;   // cos 2pi/5
    const T c2 = (T)-0.80901699437494742410229341718282;  // cos 4pi/5
    const T s1 = (T)0.95105651629515357211643933337938;   // sin 2pi/5
    const T s2 = (T)0.58778525229247312916870595463907;   // sin 4pi/5
    cx<T> a = v[0];
    cx<T> t1 = add<T>(v[1], v[4]), t2 = add<T>(v[2], v[3]);
    cx<T> u1 = sub<T>(v[1], v[4]), u2 = sub<T>(v[2], v[3]);
    v[0] = add<T>(a, add<T>(t1, t2));
    cx<T> m1 = mk<T>(a.x + c1 * t1.x + c2 * t2.x, a.y + c1 * t1.y + c2 * t2.y);
    cx<T> m2 = mk<T>(a.x + c2 * t1.x + c1 * t2.x, a.y + c2 * t1.y + c1 * t2.y);
    // -i (s1 u1 + s2 u2) and -i (s2 u1 - s1 u2)
    cx<T> n1 = mk<T>(s1 * u1.y + s2 * u2.y, -(s1 * u1.x + s2 * u2.x));
    cx<T> n2 = mk<T>(s2 * u1.y - s1 * u2.y, -(s2 * u1.x - s1 * u2.x));
    v[1] = add<T>(m1, n1);
    v[4] = sub<T>(m1, n1);
    v[2] = add<T>(m2, n2);
    v[3] = sub<T>(m2, n2);
  }
}

#define FFT_EPT 8   // complex elements per thread per pass

// One Stockham pass of radix R over nseq sequences of length M held in LDS.
// COL: element e of sequence s at buf[e*nseq + s] (sequence index fastest);
// else at buf[s*M + e].  tw: table exp(-2 pi i j / (M*twmul)).
template <typename T, int R, bool COL>
__device__ __forceinline__ void pass(cx<T>* buf, int M, int Ns, int nseq,
                                     const cx<T>* __restrict__ tw, int twmul) {
  constexpr int KM = (FFT_EPT + R - 1) / R + 1;
  const int nbf = M / R;
  const int total = nseq * nbf;
  const int span = M / (Ns * R);
  cx<T> v[KM][R];
#pragma unroll
  for (int q = 0; q < KM; q++) {
    int t = threadIdx.x + q * blockDim.x;
    if (t < total) {
      int s, j;
      if (COL) { s = t % nseq; j = t / nseq; }
      else { j = t % nbf; s = t / nbf; }
#pragma unroll
      for (int r = 0; r < R; r++) {
        int e = j + r * nbf;
        v[q][r] = buf[COL ? e * nseq + s : s * M + e];
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < KM; q++) {
    int t = threadIdx.x + q * blockDim.x;
    if (t < total) {
      int s, j;
      if (COL) { s = t % nseq; j = t / nseq; }
      else { j = t % nbf; s = t / nbf; }
      int k = j % Ns;
      if (Ns > 1) {
#pragma unroll
        for (int r = 1; r < R; r++) v[q][r] = mul<T>(v[q][r], tw[(r * k * span) * twmul]);
      }
      butterfly<T, R>(v[q]);
      int base = (j / Ns) * Ns * R + k;
#pragma unroll
      for (int r = 0; r < R; r++) {
        int e = base + r * Ns;
        buf[COL ? e * nseq + s : s * M + e] = v[q][r];
      }
    }
  }
  __syncthreads();
}

template <typename T, bool COL>
__device__ __forceinline__ void fft_lds(cx<T>* buf, int M, int nseq, const FftPlan& pl,
                                        const cx<T>* __restrict__ tw, int twmul) {
  int Ns = 1;
  for (int p = 0; p < pl.npass; p++) {
    switch (pl.radix[p]) {
      case 2: pass<T, 2, COL>(buf, M, Ns, nseq, tw, twmul); break;
      case 3: pass<T, 3, COL>(buf, M, Ns, nseq, tw, twmul); break;
      case 4: pass<T, 4, COL>(buf, M, Ns, nseq, tw, twmul); break;
      case 5: pass<T, 5, COL>(buf, M, Ns, nseq, tw, twmul); break;
      default: pass<T, 8, COL>(buf, M, Ns, nseq, tw, twmul); break;
    }
    Ns *= pl.radix[p];
  }
}

// z pass: nrows_per_wg rows of NZ reals -> NZ/2+1 complex each, in place.
template <typename T>
__global__ __launch_bounds__(1024) void k_fft_rows(T* __restrict__ g0, T* __restrict__ g1,
                                                   T* __restrict__ g2, T* __restrict__ g3,
                                                   int64_t nrows, int nz, int rows_per_wg,
                                                   FftPlan pl, const cx<T>* __restrict__ twz,
                                                   int64_t rowlen) {
  extern __shared__ unsigned char smem[];
  cx<T>* buf = reinterpret_cast<cx<T>*>(smem);
  T* gp = blockIdx.y == 0 ? g0 : blockIdx.y == 1 ? g1 : blockIdx.y == 2 ? g2 : g3;
  const int M = nz / 2;
  const int64_t r0 = (int64_t)blockIdx.x * rows_per_wg;
  const int nr = (int)min((int64_t)rows_per_wg, nrows - r0);
  // load rows as complex pairs (x[2e], x[2e+1])
  for (int t = threadIdx.x; t < rows_per_wg * M; t += blockDim.x) {
    int s = t / M, e = t - s * M;
    if (s < nr) {
      const cx<T>* row = reinterpret_cast<const cx<T>*>(gp + (r0 + s) * rowlen);
      buf[s * M + e] = row[e];
    } else {
      buf[s * M + e] = mk<T>(T(0), T(0));
    }
  }
  __syncthreads();
  fft_lds<T, false>(buf, M, rows_per_wg, pl, twz, 2);
  // even/odd split: X[k] = E[k] + w^k O[k], w = exp(-2 pi i / NZ)
  for (int t = threadIdx.x; t < nr * (M + 1); t += blockDim.x) {
    int s = t / (M + 1), k = t - s * (M + 1);
    cx<T> zk = buf[s * M + (k == M ? 0 : k)];
    cx<T> zr = buf[s * M + (k == 0 ? 0 : M - k)];
    // E = (zk + conj(zr))/2 ; O = (zk - conj(zr)) / (2i)
    cx<T> E = mk<T>(T(0.5) * (zk.x + zr.x), T(0.5) * (zk.y - zr.y));
    cx<T> O = mk<T>(T(0.5) * (zk.y + zr.y), -T(0.5) * (zk.x - zr.x));
    cx<T> X = add<T>(E, mul<T>(twz[k], O));
    cx<T>* row = reinterpret_cast<cx<T>*>(gp + (r0 + s) * rowlen);
    row[k] = X;
  }
}

// y / x pass: transforms of length N along an axis with stride astride
// (complex elements), over columns [c0, c0+B) of plane p (plane stride
// pstride), ncols columns per plane.
template <typename T, int B>
__global__ __launch_bounds__(1024) void k_fft_cols(T* __restrict__ g0, T* __restrict__ g1,
                                                   T* __restrict__ g2, T* __restrict__ g3,
                                                   int N, int64_t astride, int64_t pstride,
                                                   int64_t ncols, int64_t groups_per_plane,
                                                   FftPlan pl, const cx<T>* __restrict__ tw) {
  extern __shared__ unsigned char smem[];
  cx<T>* buf = reinterpret_cast<cx<T>*>(smem);
  T* gp = blockIdx.y == 0 ? g0 : blockIdx.y == 1 ? g1 : blockIdx.y == 2 ? g2 : g3;
  cx<T>* base = reinterpret_cast<cx<T>*>(gp);
  const int64_t plane = blockIdx.x / groups_per_plane;
  const int64_t c0 = (blockIdx.x - plane * groups_per_plane) * B;
  const int nc = (int)min((int64_t)B, ncols - c0);
  cx<T>* col0 = base + plane * pstride + c0;
  for (int t = threadIdx.x; t < N * B; t += blockDim.x) {
    int s = t % B, e = t / B;
    buf[t] = s < nc ? col0[(int64_t)e * astride + s] : mk<T>(T(0), T(0));
  }
  __syncthreads();
  fft_lds<T, true>(buf, N, B, pl, tw, 1);
  for (int t = threadIdx.x; t < N * B; t += blockDim.x) {
    int s = t % B, e = t / B;
    if (s < nc) col0[(int64_t)e * astride + s] = buf[t];
  }
}

// ---------------------------------------------------------------------------
// Register-resident two-level FFT (N = N1 * N2).
//
// x[n1 + N1 n2] -> X[k2 + N2 k1]:
//   1. for each n1: Y[n1][k2] = DFT_N2 over n2        (registers)
//   2. Y[n1][k2] *= exp(-2 pi i n1 k2 / N)           (twiddle table)
//   3. transpose through LDS (one write, one read; rows padded by one
//      element so the strided side is free of bank conflicts)
//   4. for each k2: X[k2 + N2 k1] = DFT_N1 over n1    (registers)
// A thread holds E = lcm(N1, N2) values: E/N2 first-level and E/N1
// second-level transforms.  In-register DFTs are fully unrolled mixed-radix
// (4, 2, 3) decimation in time with constant twiddles from twiddle_table.h.
// ---------------------------------------------------------------------------
constexpr int cgcd(int a, int b) { return b == 0 ? a : cgcd(b, a % b); }
constexpr int clcm(int a, int b) { return a / cgcd(a, b) * b; }

template <typename T>
__device__ __forceinline__ cx<T> twc(int R, int m) {
  // exp(-2 pi i m / R) for compile-time R | 192 (m constant after unrolling)
  m %= R;
  if (m == 0) return mk<T>(T(1), T(0));
  const int idx = m * (TWC_LEN / R);
  return mk<T>((T)TWC_RE[idx], (T)TWC_IM[idx]);
}

template <typename T>
__device__ __forceinline__ cx<T> mul_twc(cx<T> a, int R, int m) {
  m %= R;
  if (m == 0) return a;
  if (4 * m == R) return mul_mi<T>(a);                 // -i
  if (2 * m == R) return mk<T>(-a.x, -a.y);            // -1
  if (4 * m == 3 * R) return mk<T>(-a.y, a.x);         // +i
  return mul<T>(a, twc<T>(R, m));
}

template <typename T, int R>
struct RegDFT {
  static __device__ __forceinline__ void run(const cx<T>* in, cx<T>* out) {
    if constexpr (R == 1) {
      out[0] = in[0];
    } else if constexpr (R == 2 || R == 3 || R == 4 || R == 8) {
      cx<T> v[R];
#pragma unroll
      for (int i = 0; i < R; i++) v[i] = in[i];
      butterfly<T, R>(v);
#pragma unroll
      for (int i = 0; i < R; i++) out[i] = v[i];
    } else {
      constexpr int P = (R % 8 == 0 && R > 16) ? 8 : ((R % 4 == 0) ? 4 : ((R % 2 == 0) ? 2 : 3));
      constexpr int Q = R / P;
      cx<T> so[P][Q];
#pragma unroll
      for (int n1 = 0; n1 < P; n1++) {
        cx<T> si[Q];
#pragma unroll
        for (int n2 = 0; n2 < Q; n2++) si[n2] = in[n1 + P * n2];
        RegDFT<T, Q>::run(si, so[n1]);
      }
#pragma unroll
      for (int k2 = 0; k2 < Q; k2++) {
        cx<T> t[P];
#pragma unroll
        for (int n1 = 0; n1 < P; n1++) t[n1] = mul_twc<T>(so[n1][k2], R, n1 * k2);
        butterfly<T, P>(t);
#pragma unroll
        for (int k1 = 0; k1 < P; k1++) out[k2 + Q * k1] = t[k1];
      }
    }
  }
};

// LDS index of Y[n1][k2] for sequence c (B sequences interleaved).  The k2
// position is rotated by n1 so that the write side (lanes differing in n1)
// spreads over the banks; the read side (lanes differing in k2) stays
// contiguous.
template <int N1, int N2, int B>
__device__ __forceinline__ int tidx(int n1, int k2, int c) {
  return n1 * (N2 * B) + ((k2 + n1) % N2) * B + c;
}

// Columns: B consecutive complex columns, N = N1*N2 along a strided axis.
// Threads: (N / E) per column * B; lane -> (c = lane % B, t = lane / B).
// AX names the axis (1: y, 0: x) so that profiles tell the passes apart;
// AX + 2 is the inverse transform (conjugated input and output: the
// unnormalised sum with e^{+2 pi i nk/N}) used by the adjoint.
template <typename T, int N1, int N2, int B, int AX>
__global__ __launch_bounds__((N1 * N2 / clcm(N1, N2)) * B) void k_fft_cols2(
    T* __restrict__ g0, T* __restrict__ g1, T* __restrict__ g2, T* __restrict__ g3,
    int64_t astride, int64_t pstride, int64_t ncols, int64_t groups_per_plane,
    const cx<T>* __restrict__ tw, int64_t nblocks_x) {
  constexpr int N = N1 * N2;
  constexpr int E = clcm(N1, N2);
  constexpr int TS = N / E;            // threads per sequence
  constexpr int G1 = E / N2, G2 = E / N1;
  constexpr bool INV = AX >= 2;
  extern __shared__ unsigned char smem[];
  cx<T>* lds = reinterpret_cast<cx<T>*>(smem);
  T* gp = blockIdx.y == 0 ? g0 : blockIdx.y == 1 ? g1 : blockIdx.y == 2 ? g2 : g3;
  // XCD-aware block order: consecutive logical blocks on the same XCD
  int64_t bid = blockIdx.x;
  if ((nblocks_x & 7) == 0) bid = (bid & 7) * (nblocks_x >> 3) + (bid >> 3);
  const int64_t plane = bid / groups_per_plane;
  const int64_t c0 = (bid - plane * groups_per_plane) * B;
  const int c = threadIdx.x % B, t = threadIdx.x / B;
  const bool live = c0 + c < ncols;
  cx<T>* col = reinterpret_cast<cx<T>*>(gp) + plane * pstride + c0 + c;
  cx<T> v[E];
  // step 1: load x[n1 + N1 n2], n1 = t + TS g
#pragma unroll
  for (int g = 0; g < G1; g++)
#pragma unroll
    for (int n2 = 0; n2 < N2; n2++) {
      const int n = (t + TS * g) + N1 * n2;
      cx<T> q = live ? col[(int64_t)n * astride] : mk<T>(T(0), T(0));
      if (INV) q.y = -q.y;
      v[g * N2 + n2] = q;
    }
#pragma unroll
  for (int g = 0; g < G1; g++) {
    cx<T> o[N2];
    RegDFT<T, N2>::run(&v[g * N2], o);
    const int n1 = t + TS * g;
#pragma unroll
    for (int k2 = 0; k2 < N2; k2++) {
      cx<T> y = (k2 == 0 || n1 == 0) ? o[k2] : mul<T>(o[k2], tw[n1 * k2]);
      lds[tidx<N1, N2, B>(n1, k2, c)] = y;
    }
  }
  __syncthreads();
  // step 3: k2 = t + TS h, all n1
#pragma unroll
  for (int h = 0; h < G2; h++) {
    const int k2 = t + TS * h;
    cx<T> a[N1], o[N1];
#pragma unroll
    for (int n1 = 0; n1 < N1; n1++) a[n1] = lds[tidx<N1, N2, B>(n1, k2, c)];
    RegDFT<T, N1>::run(a, o);
    if (live) {
#pragma unroll
      for (int k1 = 0; k1 < N1; k1++)
        col[(int64_t)(k2 + N2 * k1) * astride] = INV ? mk<T>(o[k1].x, -o[k1].y) : o[k1];
    }
  }
}

// Rows (z pass): B rows of NZ reals -> NZ/2+1 complex, M = NZ/2 = N1*N2.
// lane -> (t = lane % TS, c = lane / TS) so a wave reads TS consecutive
// complex values of each row.
template <typename T, int N1, int N2, int B>
__global__ __launch_bounds__((N1 * N2 / clcm(N1, N2)) * B) void k_fft_rows2(
    T* __restrict__ g0, T* __restrict__ g1, T* __restrict__ g2, T* __restrict__ g3,
    int64_t nrows, const cx<T>* __restrict__ twz,
    int64_t nblocks_x, int64_t rowlen) {
  constexpr int M = N1 * N2;
  constexpr int E = clcm(N1, N2);
  constexpr int TS = M / E;
  constexpr int G1 = E / N2, G2 = E / N1;
  extern __shared__ unsigned char smem[];
  cx<T>* lds = reinterpret_cast<cx<T>*>(smem);
  T* gp = blockIdx.y == 0 ? g0 : blockIdx.y == 1 ? g1 : blockIdx.y == 2 ? g2 : g3;
  int64_t bid = blockIdx.x;
  if ((nblocks_x & 7) == 0) bid = (bid & 7) * (nblocks_x >> 3) + (bid >> 3);
  const int t = threadIdx.x % TS, c = threadIdx.x / TS;
  const int64_t row = bid * B + c;
  const bool live = row < nrows;
  cx<T>* r = reinterpret_cast<cx<T>*>(gp + (live ? row : 0) * rowlen);
  cx<T> v[E];
#pragma unroll
  for (int g = 0; g < G1; g++)
#pragma unroll
    for (int n2 = 0; n2 < N2; n2++) {
      const int n = (t + TS * g) + N1 * n2;
      v[g * N2 + n2] = live ? r[n] : mk<T>(T(0), T(0));
    }
  // LDS holds B rows of M values (padded rows for the transpose)
  constexpr int RS = N1 * (N2 + 1);   // per-row stride of the transpose buffer
#pragma unroll
  for (int g = 0; g < G1; g++) {
    cx<T> o[N2];
    RegDFT<T, N2>::run(&v[g * N2], o);
    const int n1 = t + TS * g;
#pragma unroll
    for (int k2 = 0; k2 < N2; k2++) {
      cx<T> y = (k2 == 0 || n1 == 0) ? o[k2] : mul<T>(o[k2], twz[2 * n1 * k2]);
      lds[c * RS + n1 * (N2 + 1) + k2] = y;
    }
  }
  __syncthreads();
  cx<T> z[E];
#pragma unroll
  for (int h = 0; h < G2; h++) {
    const int k2 = t + TS * h;
    cx<T> a[N1];
#pragma unroll
    for (int n1 = 0; n1 < N1; n1++) a[n1] = lds[c * RS + n1 * (N2 + 1) + k2];
    RegDFT<T, N1>::run(a, &z[h * N1]);
  }
  __syncthreads();
  // Z[k2 + N2 k1] back to LDS in natural order (row stride M + 1)
  cx<T>* zr = lds + c * (M + 1);
#pragma unroll
  for (int h = 0; h < G2; h++) {
    const int k2 = t + TS * h;
#pragma unroll
    for (int k1 = 0; k1 < N1; k1++) zr[k2 + N2 * k1] = z[h * N1 + k1];
  }
  __syncthreads();
  // even/odd split: X[k] = E[k] + w^k O[k], k = 0..M
  if (live) {
    for (int k = t; k <= M; k += TS) {
      cx<T> zk = zr[k == M ? 0 : k];
      cx<T> zm = zr[k == 0 ? 0 : M - k];
      cx<T> Ev = mk<T>(T(0.5) * (zk.x + zm.x), T(0.5) * (zk.y - zm.y));
      cx<T> Ov = mk<T>(T(0.5) * (zk.y + zm.y), -T(0.5) * (zk.x - zm.x));
      r[k] = add<T>(Ev, mul<T>(twz[k], Ov));
    }
  }
}

// Adjoint of the z pass (r2c rows) for the backward of the binning: for a
// gradient G[k], k = 0..M, on the stored half spectrum of a row, writes the
// real row g[n] = Re sum_{k=0..M} G[k] e^{+2 pi i k n / NZ}, n < NZ = 2M.
// With X[0] = Re G[0], X[M] = Re G[M] and X[k] = G[k]/2 otherwise, g is the
// inverse real transform of the Hermitian spectrum X, computed as one
// complex inverse FFT of length M of F[k] = Fe[k] + i Fo[k],
//   Fe[k] = X[k] + conj(X[M-k]),  Fo[k] = (X[k] - conj(X[M-k])) e^{+2 pi i k/NZ},
// whose real and imaginary parts are g[2m] and g[2m+1]; the inverse FFT is
// the forward register FFT on conjugated input and output.
template <typename T, int N1, int N2, int B>
__global__ __launch_bounds__((N1 * N2 / clcm(N1, N2)) * B) void k_fft_rows2_adj(
    T* __restrict__ g0, T* __restrict__ g1, T* __restrict__ g2, T* __restrict__ g3,
    int64_t nrows, const cx<T>* __restrict__ twz, int64_t nblocks_x, int64_t rowlen) {
  constexpr int M = N1 * N2;
  constexpr int E = clcm(N1, N2);
  constexpr int TS = M / E;
  constexpr int G1 = E / N2, G2 = E / N1;
  constexpr int RS = N1 * (N2 + 1);   // per-row LDS stride (>= M + 1)
  extern __shared__ unsigned char smem[];
  cx<T>* lds = reinterpret_cast<cx<T>*>(smem);
  T* gp = blockIdx.y == 0 ? g0 : blockIdx.y == 1 ? g1 : blockIdx.y == 2 ? g2 : g3;
  int64_t bid = blockIdx.x;
  if ((nblocks_x & 7) == 0) bid = (bid & 7) * (nblocks_x >> 3) + (bid >> 3);
  const int t = threadIdx.x % TS, c = threadIdx.x / TS;
  const int64_t row = bid * B + c;
  const bool live = row < nrows;
  cx<T>* r = reinterpret_cast<cx<T>*>(gp + (live ? row : 0) * rowlen);
  cx<T>* xr = lds + c * RS;
  for (int k = t; k <= M; k += TS) xr[k] = live ? r[k] : mk<T>(T(0), T(0));
  __syncthreads();
  cx<T> v[E];
#pragma unroll
  for (int g = 0; g < G1; g++)
#pragma unroll
    for (int n2 = 0; n2 < N2; n2++) {
      const int n = (t + TS * g) + N1 * n2;      // 0 .. M-1
      cx<T> a = xr[n], b = xr[M - n];
      a = (n == 0) ? mk<T>(a.x, T(0)) : mk<T>(T(0.5) * a.x, T(0.5) * a.y);
      b = (n == 0) ? mk<T>(b.x, T(0)) : mk<T>(T(0.5) * b.x, T(0.5) * b.y);   // M - n == M iff n == 0
      const cx<T> bc = mk<T>(b.x, -b.y);
      const cx<T> fe = add<T>(a, bc);
      const cx<T> w = mk<T>(twz[n].x, -twz[n].y);                           // e^{+2 pi i n/NZ}
      const cx<T> fo = mul<T>(sub<T>(a, bc), w);
      // F = Fe + i Fo, conjugated for the forward FFT
      v[g * N2 + n2] = mk<T>(fe.x - fo.y, -(fe.y + fo.x));
    }
  __syncthreads();
#pragma unroll
  for (int g = 0; g < G1; g++) {
    cx<T> o[N2];
    RegDFT<T, N2>::run(&v[g * N2], o);
    const int n1 = t + TS * g;
#pragma unroll
    for (int k2 = 0; k2 < N2; k2++) {
      cx<T> y = (k2 == 0 || n1 == 0) ? o[k2] : mul<T>(o[k2], twz[2 * n1 * k2]);
      lds[c * RS + n1 * (N2 + 1) + k2] = y;
    }
  }
  __syncthreads();
  cx<T> z[E];
#pragma unroll
  for (int h = 0; h < G2; h++) {
    const int k2 = t + TS * h;
    cx<T> a[N1];
#pragma unroll
    for (int n1 = 0; n1 < N1; n1++) a[n1] = lds[c * RS + n1 * (N2 + 1) + k2];
    RegDFT<T, N1>::run(a, &z[h * N1]);
  }
  __syncthreads();
  cx<T>* zr = lds + c * (M + 1);
#pragma unroll
  for (int h = 0; h < G2; h++) {
    const int k2 = t + TS * h;
#pragma unroll
    for (int k1 = 0; k1 < N1; k1++) zr[k2 + N2 * k1] = z[h * N1 + k1];
  }
  __syncthreads();
  if (live)
    for (int m = t; m < M; m += TS) r[m] = mk<T>(zr[m].x, -zr[m].y);   // g[2m], g[2m+1]
}

// Row pass with a transposed write (the three passes of the out-of-place
// pipeline).  A batch is B rows that are B consecutive values of an "inner"
// index at one "outer" index:
//   input row (outer, inner) at in + outer*in_so + inner*in_sb (complex),
//     contiguous, N = N1*N2 complex values (R2C: 2N reals, an r2c of length
//     2N giving N+1 outputs by the even/odd split);
//   output k of that row at out + outer*out_so + k*out_sk + inner,
// so every write is a run of B consecutive complex values (the transpose of
// the B rows, staged in LDS).  Every read is a whole row.
//
// Persistent: the batches of all grids form one list (batch b of grid g is
// item g*nbx + b); with gridDim.x a multiple of 8, workgroup w runs on XCD
// w % 8 and walks that XCD's contiguous eighth of the list, interleaved with
// the XCD's other workgroups.  The 135 KB LDS of the 1024-point FP32 pass
// leaves one workgroup per CU, so the next batch's rows are loaded (PF)
// while the current batch is transformed and written; and since a thread's
// first-stage index n1 is fixed (G1 = 1), its twiddles w^(n1 k2) are loaded
// once into registers (TWR) instead of once per row.
// First stage of a batch: the N2-point DFTs of the thread's rows, twiddled,
// into the LDS transpose (row c at c*RS).
template <typename T, int N1, int N2, bool R2C, bool TW_REG>
__device__ __forceinline__ void rt_stage1(cx<T>* v, const cx<T>* twr,
                                          const cx<T>* __restrict__ tw, cx<T>* lds, int t,
                                          int c) {
  constexpr int N = N1 * N2, E = clcm(N1, N2), TS = N / E, G1 = E / N2;
  constexpr int RS = N1 * (N2 + 1);
#pragma unroll
  for (int g = 0; g < G1; g++) {
    cx<T> o[N2];
    RegDFT<T, N2>::run(&v[g * N2], o);
    const int n1 = t + TS * g;
#pragma unroll
    for (int k2 = 0; k2 < N2; k2++) {
      // R2C: the 2N-point table, step 2; C2C: the N-point table
      cx<T> y;
      if constexpr (TW_REG) y = k2 == 0 ? o[k2] : mul<T>(o[k2], twr[k2]);
      else y = (k2 == 0 || n1 == 0) ? o[k2] : mul<T>(o[k2], tw[(R2C ? 2 : 1) * n1 * k2]);
      lds[c * RS + n1 * (N2 + 1) + k2] = y;
    }
  }
}

// Second stage after the first stage's barrier: the N1-point DFTs, then row
// c in natural order at lds + c (N + 1) (ends with a barrier).
template <typename T, int N1, int N2>
__device__ __forceinline__ void rt_stage2(cx<T>* lds, int t, int c) {
  constexpr int N = N1 * N2, E = clcm(N1, N2), TS = N / E, G2 = E / N1;
  constexpr int RS = N1 * (N2 + 1);
  cx<T> z[E];
#pragma unroll
  for (int h = 0; h < G2; h++) {
    const int k2 = t + TS * h;
    cx<T> a[N1];
#pragma unroll
    for (int n1 = 0; n1 < N1; n1++) a[n1] = lds[c * RS + n1 * (N2 + 1) + k2];
    RegDFT<T, N1>::run(a, &z[h * N1]);
  }
  __syncthreads();
  cx<T>* zr = lds + c * (N + 1);
#pragma unroll
  for (int h = 0; h < G2; h++) {
    const int k2 = t + TS * h;
#pragma unroll
    for (int k1 = 0; k1 < N1; k1++) zr[k2 + N2 * k1] = z[h * N1 + k1];
  }
  __syncthreads();
}

// The rest of a batch after the first stage's barrier: N1-point DFTs,
// natural order in LDS (row stride N + 1), the r2c split, the transposed
// write.  Ends with the LDS still in use by the write (the caller syncs
// before reusing it).
// tws (R2C): the split twiddles w^k, k = 0..N/2, in LDS (w^(N-k) =
// -conj(w^k)); null: read both from the table in global memory.
template <typename T, int N1, int N2, int B, bool R2C>
__device__ __forceinline__ void rt_rest(const cx<T>* __restrict__ tw, const cx<T>* tws,
                                        cx<T>* lds, int t, int c, cx<T>* ob, int64_t out_sk,
                                        int64_t nlive) {
  constexpr int N = N1 * N2, E = clcm(N1, N2), TS = N / E;
  constexpr int K = R2C ? N + 1 : N;  // outputs per row
  rt_stage2<T, N1, N2>(lds, t, c);
  cx<T>* zr = lds + c * (N + 1);
  if (R2C) {
    // even/odd split X[k] = E[k] + w^k O[k], k = 0..N (w = e^{-2 pi i/2N}):
    // X[k] and X[N-k] need only z[k] and z[N-k] (z[N] = z[0]), so one
    // thread computes the pair and overwrites both in place
    for (int kk = t; kk <= N / 2; kk += TS) {
      const cx<T> zk = zr[kk];
      const cx<T> zm = zr[kk == 0 ? 0 : N - kk];
      const cx<T> Ev = mk<T>(T(0.5) * (zk.x + zm.x), T(0.5) * (zk.y - zm.y));
      const cx<T> Ov = mk<T>(T(0.5) * (zk.y + zm.y), -T(0.5) * (zk.x - zm.x));
      const cx<T> Em = mk<T>(T(0.5) * (zm.x + zk.x), T(0.5) * (zm.y - zk.y));
      const cx<T> Om = mk<T>(T(0.5) * (zm.y + zk.y), -T(0.5) * (zm.x - zk.x));
      cx<T> wk, wm;
      if (tws) {
        wk = tws[kk];
        wm = mk<T>(-wk.x, wk.y);
      } else {
        wk = tw[kk];
        wm = tw[N - kk];
      }
      const cx<T> xk = add<T>(Ev, mul<T>(wk, Ov));
      const cx<T> xm = add<T>(Em, mul<T>(wm, Om));
      zr[kk] = xk;
      zr[N - kk] = xm;   // kk == 0: X[N]; kk == N/2: the same value twice
    }
    __syncthreads();
  }
  // transposed write: runs of B consecutive inner values per output k.
  // FP64 complex-to-complex passes stream it past the caches (non-temporal;
  // measured faster there, slower for FP32 and for the FP64 r2c pass).
  constexpr bool NT = sizeof(T) == 8 && !R2C;
  for (int q = threadIdx.x; q < K * B; q += TS * B) {
    const int b = q % B, kk = q / B;
    if (b < nlive) {
      const cx<T> val = lds[b * (N + 1) + kk];
      if constexpr (NT) {
        typedef double d2 __attribute__((ext_vector_type(2)));
        d2 w;
        w.x = val.x; w.y = val.y;
        __builtin_nontemporal_store(w, reinterpret_cast<d2*>(ob + (int64_t)kk * out_sk + b));
      } else {
        ob[(int64_t)kk * out_sk + b] = val;
      }
    }
  }
}

// PERSIST = false: one batch per workgroup (the same XCD mapping); FP64
// uses it, as the loop form needs about twice the registers there.
// outer_fast: batch b of a grid is (outer = b % n_outer, group = b / n_outer)
// instead of (outer = b / groups, group = b % groups).
template <typename T, int N1, int N2, int B, bool R2C, bool PERSIST, bool PF, bool TWR>
__global__ __launch_bounds__((N1 * N2 / clcm(N1, N2)) * B) void k_fft_rows_t(
    const T* __restrict__ i0, const T* __restrict__ i1, const T* __restrict__ i2,
    const T* __restrict__ i3, T* __restrict__ o0, T* __restrict__ o1, T* __restrict__ o2,
    T* __restrict__ o3, int64_t in_so, int64_t in_sb, int64_t out_so, int64_t out_sk,
    int64_t n_inner, int64_t inner_groups, const cx<T>* __restrict__ tw, int64_t nbx,
    int ngrids, int outer_fast, double* __restrict__ ssq, int ssq_slot0, int64_t ms) {
  constexpr int N = N1 * N2;
  constexpr int E = clcm(N1, N2);
  constexpr int TS = N / E;
  constexpr int G1 = E / N2;
  constexpr bool TW_REG = TWR && G1 == 1;
  extern __shared__ unsigned char smem[];
  // stacked batch: catalogue blockIdx.y, ms reals after the one before it
  // in every buffer, its own ssq rows
  if (blockIdx.y) {
    const int64_t mo = (int64_t)blockIdx.y * ms;
    i0 += mo; i1 += mo; i2 += mo; i3 += mo;
    o0 += mo; o1 += mo; o2 += mo; o3 += mo;
    if (ssq) ssq += (int64_t)blockIdx.y * gridDim.x * 4;
  }
  cx<T>* lds = reinterpret_cast<cx<T>*>(smem);
  const int t = threadIdx.x % TS, c = threadIdx.x / TS;
  const int64_t total = nbx * ngrids;
  const int64_t n_outer = nbx / inner_groups;
  int64_t L0, Lend, step;
  if ((gridDim.x & 7) == 0) {
    const int64_t M = gridDim.x >> 3;
    const int64_t chunk = (total + 7) >> 3;
    const int64_t xcd = blockIdx.x & 7;
    L0 = xcd * chunk + (blockIdx.x >> 3);
    Lend = min(total, (xcd + 1) * chunk);
    step = M;
  } else {
    L0 = blockIdx.x;
    Lend = total;
    step = gridDim.x;
  }
  cx<T> twr[TW_REG ? N2 : 1];
  if constexpr (TW_REG) {
#pragma unroll
    for (int k2 = 0; k2 < N2; k2++) twr[k2] = tw[(R2C ? 2 : 1) * t * k2];
  }
  // persistent r2c: the split twiddles once per workgroup, after the
  // transpose area (the first batch's barrier orders the fill)
  cx<T>* tws = nullptr;
  if constexpr (R2C && PERSIST) {
    tws = lds + B * (N + N1);
    for (int k = threadIdx.x; k <= N / 2; k += TS * B) tws[k] = tw[k];
  }
  cx<T> v[E];
  auto load = [&](int64_t item) {
    const int gsel = (int)(item / nbx);
    const int64_t bid = item - (int64_t)gsel * nbx;
    const T* gin = gsel == 0 ? i0 : gsel == 1 ? i1 : gsel == 2 ? i2 : i3;
    const int64_t ig = outer_fast ? bid / n_outer : bid - (bid / inner_groups) * inner_groups;
    const int64_t outer = outer_fast ? bid - ig * n_outer : bid / inner_groups;
    const int64_t inner0 = ig * B;
    const bool live = inner0 + c < n_inner;
    const cx<T>* row = reinterpret_cast<const cx<T>*>(gin) + outer * in_so +
                       (live ? inner0 + c : 0) * in_sb;
#pragma unroll
    for (int g = 0; g < G1; g++)
#pragma unroll
      for (int n2 = 0; n2 < N2; n2++) {
        const int n = (t + TS * g) + N1 * n2;
        v[g * N2 + n2] = live ? row[n] : mk<T>(T(0), T(0));
      }
  };
  // output of a batch: the start of its transposed runs, its live rows
  auto dest = [&](int64_t item, int64_t& nlive) {
    const int gsel = (int)(item / nbx);
    const int64_t bid = item - (int64_t)gsel * nbx;
    T* gout = gsel == 0 ? o0 : gsel == 1 ? o1 : gsel == 2 ? o2 : o3;
    const int64_t ig = outer_fast ? bid / n_outer : bid - (bid / inner_groups) * inner_groups;
    const int64_t outer = outer_fast ? bid - ig * n_outer : bid / inner_groups;
    nlive = n_inner - ig * B;
    return reinterpret_cast<cx<T>*>(gout) + outer * out_so + ig * B;
  };
  if constexpr (!PERSIST) {
    if (L0 >= Lend) return;
    load(L0);
    rt_stage1<T, N1, N2, R2C, TW_REG>(v, twr, tw, lds, t, c);
    __syncthreads();
    int64_t nlive;
    cx<T>* ob = dest(L0, nlive);
    rt_rest<T, N1, N2, B, R2C>(tw, nullptr, lds, t, c, ob, out_sk, nlive);
  } else {
    // ssq: sum of |v|^2 of the rows this workgroup reads, per grid (fixed
    // order: per thread in item order, then the workgroup's threads in order)
    double sq[4] = {0.0, 0.0, 0.0, 0.0};
    if (L0 < Lend) load(L0);
    for (int64_t item = L0; item < Lend; item += step) {
      if (!PF && item != L0) load(item);
      if (ssq) {
        float s2 = 0.0f;
#pragma unroll
        for (int e = 0; e < E; e++) s2 += (float)v[e].x * (float)v[e].x + (float)v[e].y * (float)v[e].y;
        const int gsel = (int)(item / nbx);
#pragma unroll
        for (int q = 0; q < 4; q++)
          if (q == gsel) sq[q] += (double)s2;
      }
      rt_stage1<T, N1, N2, R2C, TW_REG>(v, twr, tw, lds, t, c);
      __syncthreads();
      // v is free: the next batch's loads overlap the rest of this one
      if (PF && item + step < Lend) load(item + step);
      int64_t nlive;
      cx<T>* ob = dest(item, nlive);
      rt_rest<T, N1, N2, B, R2C>(tw, tws, lds, t, c, ob, out_sk, nlive);
      // the next batch's first stage overwrites the LDS
      __syncthreads();
    }
    if (ssq) {
      // the LDS holds B (N + N1) >= 32 TS B bytes: room for 4 doubles a thread
      double* red = reinterpret_cast<double*>(smem);
#pragma unroll
      for (int q = 0; q < 4; q++) red[threadIdx.x * 4 + q] = sq[q];
      __syncthreads();
      // the slots of this launch's grids only (grid-by-grid launches fill
      // the others)
      if ((int)threadIdx.x < ngrids) {
        double tot = 0.0;
        for (int i = 0; i < (int)blockDim.x; i++) tot += red[i * 4 + threadIdx.x];
        ssq[(int64_t)blockIdx.x * 4 + ssq_slot0 + threadIdx.x] = tot;
      }
    }
  }
}

// Sizes served by the register kernels: N -> (N1, N2).  2048 (64 x 32, 64
// complex values a thread) in FP32 only: FP64 would hold 256 VGPRs of data
// alone, so its 2048-point passes take the Stockham kernels.
static bool reg_split(int N, int* n1, int* n2, bool fp64 = false) {
  switch (N) {
    case 2048:
      if (fp64) return false;
      *n1 = 64; *n2 = 32; return true;
    case 16: *n1 = 4; *n2 = 4; return true;
    case 32: *n1 = 4; *n2 = 8; return true;
    case 64: *n1 = 8; *n2 = 8; return true;
    case 96: *n1 = 8; *n2 = 12; return true;
    case 128: *n1 = 8; *n2 = 16; return true;
    case 192: *n1 = 8; *n2 = 24; return true;
    case 256: *n1 = 16; *n2 = 16; return true;
    case 384: *n1 = 16; *n2 = 24; return true;
    case 512: *n1 = 16; *n2 = 32; return true;
    case 1024: *n1 = 32; *n2 = 32; return true;
  }
  return false;
}

template <typename T, int B>
static void launch_cols(FftArgs& a, const char* name, T** gp, int N, int64_t astride,
                        int64_t pstride, int64_t ncols, int64_t nplanes, const FftPlan& pl,
                        const void* tw) {
  int threads = (N * B + FFT_EPT - 1) / FFT_EPT;
  threads = min(max(((threads + 63) / 64) * 64, 64), 1024);
  int64_t gpp = (ncols + B - 1) / B;
  dim3 grid((unsigned)(gpp * nplanes), a.ngrids);
  size_t lds = (size_t)N * B * sizeof(cx<T>);
  TIMED(a.timer, name,
        (k_fft_cols<T, B><<<grid, threads, lds, a.stream>>>(gp[0], gp[1], gp[2], gp[3], N, astride,
                                                            pstride, ncols, gpp, pl,
                                                            (const cx<T>*)tw)));
}

// widest column batch B (contiguous bytes per axis index) that fits
template <typename T>
static void launch_cols_any(FftArgs& a, const char* name, T** gp, int N, int64_t astride,
                            int64_t pstride, int64_t ncols, int64_t nplanes, const FftPlan& pl,
                            const void* tw, int64_t lim) {
  const int bmax = sizeof(T) == 4 ? 8 : 4;
  if (bmax >= 8 && (int64_t)N * 8 <= lim)
    launch_cols<T, 8>(a, name, gp, N, astride, pstride, ncols, nplanes, pl, tw);
  else if ((int64_t)N * 4 <= lim)
    launch_cols<T, 4>(a, name, gp, N, astride, pstride, ncols, nplanes, pl, tw);
  else
    launch_cols<T, 2>(a, name, gp, N, astride, pstride, ncols, nplanes, pl, tw);
}
#define TIMED_COLS(a, name, gp, N, astride, pstride, ncols, nplanes, pl, tw, lim) \
  launch_cols_any<T>(a, name, gp, N, astride, pstride, ncols, nplanes, pl, tw, lim)

// Columns per workgroup for the register kernels: at least one wavefront
// of threads, 64 contiguous bytes per axis index when possible, <= 64 KB LDS.
template <typename T, int N1, int N2>
constexpr int cols_B() {
  constexpr int N = N1 * N2, E = clcm(N1, N2), TS = N / E;
  int b = sizeof(T) == 4 ? 8 : 4;
  while (TS * b < 64) b *= 2;
  while (b > 1 && (size_t)N * b * sizeof(cx<T>) > 64 * 1024) b /= 2;
  return b;
}

template <typename T, int N1, int N2>
constexpr int rows_B() {
  constexpr int M = N1 * N2, E = clcm(N1, N2), TS = M / E;
  int b = 1;
  while (TS * b < 256 && (size_t)(2 * b) * (M + N1) * sizeof(cx<T>) <= 40 * 1024) b *= 2;
  return b;
}

template <typename T, int N1, int N2, int MUL, int AX>
static void launch_cols2(FftArgs& a, const char* name, T** gp, int64_t astride, int64_t pstride,
                         int64_t ncols, int64_t nplanes, const void* tw) {
  constexpr int N = N1 * N2, E = clcm(N1, N2), TS = N / E;
  constexpr int B = cols_B<T, N1, N2>() * MUL;
  const size_t lds = (size_t)N * B * sizeof(cx<T>);
  if (lds > 64 * 1024)
    hipFuncSetAttribute((const void*)k_fft_cols2<T, N1, N2, B, AX>,
                        hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  const int64_t gpp = (ncols + B - 1) / B;
  const int64_t nbx = gpp * nplanes;
  dim3 grid((unsigned)nbx, a.ngrids);
  TIMED(a.timer, name,
        (k_fft_cols2<T, N1, N2, B, AX><<<grid, TS * B, lds, a.stream>>>(
            gp[0], gp[1], gp[2], gp[3], astride, pstride, ncols, gpp, (const cx<T>*)tw, nbx)));
}

template <typename T, int N1, int N2>
static void launch_rows2(FftArgs& a, T** gp, int64_t nrows, const void* twz, int64_t rowlen) {
  constexpr int M = N1 * N2, E = clcm(N1, N2), TS = M / E;
  constexpr int B = rows_B<T, N1, N2>();
  const size_t lds = (size_t)B * (M + N1) * sizeof(cx<T>);
  const int64_t nbx = (nrows + B - 1) / B;
  dim3 grid((unsigned)nbx, a.ngrids);
  TIMED(a.timer, "fft_z",
        (k_fft_rows2<T, N1, N2, B><<<grid, TS * B, lds, a.stream>>>(
            gp[0], gp[1], gp[2], gp[3], nrows, (const cx<T>*)twz, nbx, rowlen)));
}

#define CAPOW_FFT_SIZES(X) X(4, 4) X(4, 8) X(8, 8) X(8, 12) X(8, 16) X(8, 24) X(16, 16) \
  X(16, 24) X(16, 32) X(32, 32) X(64, 32)
// split (A, B_) instantiated for precision T (64 x 32: FP32 only)
template <typename T, int A, int B_>
constexpr bool split_for() { return sizeof(T) == 4 || A * B_ != 2048; }

template <typename T, int AX>
static bool try_cols2(FftArgs& a, const char* name, T** gp, int N, int64_t astride,
                      int64_t pstride, int64_t ncols, int64_t nplanes, const void* tw) {
  int n1, n2;
  if (!reg_split(N, &n1, &n2, sizeof(T) == 8)) return false;
#define X(A, B_)                                                                     \
  if constexpr (split_for<T, A, B_>())                                                \
  if (n1 == A && n2 == B_) {                                                          \
    if (a.col_mul == 2 && (size_t)A * B_ * cols_B<T, A, B_>() * 2 * sizeof(cx<T>) <= 160 * 1024) \
      launch_cols2<T, A, B_, 2, AX>(a, name, gp, astride, pstride, ncols, nplanes, tw); \
    else                                                                              \
      launch_cols2<T, A, B_, 1, AX>(a, name, gp, astride, pstride, ncols, nplanes, tw); \
    return true;                                                                      \
  }
  CAPOW_FFT_SIZES(X)
#undef X
  return false;
}

template <typename T>
static bool try_rows2(FftArgs& a, T** gp, int M, int64_t nrows, const void* twz, int64_t rowlen) {
  int n1, n2;
  if (!reg_split(M, &n1, &n2, sizeof(T) == 8)) return false;
#define X(A, B_)                                   \
  if constexpr (split_for<T, A, B_>())              \
  if (n1 == A && n2 == B_) {                        \
    launch_rows2<T, A, B_>(a, gp, nrows, twz, rowlen); \
    return true;                                    \
  }
  CAPOW_FFT_SIZES(X)
#undef X
  return false;
}

template <typename T, int N1, int N2, bool R2C, int MUL>
static void launch_rows_t(FftArgs& a, const char* name, T* const* in, T* const* out, int64_t n_outer,
                          int64_t n_inner, int64_t in_so, int64_t in_sb, int64_t out_so,
                          int64_t out_sk, const void* tw, double* ssq) {
  constexpr int N = N1 * N2, E = clcm(N1, N2), TS = N / E;
  constexpr int B = (sizeof(T) == 4 ? 8 : 4) * MUL;   // 64- (MUL 2: 128-) byte write runs
  // FP32: persistent, prefetch and register twiddles (FP64 has no registers
  // to spare: one batch per workgroup)
  constexpr bool PS = sizeof(T) == 4, PF = sizeof(T) == 4, TWR = sizeof(T) == 4;
  const void* fn = (const void*)k_fft_rows_t<T, N1, N2, B, R2C, PS, PF, TWR>;
  const size_t lds = ((size_t)B * (N + N1) + (R2C && PS ? N / 2 + 1 : 0)) * sizeof(cx<T>);
  if (lds > 64 * 1024)
    hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  const int64_t groups = (n_inner + B - 1) / B;
  const int64_t nbx = n_outer * groups;
  const int64_t total = nbx * a.ngrids;
  int per_cu = 1;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, TS * B, lds) != hipSuccess ||
      per_cu < 1)
    per_cu = 1;
  int64_t nblk = (int64_t)(a.ncu > 0 ? a.ncu : 256) * per_cu;
  if (!PS || a.persist == 0 || nblk > total) nblk = total;
  if (ssq) a.ssq_rows = (int)nblk;
  const int ofast = (a.order >> (name[4] == 'z' ? 0 : name[4] == 'y' ? 1 : 2)) & 1;
  const dim3 grid((unsigned)nblk, (unsigned)a.nmock);
  TIMED(a.timer, name,
        (k_fft_rows_t<T, N1, N2, B, R2C, PS, PF, TWR><<<grid, TS * B, lds, a.stream>>>(
            in[0], in[1], in[2], in[3], out[0], out[1], out[2], out[3], in_so, in_sb, out_so,
            out_sk, n_inner, groups, (const cx<T>*)tw, nbx, a.ngrids, ofast, ssq, a.ssq_slot0,
            a.mock_stride)));
}

template <typename T, bool R2C>
static bool try_rows_t(FftArgs& a, const char* name, int N, T* const* in, T* const* out,
                       int64_t n_outer, int64_t n_inner, int64_t in_so, int64_t in_sb,
                       int64_t out_so, int64_t out_sk, const void* tw, int mul,
                       double* ssq = nullptr) {
  int n1, n2;
  if (!reg_split(N, &n1, &n2, sizeof(T) == 8)) return false;
#define X(A, B_)                                                                               \
  if constexpr (split_for<T, A, B_>())                                                          \
  if (n1 == A && n2 == B_) {                                                                    \
    if (mul == 2 && (size_t)(sizeof(T) == 4 ? 16 : 8) * sizeof(cx<T>) * (A * B_ + A) +           \
                        (R2C ? (size_t)(A * B_ / 2 + 1) * sizeof(cx<T>) : 0) <= 160 * 1024)      \
      launch_rows_t<T, A, B_, R2C, 2>(a, name, in, out, n_outer, n_inner, in_so, in_sb, out_so,  \
                                      out_sk, tw, ssq);                                        \
    else                                                                                       \
      launch_rows_t<T, A, B_, R2C, 1>(a, name, in, out, n_outer, n_inner, in_so, in_sb, out_so,  \
                                      out_sk, tw, ssq);                                        \
    return true;                                                                               \
  }
  CAPOW_FFT_SIZES(X)
#undef X
  return false;
}

// ---------------------------------------------------------------------------
// The x pass fused with the mirror-group binning (FP32, 1-D bins).
//
// After the z and y passes the grids hold u = [ky][kz][x] (x rows
// contiguous, pitch px).  A work item is the x rows (j, kz) and (NY - j, kz)
// of every grid for KZ consecutive kz (B = 2 G KZ rows, G grids): they are
// transformed along x as the x pass does (two-level register FFT, natural
// order in LDS), and the mirror groups (i, j), (NX - i, j), (i, NY - j),
// (NX - i, NY - j) of those kz -- all in LDS -- are binned at once, so the
// x pass output is never written.  The sums are exact int64 quanta, one
// quantum per spectrum, added into ONE table per workgroup (integer adds do
// not depend on their order: the bins are bit-reproducible).  With S_A the
// sum of |A(k)|^2 + |A_int(k)|^2 over the stored modes (NX NY times the sum
// of |v|^2 over the y pass input of those grids: Parseval along y and x),
// S_B the same for the B grids, and |delta|^2 <= 2 (|A|^2 + |A_int|^2):
//   auto A:  sum |NM cor L_l |delta|^2|       <= 4 cor_max S_A
//   auto B:  sum |NM cor L_l |beta|^2|        <= 4 cor_max S_B
//   cross:   sum |NM cor L_l Re/Im delta* beta| <= 2 cor_max sum |delta||beta|
//                                              <= 4 cor_max sqrt(S_A S_B)
// (Cauchy-Schwarz), so with that bound T_s < 2^e and q_s = 2^(e - 62) no
// partial sum of any bin leaves int64 (each rounded add moves it by at most
// 1/2), and q_s is about 2^-62 T_s: a spectrum's quantum follows its own
// power (the density spectrum is not rounded on the momentum's scale).
// ---------------------------------------------------------------------------

// ssq rows (4 per workgroup of the y pass, one per transformed grid in the
// order A, A_int, B, B_int of the grids present) -> xq[2 s], xq[2 s + 1] =
// q_s, 1/q_s per spectrum s, fixed order.  factor = 4 cor_max NX NY.
// Stacked batch: catalogue blockIdx.x, its ssq rows and quanta after those
// of the catalogues before it.
__global__ __launch_bounds__(256) void k_xbin_scale(const double* __restrict__ ssq, int rows,
                                                    double factor, int il, int hb, BinArgsDev a,
                                                    double* __restrict__ xq) {
  ssq += (int64_t)blockIdx.x * rows * 4;
  xq += (int64_t)blockIdx.x * 6;
  // each thread sums a strided subset of the rows, then a fixed tree
  __shared__ double red[4][256];
  double part[4] = {0.0, 0.0, 0.0, 0.0};
  for (int r = threadIdx.x; r < rows; r += 256)
#pragma unroll
    for (int q = 0; q < 4; q++) part[q] += ssq[r * 4 + q];
#pragma unroll
  for (int q = 0; q < 4; q++) red[q][threadIdx.x] = part[q];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
#pragma unroll
      for (int q = 0; q < 4; q++) red[q][threadIdx.x] += red[q][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x != 0) return;
  const double g[4] = {red[0][0], red[1][0], red[2][0], red[3][0]};
  const double sa = g[0] + (il ? g[1] : 0.0);
  const double sb = hb ? (il ? g[2] + g[3] : g[1]) : 0.0;
  for (int s = 0; s < 3; s++) {
    double T = 0.0;
    if (s < a.nspec) {
      const int k = a.kind[s];
      T = factor * (k == SPEC_AUTO_A ? sa : (k == SPEC_AUTO_B ? sb : sqrt(sa * sb)));
      T *= 1.0 + 1e-5;   // the float partial sums of |v|^2 round by < 1e-5
    }
    int e = 0;
    if (T > 0) frexp(T, &e);   // T < 2^e
    xq[2 * s] = T > 0 ? ldexp(1.0, e - 62) : 1.0;
    xq[2 * s + 1] = T > 0 ? ldexp(1.0, 62 - e) : 1.0;
  }
}

template <int N1, int N2, int G, int KZ, bool IL, bool HB, bool ODD>
__global__ __launch_bounds__((N1 * N2 / clcm(N1, N2)) * (2 * G * KZ)) void k_fft_x_bin(
    const float* __restrict__ i0, const float* __restrict__ i1, const float* __restrict__ i2,
    const float* __restrict__ i3, int64_t px, int64_t plane, const cx<float>* __restrict__ tw,
    const double* __restrict__ tabs, BinArgsDev a, const double* __restrict__ xq,
    long long* __restrict__ xpart, int nkzb, int64_t ms) {
  constexpr int N = N1 * N2, E = clcm(N1, N2), TS = N / E, G1 = E / N2;
  constexpr int B = 2 * G * KZ;
  constexpr bool TW_REG = G1 == 1;
  extern __shared__ unsigned char smem[];
  // stacked batch: catalogue blockIdx.y (grids ms reals apart, its own
  // quanta and table rows)
  if (blockIdx.y) {
    const int64_t mo = (int64_t)blockIdx.y * ms;
    i0 += mo; i1 += mo; i2 += mo; i3 += mo;
    xq += (int64_t)blockIdx.y * 6;
  }
  cx<float>* lds = reinterpret_cast<cx<float>*>(smem);
  unsigned long long* tab = reinterpret_cast<unsigned long long*>(smem + (size_t)B * (N + N1) * 8);
  constexpr int NL = ODD ? 5 : 3;   // multipoles summed: 0, 2, 4 [, 1, 3]
  const int NK = a.nk, nl = NL, F = a.nspec * nl;
  for (int q = threadIdx.x; q < F * NK; q += blockDim.x) tab[q] = 0ull;
  const int t = threadIdx.x % TS, c = threadIdx.x / TS;
  // row c of an item: grid c % G, j slot (c / G) % 2 (0: j, 1: NY - j),
  // kz slot c / (2 G)
  const int cg = c % G, cjs = (c / G) & 1, cks = c / (2 * G);
  const int hy1 = a.ny / 2 + 1;
  const int64_t total = (int64_t)hy1 * nkzb;
  int64_t L0, Lend, step;
  if ((gridDim.x & 7) == 0) {
    const int64_t M = gridDim.x >> 3;
    const int64_t chunk = (total + 7) >> 3;
    const int64_t xcd = blockIdx.x & 7;
    L0 = xcd * chunk + (blockIdx.x >> 3);
    Lend = min(total, (xcd + 1) * chunk);
    step = M;
  } else {
    L0 = blockIdx.x;
    Lend = total;
    step = gridDim.x;
  }
  cx<float> twr[TW_REG ? N2 : 1];
  if constexpr (TW_REG) {
#pragma unroll
    for (int k2 = 0; k2 < N2; k2++) twr[k2] = tw[t * k2];
  }
  const float* gin = cg == 0 ? i0 : cg == 1 ? i1 : cg == 2 ? i2 : i3;
  cx<float> v[E];
  auto load = [&](int64_t item) {
    const int jc = (int)(item / nkzb), kz = (int)(item - (int64_t)jc * nkzb) * KZ + cks;
    const bool live = kz < a.nzh && (cjs == 0 || (jc > 0 && 2 * jc < a.ny));
    const int j = cjs ? a.ny - jc : jc;
    const cx<float>* row = reinterpret_cast<const cx<float>*>(gin) +
                           (live ? (int64_t)j * plane + (int64_t)kz * px : 0);
#pragma unroll
    for (int g = 0; g < G1; g++)
#pragma unroll
      for (int n2 = 0; n2 < N2; n2++) {
        const int n = (t + TS * g) + N1 * n2;
        v[g * N2 + n2] = live ? row[n] : mk<float>(0.f, 0.f);
      }
  };
  // binning split: the item's KZ x HX groups (kz slot, canonical
  // i = 0 .. NX/2 - 1) numbered q = slot HX + i, thread t taking the S
  // consecutive q from t S.  An even S visits them with the lowest bit
  // flipped in every other run of 8 lanes, so that the lanes' 8-byte LDS
  // reads (S complex apart) spread over the banks.  The x Nyquist groups
  // (i = NX/2, one per kz slot) lie above every Nyquist frequency except at
  // ky = kz = 0; threads 0 .. KZ-1 take them after an exact check of the bin
  // condition, so they cost no round of their own.
  static_assert((TS * B) % 64 == 0, "whole waves");
  static_assert(N % 2 == 0, "even x extent: one x Nyquist group per kz");
  constexpr int HX = N / 2;
  constexpr int GI = KZ * HX;
  constexpr int S = (GI + TS * B - 1) / (TS * B);
  const int lane = threadIdx.x & 63;
  const int uflip = (S % 2) ? 0 : (((lane >> 3) ^ (lane >> 4)) & 1);
  const double iqs[3] = {xq[1], xq[3], xq[5]};
  const typename CT<float>::t zero = {0.f, 0.f};
  if (L0 < Lend) load(L0);
  for (int64_t item = L0; item < Lend; item += step) {
    rt_stage1<float, N1, N2, false, TW_REG>(v, twr, tw, lds, t, c);
    __syncthreads();
    if (item + step < Lend) load(item + step);
    rt_stage2<float, N1, N2>(lds, t, c);
    // bin the item's groups from LDS.  The item's y quantities (rows j and
    // NY - j) are the same for every group; a thread takes S consecutive
    // groups, whose modes mostly share a k bin, and sums them in FP64
    // registers until the bin changes (one conversion to quanta and one set
    // of LDS atomics per run).
    const int jc = (int)(item / nkzb);
    const int kz0 = (int)(item - (int64_t)jc * nkzb) * KZ;
    const bool my = jc > 0 && 2 * jc < a.ny;
    const int j2 = my ? a.ny - jc : jc;
    const double fy = tabs[a.ofy + jc], cory = tabs[a.ocy + jc];
    const double pyr = tabs[a.opy + 2 * jc], pyi = tabs[a.opy + 2 * jc + 1];
    const double pyr2 = tabs[a.opy + 2 * j2], pyi2 = tabs[a.opy + 2 * j2 + 1];
    int cur = -1;
    double acc[3][NL];
#pragma unroll
    for (int sp = 0; sp < 3; sp++)
#pragma unroll
      for (int q = 0; q < NL; q++) acc[sp][q] = 0.0;
    auto flush = [&]() {
      if (cur < 0) return;
#pragma unroll
      for (int sp = 0; sp < 3; sp++) {
        if (sp >= a.nspec) break;
        unsigned long long* b = tab + sp * nl * NK + cur;
#pragma unroll
        for (int q = 0; q < NL; q++) {
          atomicAdd(&b[q * NK], (unsigned long long)__double2ll_rn(acc[sp][q] * iqs[sp]));
          acc[sp][q] = 0.0;
        }
      }
    };
    // group (i, jc) of kz slot wks (kz = kz0 + wks < NZ/2 + 1)
    auto bin_group = [&](int wks, int i) {
      const int kz = kz0 + wks;
      const double fz = tabs[a.ofz + kz], corz = tabs[a.ocz + kz];
      const double pzr = tabs[a.opz + 2 * kz], pzi = tabs[a.opz + 2 * kz + 1];
      const bool mx = i > 0 && 2 * i < a.nx;
      const int i2 = mx ? a.nx - i : i;
      Group Gp;
      Gp.i = i;
      Gp.j = jc;
      Gp.fx = tabs[a.ofx + i];
      Gp.fy = fy;
      Gp.fxy = dadd(dmul(Gp.fx, Gp.fx), dmul(fy, fy));
      Gp.corxy = tabs[a.ocx + i] * cory;
      Gp.live[0] = true;
      Gp.live[1] = mx;
      Gp.live[2] = my;
      Gp.live[3] = mx && my;
#pragma unroll
      for (int r = 0; r < 4; r++) {
        Gp.qr[r] = 1.0;
        Gp.qi[r] = 0.0;
      }
      if constexpr (IL) {
        const double pxr = tabs[a.opx + 2 * i], pxi = tabs[a.opx + 2 * i + 1];
        const double pxr2 = tabs[a.opx + 2 * i2], pxi2 = tabs[a.opx + 2 * i2 + 1];
        const double xr[4] = {pxr, pxr2, pxr, pxr2}, xi[4] = {pxi, pxi2, pxi, pxi2};
        const double yr[4] = {pyr, pyr, pyr2, pyr2}, yi[4] = {pyi, pyi, pyi2, pyi2};
#pragma unroll
        for (int r = 0; r < 4; r++) {
          Gp.qr[r] = xr[r] * yr[r] - xi[r] * yi[r];
          Gp.qi[r] = xr[r] * yi[r] + xi[r] * yr[r];
        }
      }
      typename CT<float>::t vv[4][4];
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int js = r >> 1, ii = (r & 1) ? i2 : i;
        const bool lv = Gp.live[r];
#pragma unroll
        for (int q = 0; q < 4; q++) {
          // grid slot q (A, A_int, B, B_int) -> loaded grid index
          const int gi = q == 0 ? 0 : q == 1 ? (IL ? 1 : -1) : q == 2 ? (HB ? (IL ? 2 : 1) : -1)
                                                                      : (IL && HB ? 3 : -1);
          vv[r][q] = (lv && gi >= 0) ? lds[((wks * 2 + js) * G + gi) * (N + 1) + ii] : zero;
        }
      }
      ModeVals mv =
          group_mode<typename CT<float>::t, IL, HB, true>(Gp, kz, fz, corz, pzr, pzi, vv, a);
      if (mv.near) {
        const KBin e = kbin_exact(Gp.fxy, fz, a);
        mv.kbin = e.kbin;
        mv.ok = e.ok;
      }
      if (!mv.ok) return;
      if (mv.kbin != cur) {
        flush();
        cur = mv.kbin;
      }
#pragma unroll
      for (int sp = 0; sp < 3; sp++) {
        const double np = mv.np[sp];
        acc[sp][0] += np;
        acc[sp][1] += mv.L2 * np;
        acc[sp][2] += mv.L4 * np;
        if constexpr (ODD) {
          acc[sp][3] += mv.mu * np;
          acc[sp][4] += (2.5 * mv.mu * mv.mu * mv.mu - 1.5 * mv.mu) * np;
        }
      }
    };
#pragma unroll
    for (int u = 0; u < S; u++) {
      const int q = (int)threadIdx.x * S + (u ^ uflip);
      if (q >= GI) continue;
      const int wks = q / HX;
      if (kz0 + wks >= a.nzh) continue;
      bin_group(wks, q - wks * HX);
    }
    // the x Nyquist group of kz slot t: binned only where its |k| passes the
    // bin condition (the same arithmetic as the group's own)
    if ((int)threadIdx.x < KZ && kz0 + (int)threadIdx.x < a.nzh) {
      const double fxn = tabs[a.ofx + HX];
      const KBin kb = kbin_any(dadd(dmul(fxn, fxn), dmul(fy, fy)), tabs[a.ofz + kz0 + threadIdx.x], a);
      if (kb.ok || kb.near) bin_group((int)threadIdx.x, HX);
    }
    flush();
    // the next item's first stage overwrites the rows
    __syncthreads();
  }
  long long* out = xpart + ((int64_t)blockIdx.y * gridDim.x + blockIdx.x) * F * NK;
  for (int q = threadIdx.x; q < F * NK; q += blockDim.x) out[q] = (long long)tab[q];
}

// Workgroup tables summed exactly (integer sums: any order gives the same
// result), times their spectrum's quantum: the sums of L.out ([s][l][k],
// nlk = nl NK per spectrum), then the mode counts (the k_bin_final layout of
// the branch-free forms).  One block per entry; stacked batch: catalogue
// blockIdx.y, its outputs out_stride doubles after the one before it.
__global__ __launch_bounds__(256) void k_xbin_final(const long long* __restrict__ xpart, int nwg,
                                                    int fnk, int nk, int nlk,
                                                    const double* __restrict__ counts,
                                                    const double* __restrict__ xq,
                                                    double* __restrict__ out, int64_t out_stride) {
  const int i = blockIdx.x;
  xpart += (int64_t)blockIdx.y * nwg * fnk;
  xq += (int64_t)blockIdx.y * 6;
  out += (int64_t)blockIdx.y * out_stride;
  if (i >= fnk) {
    if (threadIdx.x == 0 && i < fnk + nk) out[i] = counts[i - fnk];
    return;
  }
  __shared__ long long red[256];
  long long t = 0;
  for (int w = threadIdx.x; w < nwg; w += 256) t += xpart[(int64_t)w * fnk + i];
  red[threadIdx.x] = t;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[i] = (double)red[0] * xq[2 * (i / nlk)];
}

// rows per item for G grids: 16 (KZ = 8/G) when the rows and the table fit
// the LDS beside each other, else 8
static int xbin_rows(int N, int N1, const BinArgsDev& b, size_t* lds) {
  const int F = b.nspec * (b.odd ? 5 : 3);
  const size_t tabb = (size_t)F * b.nk * 8;
  for (int B = 16; B >= 8; B /= 2) {
    const size_t l = (size_t)B * (N + N1) * 8 + tabb;
    if (l <= 160 * 1024) {
      *lds = l;
      return B;
    }
  }
  *lds = 0;
  return 0;
}

size_t xbin_lds(const Geom& g, const BinArgsDev& b) {
  int n1, n2;
  if (!reg_split(g.n[0], &n1, &n2) || b.out2d) return 0;
  size_t l;
  return xbin_rows(g.n[0], n1, b, &l) ? l : 0;
}

template <int N1, int N2, int G, int KZ, bool IL, bool HB, bool ODD>
static int launch_xbin(FftArgs& f, BinLaunch& L, long long* xpart, int xpart_rows, size_t lds) {
  constexpr int N = N1 * N2, E = clcm(N1, N2), TS = N / E, B = 2 * G * KZ;
  const void* fn = (const void*)k_fft_x_bin<N1, N2, G, KZ, IL, HB, ODD>;
  if (lds > 64 * 1024) hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  int per_cu = 1;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, TS * B, lds) != hipSuccess || per_cu < 1)
    per_cu = 1;
  const int nzh = f.g->n[2] / 2 + 1;
  const int nkzb = (nzh + KZ - 1) / KZ;
  const int64_t total = (int64_t)(f.g->n[1] / 2 + 1) * nkzb;
  // at least a few items per workgroup (each workgroup's table is summed at
  // the end), at most what fits the CUs at once
  int64_t nblk = (int64_t)(f.ncu > 0 ? f.ncu : 256) * per_cu;
  if (nblk > (total + 3) / 4) nblk = (total + 3) / 4;
  if (nblk < 1) nblk = 1;
  if (nblk > xpart_rows) nblk = xpart_rows;
  const float* g[4];
  for (int q = 0; q < 4; q++) g[q] = (const float*)(q < f.ngrids ? f.grid[q] : f.grid[0]);
  TIMED(f.timer, "fft_x_bin",
        (k_fft_x_bin<N1, N2, G, KZ, IL, HB, ODD><<<dim3((unsigned)nblk, (unsigned)f.nmock), TS * B,
                                                   lds, f.stream>>>(
            g[0], g[1], g[2], g[3], f.px, (int64_t)nzh * f.px, (const cx<float>*)f.tw[0], L.tabs,
            L.dev, f.xq, xpart, nkzb, f.mock_stride)));
  const int nlk = (L.dev.odd ? 5 : 3) * L.dev.nk, fnk = L.dev.nspec * nlk;
  TIMED(f.timer, "bin_final",
        (k_xbin_final<<<dim3((unsigned)(fnk + L.dev.nk), (unsigned)f.nmock), 256, 0, f.stream>>>(
            xpart, (int)nblk, fnk, L.dev.nk, nlk, L.counts, f.xq, L.out, L.out_stride)));
  return hipGetLastError() == hipSuccess ? CAPOW_OK : CAPOW_ERR_HIP;
}

template <int N1, int N2, bool IL, bool HB>
static int launch_xbin_rows(FftArgs& f, BinLaunch& L, long long* xpart, int xpart_rows) {
  constexpr int G = 1 + (IL ? 1 : 0) + (HB ? 1 : 0) + (IL && HB ? 1 : 0);
  constexpr int TS = N1 * N2 / clcm(N1, N2);
  size_t lds;
  const int B = xbin_rows(N1 * N2, N1, L.dev, &lds);
  const bool odd = L.dev.odd != 0;
  if constexpr ((TS * 16) % 64 == 0)
    if (B == 16)
      return odd ? launch_xbin<N1, N2, G, 8 / G, IL, HB, true>(f, L, xpart, xpart_rows, lds)
                 : launch_xbin<N1, N2, G, 8 / G, IL, HB, false>(f, L, xpart, xpart_rows, lds);
  if constexpr ((TS * 8) % 64 == 0 && G <= 4)
    if (B == 8)
      return odd ? launch_xbin<N1, N2, G, 4 / G, IL, HB, true>(f, L, xpart, xpart_rows, lds)
                 : launch_xbin<N1, N2, G, 4 / G, IL, HB, false>(f, L, xpart, xpart_rows, lds);
  return CAPOW_ERR_UNSUPPORTED;
}

int run_fft_x_bin(FftArgs& f, BinLaunch& L, long long* xpart, int xpart_rows) {
  int n1, n2;
  if (!reg_split(f.g->n[0], &n1, &n2) || L.dev.out2d) return CAPOW_ERR_UNSUPPORTED;
  if (!f.ssq || f.ssq_rows < 1 || f.ssq_rows > XBIN_SSQ_MAX || !f.xq) return CAPOW_ERR_UNSUPPORTED;
  const int rc = run_bin_counts(L);
  if (rc) return rc;
  const bool il = L.dev.interlace != 0, hb = L.dev.hasB != 0;
  k_xbin_scale<<<f.nmock, 256, 0, f.stream>>>(f.ssq, f.ssq_rows, f.xbin_factor, il, hb, L.dev, f.xq);
#define X(A_, B_)                                                                          \
  if (n1 == A_ && n2 == B_) {                                                              \
    if (il && hb) return launch_xbin_rows<A_, B_, true, true>(f, L, xpart, xpart_rows);    \
    if (il) return launch_xbin_rows<A_, B_, true, false>(f, L, xpart, xpart_rows);         \
    if (hb) return launch_xbin_rows<A_, B_, false, true>(f, L, xpart, xpart_rows);         \
    return launch_xbin_rows<A_, B_, false, false>(f, L, xpart, xpart_rows);                \
  }
  CAPOW_FFT_SIZES(X)
#undef X
  return CAPOW_ERR_UNSUPPORTED;
}

// Out-of-place pipeline: every pass reads whole rows and writes runs of the
// next pass's row axis.
//   z (r2c): grid [x][y][z]   -> scratch [x][kz][y]   (row pitch py)
//   y (c2c): scratch          -> grid    [ky][kz][x]  (row pitch px; FP64:
//                                          [kz][ky][x])
//   x (c2c): grid             -> scratch [kx][ky][kz] (row pitch rowc)
// ([ky][kz][x] puts the rows of one x-pass batch -- consecutive kz at one
// ky -- next to each other: FP32 x pass 7.6 -> 7.0 ms at 1024^3.)
// The result is in the scratch buffers; the caller swaps them with the grids.
template <typename T>
static bool run_fft_transposed(FftArgs& a) {
  const Geom& g = *a.g;
  const int nx = g.n[0], ny = g.n[1], nz = g.n[2];
  const int64_t nzh = nz / 2 + 1, rowc = g.rowc, px = a.px, py = a.py;
  int n1, n2;
  constexpr bool F64 = sizeof(T) == 8;
  if (!reg_split(nz / 2, &n1, &n2, F64) || !reg_split(ny, &n1, &n2, F64) || !reg_split(nx, &n1, &n2, F64))
    return false;
  // write runs: 128 bytes (FP64: 64 for the y pass; measured best at 1024^3);
  // rows_mul overrides all three
  const int mz = a.rows_mul ? a.rows_mul : 2, my = a.rows_mul ? a.rows_mul : (sizeof(T) == 4 ? 2 : 1),
            mx = a.rows_mul ? a.rows_mul : 2;
  if (a.one_spare && a.ngrids > 1) {
    // grid by grid through the one spare buffer: z grid -> spare, y spare
    // -> grid, [x grid -> spare, then the two swap]
    const bool kyz = sizeof(T) == 4;
    const int64_t y_so = kyz ? px : (int64_t)ny * px, y_sk = kyz ? nzh * px : px;
    const int64_t x_so = kyz ? nzh * px : px, x_sb = kyz ? px : (int64_t)ny * px;
    const bool defer = a.defer_x && kyz && a.ssq && a.xq;
    FftArgs b = a;
    b.ngrids = 1;
    int rows = 0;
    if (defer && zero_u32((uint32_t*)a.ssq, (int64_t)XBIN_SSQ_MAX * 4 * 2, a.stream) != CAPOW_OK)
      return false;
    for (int q = 0; q < a.ngrids; q++) {
      T* Gq[4] = {(T*)a.grid[q], (T*)a.grid[q], (T*)a.grid[q], (T*)a.grid[q]};
      T* Sp[4] = {(T*)a.scratch[0], (T*)a.scratch[0], (T*)a.scratch[0], (T*)a.scratch[0]};
      if (!try_rows_t<T, true>(b, "fft_z", nz / 2, Gq, Sp, nx, ny, (int64_t)ny * rowc, rowc,
                               nzh * py, py, a.tw[2], mz))
        return false;
      b.ssq_slot0 = q;
      if (!try_rows_t<T, false>(b, "fft_y", ny, Sp, Gq, nzh, nx, py, nzh * py, y_so, y_sk, a.tw[1],
                                my, defer ? a.ssq : nullptr))
        return false;
      rows = max(rows, b.ssq_rows);
      if (!defer) {
        if (!try_rows_t<T, false>(b, "fft_x", nx, Gq, Sp, ny, nzh, x_so, x_sb, rowc,
                                  (int64_t)ny * rowc, a.tw[0], mx))
          return false;
        void* t = a.grid[q];
        a.grid[q] = a.scratch[0];
        a.scratch[0] = t;
      }
    }
    a.ssq_rows = rows;
    a.rotated = true;
    if (defer && rows <= XBIN_SSQ_MAX) {
      a.x_deferred = true;
      return true;
    }
    if (defer) {
      // (more y-pass workgroups than ssq rows: the x passes run now)
      for (int q = 0; q < a.ngrids; q++) {
        T* Gq[4] = {(T*)a.grid[q], (T*)a.grid[q], (T*)a.grid[q], (T*)a.grid[q]};
        T* Sp[4] = {(T*)a.scratch[0], (T*)a.scratch[0], (T*)a.scratch[0], (T*)a.scratch[0]};
        if (!try_rows_t<T, false>(b, "fft_x", nx, Gq, Sp, ny, nzh, x_so, x_sb, rowc,
                                  (int64_t)ny * rowc, a.tw[0], mx))
          return false;
        void* t = a.grid[q];
        a.grid[q] = a.scratch[0];
        a.scratch[0] = t;
      }
    }
    return true;
  }
  T* G[4];
  T* S[4];
  for (int q = 0; q < 4; q++) {
    G[q] = q < a.ngrids ? (T*)a.grid[q] : (T*)a.grid[0];
    S[q] = q < a.ngrids ? (T*)a.scratch[q] : (T*)a.scratch[0];
  }
  if (!try_rows_t<T, true>(a, "fft_z", nz / 2, G, S, nx, ny, (int64_t)ny * rowc, rowc,
                           nzh * py, py, a.tw[2], mz))
    return false;
  // FP32: [ky][kz][x] between the y and x passes; FP64 (64-byte y runs):
  // [kz][ky][x], which keeps a y batch's output runs in one plane
  const bool kyz = sizeof(T) == 4;
  const int64_t y_so = kyz ? px : (int64_t)ny * px, y_sk = kyz ? nzh * px : px;
  const int64_t x_so = kyz ? nzh * px : px, x_sb = kyz ? px : (int64_t)ny * px;
  const bool defer = a.defer_x && kyz && a.ssq && a.xq;
  if (!try_rows_t<T, false>(a, "fft_y", ny, S, G, nzh, nx, py, nzh * py, y_so, y_sk, a.tw[1],
                            my, defer ? a.ssq : nullptr))
    return false;
  if (defer && a.ssq_rows <= XBIN_SSQ_MAX) {
    a.x_deferred = true;
    return true;
  }
  // (more y-pass workgroups than ssq rows: the x pass runs now)
  if (!try_rows_t<T, false>(a, "fft_x", nx, G, S, ny, nzh, x_so, x_sb, rowc, (int64_t)ny * rowc,
                            a.tw[0], mx))
    return false;
  return true;
}

// The x pass left pending by a deferred run_fft: grid [ky][kz][x] ->
// scratch [kx][ky][kz] (FP32 layout of run_fft_transposed).
int run_fft_x(FftArgs& a) {
  const Geom& g = *a.g;
  const int nx = g.n[0], ny = g.n[1];
  const int64_t nzh = g.n[2] / 2 + 1, rowc = g.rowc, px = a.px;
  const int mx = a.rows_mul ? a.rows_mul : 2;
  if (a.one_spare && a.ngrids > 1) {
    FftArgs b = a;
    b.ngrids = 1;
    for (int q = 0; q < a.ngrids; q++) {
      float* Gq[4] = {(float*)a.grid[q], (float*)a.grid[q], (float*)a.grid[q], (float*)a.grid[q]};
      float* Sp[4] = {(float*)a.scratch[0], (float*)a.scratch[0], (float*)a.scratch[0],
                      (float*)a.scratch[0]};
      if (!try_rows_t<float, false>(b, "fft_x", nx, Gq, Sp, ny, nzh, nzh * px, px, rowc,
                                    (int64_t)ny * rowc, a.tw[0], mx))
        return CAPOW_ERR_UNSUPPORTED;
      void* t = a.grid[q];
      a.grid[q] = a.scratch[0];
      a.scratch[0] = t;
    }
    a.rotated = true;
    a.x_deferred = false;
    return hipGetLastError() == hipSuccess ? CAPOW_OK : CAPOW_ERR_HIP;
  }
  float* G[4];
  float* S[4];
  for (int q = 0; q < 4; q++) {
    G[q] = q < a.ngrids ? (float*)a.grid[q] : (float*)a.grid[0];
    S[q] = q < a.ngrids ? (float*)a.scratch[q] : (float*)a.scratch[0];
  }
  if (!try_rows_t<float, false>(a, "fft_x", nx, G, S, ny, nzh, nzh * px, px, rowc,
                                (int64_t)ny * rowc, a.tw[0], mx))
    return CAPOW_ERR_UNSUPPORTED;
  a.x_deferred = false;
  return hipGetLastError() == hipSuccess ? CAPOW_OK : CAPOW_ERR_HIP;
}

template <typename T>
static int run_fft_t(FftArgs& a) {
  const Geom& g = *a.g;
  hipStream_t st = a.stream;
  a.swapped = false;
  a.x_deferred = false;
  if (a.fast && a.scratch[0] && run_fft_transposed<T>(a)) {
    a.swapped = !a.x_deferred && !a.rotated;
    return hipGetLastError() == hipSuccess ? CAPOW_OK : CAPOW_ERR_HIP;
  }
  // stacked catalogues: the transposed register passes only
  if (a.nmock > 1) return CAPOW_ERR_UNSUPPORTED;
  T* gp[4];
  for (int q = 0; q < 4; q++) gp[q] = q < a.ngrids ? (T*)a.grid[q] : (T*)a.grid[0];
  const int nx = g.n[0], ny = g.n[1], nz = g.n[2];
  const int nzh = nz / 2 + 1;
  const int64_t nrows = (int64_t)nx * ny;
  // z pass
  const int64_t rowc = g.rowc;
  if (!(a.fast && try_rows2<T>(a, gp, nz / 2, nrows, a.tw[2], g.rowlen))) {
    const int M = nz / 2;
    int rows = max(1, (FFT_EPT * 1024) / max(M, 1));
    rows = min(rows, 16);
    // keep the row batch within 64 KB of LDS
    rows = max(1, min(rows, (int)((64 * 1024) / ((size_t)M * sizeof(cx<T>)))));
    int threads = (rows * M + FFT_EPT - 1) / FFT_EPT;
    threads = ((threads + 63) / 64) * 64;
    threads = min(max(threads, 64), 1024);
    while (rows > 1 && (int64_t)threads * FFT_EPT < (int64_t)rows * M) rows--;
    dim3 grid((unsigned)((nrows + rows - 1) / rows), a.ngrids);
    size_t lds = (size_t)rows * M * sizeof(cx<T>);
    TIMED(a.timer, "fft_z",
          (k_fft_rows<T><<<grid, threads, lds, st>>>(gp[0], gp[1], gp[2], gp[3], nrows, nz, rows,
                                                     a.plan[2], (const cx<T>*)a.tw[2], g.rowlen)));
  }
  // y pass: planes = x index, columns = kz
  if (!(a.fast && try_cols2<T, 1>(a, "fft_y", gp, ny, rowc, (int64_t)ny * rowc,
                               (int64_t)nzh, nx, a.tw[1]))) {
    int64_t lim = 8192;  // N*B complex per workgroup (FFT_EPT * 1024)
    TIMED_COLS(a, "fft_y", gp, ny, rowc, (int64_t)ny * rowc, (int64_t)nzh, nx, a.plan[1],
               a.tw[1], lim);
  }
  // x pass: one plane, columns = (y, kz) over the padded rows (the padding
  // columns share the rows' last cache line and are never read as modes)
  const int64_t ncols = (int64_t)ny * rowc;
  if (!(a.fast && try_cols2<T, 0>(a, "fft_x", gp, nx, ncols, 0, ncols, 1, a.tw[0]))) {
    int64_t lim = 8192;
    TIMED_COLS(a, "fft_x", gp, nx, ncols, 0, ncols, 1, a.plan[0], a.tw[0], lim);
  }
  return hipGetLastError() == hipSuccess ? CAPOW_OK : CAPOW_ERR_HIP;
}

template <typename T, int N1, int N2>
static void launch_rows2_adj(FftArgs& a, T** gp, int64_t nrows, const void* twz, int64_t rowlen) {
  constexpr int M = N1 * N2, E = clcm(N1, N2), TS = M / E;
  constexpr int B = rows_B<T, N1, N2>();
  const size_t lds = (size_t)B * (M + N1) * sizeof(cx<T>);
  const int64_t nbx = (nrows + B - 1) / B;
  dim3 grid((unsigned)nbx, a.ngrids);
  TIMED(a.timer, "ifft_z",
        (k_fft_rows2_adj<T, N1, N2, B><<<grid, TS * B, lds, a.stream>>>(
            gp[0], gp[1], gp[2], gp[3], nrows, (const cx<T>*)twz, nbx, rowlen)));
}

// Adjoint of run_fft on gradient grids: inverse x and y passes (conjugated
// register kernels), then the z-pass adjoint.  Register sizes only.
template <typename T>
static int run_ifft_t(FftArgs& a) {
  const Geom& g = *a.g;
  T* gp[4];
  for (int q = 0; q < 4; q++) gp[q] = q < a.ngrids ? (T*)a.grid[q] : (T*)a.grid[0];
  const int nx = g.n[0], ny = g.n[1], nz = g.n[2];
  const int nzh = nz / 2 + 1;
  const int64_t rowc = g.rowc, nrows = (int64_t)nx * ny;
  int n1, n2;
  constexpr bool F64 = sizeof(T) == 8;
  if (!reg_split(nx, &n1, &n2, F64) || !reg_split(ny, &n1, &n2, F64) || !reg_split(nz / 2, &n1, &n2, F64))
    return CAPOW_ERR_UNSUPPORTED;
  const int64_t ncols = (int64_t)ny * rowc;
  if (!try_cols2<T, 2>(a, "ifft_x", gp, nx, ncols, 0, ncols, 1, a.tw[0])) return CAPOW_ERR_UNSUPPORTED;
  if (!try_cols2<T, 3>(a, "ifft_y", gp, ny, rowc, (int64_t)ny * rowc, (int64_t)nzh, nx, a.tw[1]))
    return CAPOW_ERR_UNSUPPORTED;
  reg_split(nz / 2, &n1, &n2);
  bool done = false;
#define X(A, B_)                                                   \
  if constexpr (split_for<T, A, B_>())                             \
  if (!done && n1 == A && n2 == B_) {                              \
    launch_rows2_adj<T, A, B_>(a, gp, nrows, a.tw[2], g.rowlen);   \
    done = true;                                                   \
  }
  CAPOW_FFT_SIZES(X)
#undef X
  if (!done) return CAPOW_ERR_UNSUPPORTED;
  return hipGetLastError() == hipSuccess ? CAPOW_OK : CAPOW_ERR_HIP;
}

int run_ifft(FftArgs& a, bool fp64) { return fp64 ? run_ifft_t<double>(a) : run_ifft_t<float>(a); }

int run_fft(FftArgs& a, bool fp64) { return fp64 ? run_fft_t<double>(a) : run_fft_t<float>(a); }

}  // namespace capow
