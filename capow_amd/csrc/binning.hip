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
// Accumulation: every wavefront owns a private copy of the bins in LDS (adds
// in program order => bit-reproducible); workgroup partials are then summed
// in a fixed order.
#include "binning_math.h"

namespace capow {

// Per-mode binning: geometry (mode_geom), interlace combine, power of each
// spectrum, Legendre weights, accumulation into this wave's bins.
template <typename T>
__device__ __forceinline__ void bin_mode(const RowVals& R, int k, double fz, double corz,
                                         double pzr, double pzi, typename CT<T>::t va,
                                         typename CT<T>::t vai, typename CT<T>::t vb,
                                         typename CT<T>::t vbi, const BinArgsDev& a,
                                         double* accS, double* acc2, double* cnt, double* cnt2,
                                         int nl, int n2d) {
  const ModeGeom G = mode_geom(R, k, fz, corz, pzr, pzi, a);
  if (!G.ok) return;
  const int NK = a.nk;
  const int kbin = G.kbin;
  const double mu = G.mu, NM = G.NM, cor = G.cor, cr = G.cr, ci = G.ci;
  double ar = (double)va.x, ai = (double)va.y, br = 0, bi = 0;
  if (a.interlace) {
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
  const int b2 = G.b2;
  const bool in2d = G.in2d;
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

// Flat-index form: r is the offset within plane i of the padded layout.
template <typename T>
__device__ __forceinline__ void bin_one(int i, uint32_t r, typename CT<T>::t va,
                                        typename CT<T>::t vai, typename CT<T>::t vb,
                                        typename CT<T>::t vbi, const double* __restrict__ tabs,
                                        const BinArgsDev& a, double* accS, double* acc2,
                                        double* cnt, double* cnt2, int nl, int n2d) {
  const uint32_t j = a.fd_rowc.div(r);
  const int k = (int)(r - j * (uint32_t)a.rowc);
  if (k >= a.nzh) return;   // row padding
  const RowVals R = row_vals(i, (int)j, tabs, a);
  bin_mode<T>(R, k, tabs[a.ofz + k], tabs[a.ocz + k], tabs[a.opz + 2 * k], tabs[a.opz + 2 * k + 1],
              va, vai, vb, vbi, a, accS, acc2, cnt, cnt2, nl, n2d);
}

// One workgroup = `copies` wavefronts, each with a private copy of the bins
// in LDS (adds in program order => bit-reproducible).  A lane handles U
// consecutive modes per iteration; the next iteration's loads are issued
// before the current modes are binned, so every wave keeps two groups of
// loads in flight (the LDS copies leave few waves per CU to hide latency).
template <typename T, int U>
__global__ __launch_bounds__(512) void k_bin(const T* __restrict__ gA, const T* __restrict__ gAi,
                                             const T* __restrict__ gB, const T* __restrict__ gBi,
                                             const double* __restrict__ tabs, BinArgsDev a,
                                             double* __restrict__ partials) {
  typedef typename CT<T>::t C;
  extern __shared__ double sm[];
  const int NK = a.nk;
  const int nl = a.odd ? 5 : 3;
  const int F = a.nspec * nl;                       // sums per bin
  const int n2d = a.out2d ? a.nk * a.nmu : 0;
  const int per_copy = F * NK + a.nspec * n2d + NK + n2d;  // doubles (counts as double)
  const int wave = threadIdx.x >> 6;
  double* my = sm + wave * per_copy;
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
  const bool il = a.interlace != 0, hb = a.hasB != 0;
  // lane handles U consecutive modes; groups are strided over the grid
  const int64_t gstride = (int64_t)gridDim.x * blockDim.x * U;
  int64_t m0 = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) * U;
  C va[U], vai[U], vb[U], vbi[U];
#pragma unroll
  for (int u = 0; u < U; u++) {
    const int64_t m = m0 + u;
    const bool ok = m < total;
    va[u] = ok ? A[m] : zero;
    vai[u] = (ok && il) ? Ai[m] : zero;
    vb[u] = (ok && hb) ? B[m] : zero;
    vbi[u] = (ok && hb && il) ? Bi[m] : zero;
  }
  for (; m0 < total; m0 += gstride) {
    // prefetch the next group
    C na[U], nai[U], nb[U], nbi[U];
    const int64_t m1 = m0 + gstride;
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int64_t m = m1 + u;
      const bool ok = m < total;
      na[u] = ok ? A[m] : zero;
      nai[u] = (ok && il) ? Ai[m] : zero;
      nb[u] = (ok && hb) ? B[m] : zero;
      nbi[u] = (ok && hb && il) ? Bi[m] : zero;
    }
    int i = (int)(m0 / plane);
    uint32_t r = (uint32_t)(m0 - (int64_t)i * plane);
#pragma unroll
    for (int u = 0; u < U; u++) {
      if (m0 + u < total) bin_one<T>(i, r, va[u], vai[u], vb[u], vbi[u], tabs, a, accS, acc2, cnt, cnt2, nl, n2d);
      if (++r == plane) { r = 0; i++; }
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      va[u] = na[u];
      vai[u] = nai[u];
      vb[u] = nb[u];
      vbi[u] = nbi[u];
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

template <typename T>
__device__ __forceinline__ ModeVals mode_vals(const RowVals& R, int k, double fz, double corz,
                                              double pzr, double pzi, typename CT<T>::t va,
                                              typename CT<T>::t vai, typename CT<T>::t vb,
                                              typename CT<T>::t vbi, const BinArgsDev& a) {
  ModeVals V;
  const double twopi = 2.0 * 3.14159265358979323846;
  const KBin kb = kbin_fast(R.fxy, fz, a);
  const double fktot = kb.fktot;
  V.near = kb.near;
  V.kbin = kb.kbin;
  V.ok = kb.ok;
  const double fl = a.los == 1 ? R.fx : (a.los == 2 ? R.fy : fz);
  double rk = __builtin_amdgcn_rcp(fktot > 0 ? fktot : 1.0);
  rk = rk * (2.0 - fktot * rk);
  rk = rk * (2.0 - fktot * rk);
  const double mu = fktot > 0 ? (twopi * fl) * rk : 0.0;
  const double wc = nm_weight(R.i, R.j, k, a.nx) * (R.corxy * corz);
  double cr = 1.0, ci = 0.0;
  double ar = (double)va.x, ai = (double)va.y, br = 0, bi = 0;
  if (a.interlace) {
    cr = R.qr * pzr - R.qi * pzi;
    ci = R.qr * pzi + R.qi * pzr;
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
#pragma unroll
  for (int s = 0; s < 3; s++) {
    double p;
    switch (a.kind[s]) {
      case SPEC_AUTO_A: p = ar * ar + ai * ai; break;
      case SPEC_AUTO_B: p = br * br + bi * bi; break;
      case SPEC_CROSS_IM: p = ar * bi - ai * br; break;
      default: p = ar * br + ai * bi; break;
    }
    V.np[s] = wc * p;
  }
  const double mu2 = mu * mu;
  V.mu = mu;
  V.L2 = 1.5 * mu2 - 0.5;
  V.L4 = 4.375 * mu2 * mu2 - 3.75 * mu2 + 0.375;
  return V;
}

// V consecutive complex values from p (16-byte loads where they fit: p is
// 16-byte aligned whenever V > 1).  FP32: non-temporal loads (every value is
// read once; measured 3.11 -> 3.03 ms for the C2 binning).
template <typename C, int V>
__device__ __forceinline__ void load_run(const C* __restrict__ p, C* out) {
  if constexpr (sizeof(C) == 8 && V >= 2) {
#pragma unroll
    for (int u = 0; u < V; u += 2) {
      typedef float f4v __attribute__((ext_vector_type(4)));
      const f4v q = __builtin_nontemporal_load(reinterpret_cast<const f4v*>(p + u));
      out[u].x = q.x;
      out[u].y = q.y;
      out[u + 1].x = q.z;
      out[u + 1].y = q.w;
    }
  } else {
#pragma unroll
    for (int u = 0; u < V; u++) out[u] = p[u];
  }
}

// Row form.  A wave takes batches of 64 consecutive (i, j) rows, with the
// row's x/y quantities wave-uniform.  Lane l covers kz = p 64V + V l + u
// (slot s = p V + u, u < V, p < NUF / V): V consecutive modes per lane, so
// that one instruction's 64 lanes span the whole row and fall into many
// different k bins (adjacent kz share bins wherever kz << k, and same-bin
// lanes serialise the LDS adds).  The z tables sit in LDS in slot-major
// order [s][lane].  The rest of each row, kz in [64 NUF, NZ/2+1), is binned
// once per batch with lane l on row l of the batch.  PREF: the next row's
// loads are issued before the current row is binned.
template <typename T, int NUF, bool PREF, bool FAST>
__global__ __launch_bounds__(512) void k_bin_rows(const T* __restrict__ gA,
                                                  const T* __restrict__ gAi,
                                                  const T* __restrict__ gB,
                                                  const T* __restrict__ gBi,
                                                  const double* __restrict__ tabs, BinArgsDev a,
                                                  double* __restrict__ partials) {
  typedef typename CT<T>::t C;
  extern __shared__ double sm[];
  const int NK = a.nk;
  const int nl = a.odd ? 5 : 3;
  const int F = a.nspec * nl;
  const int n2d = a.out2d ? a.nk * a.nmu : 0;
  const int per_copy = F * NK + a.nspec * n2d + NK + n2d;
  // LDS doubles per wave: the branch-free form holds only the sums (its mode
  // counts come from k_bin_counts)
  const int lcp = FAST ? F * NK : per_copy;
  const int copies = a.copies;
  const int nzh = a.nzh;
  constexpr int V = NUF < (sizeof(T) == 4 ? 8 : 4) ? NUF : (sizeof(T) == 4 ? 8 : 4);
  const int rem0 = 64 * NUF;
  // z tables, 4 doubles per kz (fz, corz, phz.re, phz.im): slots [s][lane]
  // for kz < 64 NUF, then the rest of the row in kz order
  double* zt = sm + copies * lcp;
  double* ztr = zt + 4 * rem0;
  for (int q = threadIdx.x; q < lcp * copies; q += blockDim.x) sm[q] = 0.0;
  for (int q = threadIdx.x; q < nzh; q += blockDim.x) {
    int k = q;
    if (q < rem0) {
      const int sl = q >> 6, l = q & 63;
      k = (sl / V) * 64 * V + V * l + (sl % V);
    }
    zt[4 * q] = tabs[a.ofz + k];
    zt[4 * q + 1] = tabs[a.ocz + k];
    zt[4 * q + 2] = tabs[a.opz + 2 * k];
    zt[4 * q + 3] = tabs[a.opz + 2 * k + 1];
  }
  __syncthreads();
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int lane = threadIdx.x & 63;
  double* my = sm + wave * lcp;
  double* accS = my;
  double* acc2 = my + F * NK;
  double* cnt = acc2 + a.nspec * n2d;
  double* cnt2 = cnt + NK;
  const C* A = reinterpret_cast<const C*>(gA);
  const C* Ai = reinterpret_cast<const C*>(gAi);
  const C* B = reinterpret_cast<const C*>(gB);
  const C* Bi = reinterpret_cast<const C*>(gBi);
  const C zero = {T(0), T(0)};
  const bool il = a.interlace != 0, hb = a.hasB != 0;
  const int64_t nrows = (int64_t)a.nx * a.ny;
  const int64_t nbatch = (nrows + 63) / 64;
  const int64_t rowc = a.rowc;
  const int nw = gridDim.x * copies;
  for (int64_t bt = (int64_t)blockIdx.x * copies + wave; bt < nbatch; bt += nw) {
    const int64_t r0 = bt * 64;
    const int nr = (int)min((int64_t)64, nrows - r0);
    if constexpr (NUF > 0) {
      C va[NUF], vai[NUF], vb[NUF], vbi[NUF];
      auto load = [&](int64_t row, C* xa, C* xai, C* xb, C* xbi) {
        const int64_t base = row * rowc + V * lane;
#pragma unroll
        for (int p = 0; p < NUF / V; p++) {
          const int64_t m = base + 64 * V * p;
          load_run<C, V>(A + m, xa + p * V);
          if (il) load_run<C, V>(Ai + m, xai + p * V);
          else
#pragma unroll
            for (int u = 0; u < V; u++) xai[p * V + u] = zero;
          if (hb) load_run<C, V>(B + m, xb + p * V);
          else
#pragma unroll
            for (int u = 0; u < V; u++) xb[p * V + u] = zero;
          if (hb && il) load_run<C, V>(Bi + m, xbi + p * V);
          else
#pragma unroll
            for (int u = 0; u < V; u++) xbi[p * V + u] = zero;
        }
      };
      load(r0, va, vai, vb, vbi);
      for (int rr = 0; rr < nr; rr++) {
        const int64_t row = r0 + rr;
        C na[NUF], nai[NUF], nb[NUF], nbi[NUF];
        if (PREF && rr + 1 < nr) load(row + 1, na, nai, nb, nbi);
        const int i = (int)(row / a.ny), j = (int)(row - (int64_t)i * a.ny);
        const RowVals R = row_vals(i, j, tabs, a);
        if (FAST) {
          // groups of modes evaluated branch-free, then accumulated
          // modes per group: as many independent chains as the registers allow
          constexpr int GRP = NUF < 2 ? NUF : (NUF >= 8 ? 2 : (NUF < 4 ? NUF : 4));
#pragma unroll
          for (int u0 = 0; u0 < NUF; u0 += GRP) {
            ModeVals MV[GRP];
#pragma unroll
            for (int q = 0; q < GRP; q++) {
              const int u = u0 + q;
              const int k = (u / V) * 64 * V + V * lane + (u % V);
              const double4 z = *reinterpret_cast<const double4*>(zt + 4 * (64 * u + lane));
              MV[q] = mode_vals<T>(R, k, z.x, z.y, z.z, z.w, va[u], vai[u], vb[u], vbi[u], a);
            }
            bool anynear = false;
#pragma unroll
            for (int q = 0; q < GRP; q++) anynear |= MV[q].near;
            if (__any(anynear)) {
#pragma unroll
              for (int q = 0; q < GRP; q++)
                if (MV[q].near) {
                  const KBin e = kbin_exact(R.fxy, zt[4 * (64 * (u0 + q) + lane)], a);
                  MV[q].kbin = e.kbin;
                  MV[q].ok = e.ok;
                }
            }
#pragma unroll
            for (int q = 0; q < GRP; q++)
              if (MV[q].ok) mode_add(MV[q], a, accS, nl);
          }
        } else {
#pragma unroll
          for (int u = 0; u < NUF; u++) {
            const int k = (u / V) * 64 * V + V * lane + (u % V);
            const double4 z = *reinterpret_cast<const double4*>(zt + 4 * (64 * u + lane));
            bin_mode<T>(R, k, z.x, z.y, z.z, z.w, va[u], vai[u], vb[u], vbi[u], a, accS, acc2, cnt,
                        cnt2, nl, n2d);
          }
        }
        if (rr + 1 < nr) {
          if (PREF) {
#pragma unroll
            for (int u = 0; u < NUF; u++) {
              va[u] = na[u];
              vai[u] = nai[u];
              vb[u] = nb[u];
              vbi[u] = nbi[u];
            }
          } else {
            load(row + 1, va, vai, vb, vbi);
          }
        }
      }
    }
    // the rest of each row: lane l on row r0 + l
    if (lane < nr) {
      const int64_t row = r0 + lane;
      const int i = (int)(row / a.ny), j = (int)(row - (int64_t)i * a.ny);
      const RowVals R = row_vals(i, j, tabs, a);
      for (int k = rem0; k < nzh; k++) {
        const int64_t m = row * rowc + k;
        const C xa = A[m];
        const C xai = il ? Ai[m] : zero;
        const C xb = hb ? B[m] : zero;
        const C xbi = (hb && il) ? Bi[m] : zero;
        const double4 z = *reinterpret_cast<const double4*>(ztr + 4 * (k - rem0));
        if (FAST) {
          ModeVals mv = mode_vals<T>(R, k, z.x, z.y, z.z, z.w, xa, xai, xb, xbi, a);
          if (mv.near) {
            const KBin e = kbin_exact(R.fxy, z.x, a);
            mv.kbin = e.kbin;
            mv.ok = e.ok;
          }
          if (mv.ok) mode_add(mv, a, accS, nl);
        } else {
          bin_mode<T>(R, k, z.x, z.y, z.z, z.w, xa, xai, xb, xbi, a, accS, acc2, cnt, cnt2, nl,
                      n2d);
        }
      }
    }
  }
  __syncthreads();
  double* out = partials + (int64_t)blockIdx.x * per_copy;
  for (int q = threadIdx.x; q < per_copy; q += blockDim.x) {
    double t = 0.0;
    if (q < lcp)
      for (int c = 0; c < copies; c++) t += sm[c * lcp + q];
    out[q] = t;
  }
}

// The four grids of the live rows of a group at MPL consecutive modes
// k0 .. k0 + MPL - 1 of every row (one 8 * MPL-byte (FP32) or 16-byte (FP64)
// load per row and grid; lanes beyond the row load nothing).
template <typename C, int MPL>
__device__ __forceinline__ void group_load(const C* __restrict__ A, const C* __restrict__ Ai,
                                           const C* __restrict__ B, const C* __restrict__ Bi,
                                           const int64_t* row, const bool* live, int64_t rowc,
                                           int k0, int nzh, bool il, bool hb, C (&v)[4][4][MPL]) {
  const C zero = {0, 0};
  const bool in = k0 < nzh;
#pragma unroll
  for (int r = 0; r < 4; r++) {
    const int64_t m = row[r] * rowc + k0;
    const bool lv = live[r] && in;
    const C* src[4] = {A, Ai, B, Bi};
    const bool use[4] = {true, il, hb, hb && il};
#pragma unroll
    for (int g = 0; g < 4; g++) {
      if (lv && use[g]) {
        load_run<C, MPL>(src[g] + m, v[r][g]);
      } else {
#pragma unroll
        for (int u = 0; u < MPL; u++) v[r][g][u] = zero;
      }
    }
  }
}

// A wave walks groups g = w0, w0 + nw, ...; each group in nch chunks of
// 64 MPL modes (lane l: kz = 64 MPL c + MPL l + u; the last chunk may end
// past the row, its lanes there bin nothing); the next chunk's loads are in
// flight while the current one is binned.  Modes kz >= 64 MPL nch (a few at
// the end of the row that do not fill a chunk) are binned afterwards with one
// group per lane.
// Accumulation of one evaluated group mode in the mirror-group kernel.
// MODE 0: the 1-D sums into this wave's LDS copy; MODE 1: also the mode's
// bound into the copy's extra NK entries (the per-k scale of MODE 2);
// MODE 2: the (k, mu) sums as integer quanta of their k bin's scale
// (np * qinv[kbin] rounded) into the workgroup's one int64 table: integer
// adds do not depend on their order, so all waves share it.
template <int MODE>
__device__ __forceinline__ void mirror_accum(const ModeVals& mv, double fl, const BinArgsDev& a,
                                             double* accS, unsigned long long* acc2i,
                                             const double* __restrict__ qinv, int nl, int FNK,
                                             int n2d) {
  if constexpr (MODE < 2) {
    mode_add(mv, a, accS, nl);
    if constexpr (MODE == 1) atomicAdd(&accS[FNK + mv.kbin], mv.bound);
  } else {
    bool in2d;
    const int b2 = mu_bin2d(mv.kbin, fl, mv.fktot, a, in2d);
    if (!in2d) return;
    const double iq = qinv[mv.kbin];
#pragma unroll
    for (int s = 0; s < 3; s++) {
      if (s >= a.nspec) break;
      atomicAdd(&acc2i[s * n2d + b2], (unsigned long long)__double2ll_rn(mv.np[s] * iq));
    }
  }
}

template <typename T, int MPL, bool IL, bool HB, int MODE>
__global__ __launch_bounds__(512) void k_bin_mirror(const T* __restrict__ gA,
                                                    const T* __restrict__ gAi,
                                                    const T* __restrict__ gB,
                                                    const T* __restrict__ gBi,
                                                    const double* __restrict__ tabs, BinArgsDev a,
                                                    int nch, double* __restrict__ partials,
                                                    const double* __restrict__ qinv,
                                                    long long* __restrict__ partials2) {
  typedef typename CT<T>::t C;
  extern __shared__ double sm[];
  constexpr int CH = 64 * MPL;
  const int NK = a.nk;
  const int nl = a.odd ? 5 : 3;
  const int F = a.nspec * nl;
  const int n2d = a.out2d ? a.nk * a.nmu : 0;
  // MODE < 2: one copy of the sums (and bounds) per wave; MODE 2: one int64
  // (k, mu) table per workgroup
  const int lcp = MODE == 2 ? a.nspec * n2d : F * NK + (MODE == 1 ? NK : 0);
  const int per_copy = F * NK + a.nspec * n2d + NK + n2d;
  const int copies = MODE == 2 ? 1 : a.copies;
  const int nzh = a.nzh;
  const int ktail = nch * CH;
  double* zt = sm + copies * lcp;   // fz, corz, phz.re, phz.im per kz
  unsigned long long* acc2i = reinterpret_cast<unsigned long long*>(sm);
  for (int q = threadIdx.x; q < lcp * copies; q += blockDim.x) sm[q] = 0.0;
  for (int q = threadIdx.x; q < nzh; q += blockDim.x) {
    zt[4 * q] = tabs[a.ofz + q];
    zt[4 * q + 1] = tabs[a.ocz + q];
    zt[4 * q + 2] = tabs[a.opz + 2 * q];
    zt[4 * q + 3] = tabs[a.opz + 2 * q + 1];
  }
  __syncthreads();
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int lane = threadIdx.x & 63;
  double* accS = sm + (MODE == 2 ? 0 : wave) * lcp;
  const C* A = reinterpret_cast<const C*>(gA);
  const C* Ai = reinterpret_cast<const C*>(gAi);
  const C* B = reinterpret_cast<const C*>(gB);
  const C* Bi = reinterpret_cast<const C*>(gBi);
  constexpr bool il = IL, hb = HB;
  const int64_t rowc = a.rowc;
  const int hy1 = a.ny / 2 + 1;
  const int ng = (a.nx / 2 + 1) * hy1;
  const int wpb = blockDim.x >> 6;   // waves per workgroup
  const int nw = gridDim.x * wpb;
  const int w0 = blockIdx.x * wpb + wave;
  if (nch > 0 && w0 < ng) {
    C cur[4][4][MPL];
    {
      int i0, j0;
      int64_t row[4];
      bool live[4];
      group_rows(w0, hy1, a, i0, j0, row, live);
      group_load<C, MPL>(A, Ai, B, Bi, row, live, rowc, MPL * lane, nzh, il, hb, cur);
    }
    for (int g = w0; g < ng; g += nw) {
      const Group G = group_of(g, hy1, tabs, a);
      for (int c = 0; c < nch; c++) {
        const int gn = c + 1 < nch ? g : g + nw, cn = c + 1 < nch ? c + 1 : 0;
        C nxt[4][4][MPL];
        if (gn < ng) {
          int in_, jn;
          int64_t row[4];
          bool live[4];
          group_rows(gn, hy1, a, in_, jn, row, live);
          group_load<C, MPL>(A, Ai, B, Bi, row, live, rowc, cn * CH + MPL * lane, nzh, il, hb,
                             nxt);
        }
        ModeVals MV[MPL];
#pragma unroll
        for (int u = 0; u < MPL; u++) {
          const int k = c * CH + MPL * lane + u;
          const double4 z = *reinterpret_cast<const double4*>(zt + 4 * (k < nzh ? k : nzh - 1));
          C v[4][4];
#pragma unroll
          for (int r = 0; r < 4; r++)
#pragma unroll
            for (int q = 0; q < 4; q++) v[r][q] = cur[r][q][u];
          MV[u] = group_mode<C, IL, HB>(G, k, z.x, z.y, z.z, z.w, v, a);
        }
        bool anynear = false;
#pragma unroll
        for (int u = 0; u < MPL; u++) anynear |= MV[u].near;
        if (__any(anynear)) {
#pragma unroll
          for (int u = 0; u < MPL; u++)
            if (MV[u].near) {
              const int k = c * CH + MPL * lane + u;
              const KBin e = kbin_exact(G.fxy, zt[4 * (k < nzh ? k : nzh - 1)], a);
              MV[u].kbin = e.kbin;
              MV[u].ok = e.ok;
            }
        }
#pragma unroll
        for (int u = 0; u < MPL; u++)
          if (MV[u].ok && c * CH + MPL * lane + u < nzh) {
            const int k = c * CH + MPL * lane + u;
            mirror_accum<MODE>(MV[u], zt[4 * (k < nzh ? k : nzh - 1)], a, accS, acc2i, qinv, nl,
                               F * NK, n2d);
          }
        if (gn < ng) {
#pragma unroll
          for (int r = 0; r < 4; r++)
#pragma unroll
            for (int q = 0; q < 4; q++)
#pragma unroll
              for (int u = 0; u < MPL; u++) cur[r][q][u] = nxt[r][q][u];
        }
      }
    }
  }
  if (ktail < nzh) {
    // the end of every row: one group per lane
    const C zero = {0, 0};
    for (int gb = w0 * 64; gb < ng; gb += nw * 64) {
      const int g = gb + lane;
      if (g < ng) {
        const Group G = group_of(g, hy1, tabs, a);
        for (int k = ktail; k < nzh; k++) {
          C v[4][4];
#pragma unroll
          for (int r = 0; r < 4; r++) {
            const int64_t m = G.row[r] * rowc + k;
            const bool lv = G.live[r];
            v[r][0] = lv ? A[m] : zero;
            v[r][1] = lv && il ? Ai[m] : zero;
            v[r][2] = lv && hb ? B[m] : zero;
            v[r][3] = lv && hb && il ? Bi[m] : zero;
          }
          const double4 z = *reinterpret_cast<const double4*>(zt + 4 * k);
          ModeVals mv = group_mode<C, IL, HB>(G, k, z.x, z.y, z.z, z.w, v, a);
          if (mv.near) {
            const KBin e = kbin_exact(G.fxy, z.x, a);
            mv.kbin = e.kbin;
            mv.ok = e.ok;
          }
          if (mv.ok) mirror_accum<MODE>(mv, z.x, a, accS, acc2i, qinv, nl, F * NK, n2d);
        }
      }
    }
  }
  __syncthreads();
  if constexpr (MODE == 2) {
    long long* out = partials2 + (int64_t)blockIdx.x * lcp;
    for (int q = threadIdx.x; q < lcp; q += blockDim.x) out[q] = (long long)acc2i[q];
  } else {
    double* out = partials + (int64_t)blockIdx.x * per_copy;
    for (int q = threadIdx.x; q < per_copy; q += blockDim.x) {
      double t = 0.0;
      if (q < lcp)
        for (int c = 0; c < copies; c++) t += sm[c * lcp + q];
      out[q] = t;
    }
  }
}

// Per-k scales of the (k, mu) quanta from the bounds (sums of NM cor
// (|A|^2 + |B|^2) over each k bin, which no (k, mu) partial sum of any
// spectrum exceeds in magnitude): q = B 2^-61 (exact), qinv = 2^61 / B, so
// the int64 tables cannot overflow.
__global__ void k_bin_scale(const double* __restrict__ bound, int nk, double* __restrict__ q) {
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < nk; k += gridDim.x * blockDim.x) {
    const double b = bound[k];
    q[k] = b > 0 ? ldexp(b, -61) : 0.0;
    q[nk + k] = b > 0 ? ldexp(1.0, 61) / b : 0.0;
  }
}

// (k, mu) sums: the int64 workgroup tables summed exactly, times their k
// bin's quantum, into out[0 .. nspec * n2d).
__global__ void k_bin_final2d(const long long* __restrict__ partials2, int nwg, int len, int n2d,
                              int nmu, const double* __restrict__ q, double* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= len) return;
  long long t = 0;
  for (int w = 0; w < nwg; w++) t += partials2[(int64_t)w * len + i];
  out[i] = (double)t * q[(i % n2d) / nmu];
}

// Mode counts of the branch-free form (they depend on the geometry and the
// k bins only): one wave per row, kz over the lanes; the same bin decisions
// as k_bin_rows (kbin_fast, kbin_exact near integers).  Counts are integers
// held exactly in doubles, so the atomic order does not change the result.
__global__ __launch_bounds__(256) void k_bin_counts(const double* __restrict__ tabs,
                                                    BinArgsDev a, double* __restrict__ counts) {
  extern __shared__ double sc[];
  const int n2d = a.out2d ? a.nk * a.nmu : 0;   // (k, mu) counts after the k counts
  for (int q = threadIdx.x; q < a.nk + n2d; q += blockDim.x) sc[q] = 0.0;
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int64_t nrows = (int64_t)a.nx * a.ny;
  const int64_t nw = (int64_t)gridDim.x * (blockDim.x >> 6);
  for (int64_t row = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); row < nrows;
       row += nw) {
    const int i = (int)(row / a.ny), j = (int)(row - (int64_t)i * a.ny);
    const double fx = tabs[a.ofx + i], fy = tabs[a.ofy + j];
    const double fxy = dadd(dmul(fx, fx), dmul(fy, fy));
    for (int k = lane; k < a.nzh; k += 64) {
      KBin kb = kbin_any(fxy, tabs[a.ofz + k], a);
      if (kb.near) kb = kbin_exact(fxy, tabs[a.ofz + k], a);
      if (kb.ok) {
        const double nm = nm_weight(i, j, k, a.nx);
        atomicAdd(&sc[kb.kbin], nm);
        if (n2d) {
          const double fz = tabs[a.ofz + k];
          bool in2d;
          const int b2 = mu_bin2d(kb.kbin, a.los == 1 ? fx : (a.los == 2 ? fy : fz), kb.fktot, a,
                                  in2d);
          if (in2d) atomicAdd(&sc[a.nk + b2], nm);
        }
      }
    }
  }
  __syncthreads();
  for (int q = threadIdx.x; q < a.nk + n2d; q += blockDim.x)
    if (sc[q] != 0.0) atomicAdd(&counts[q], sc[q]);
}

// Fixed-order sum over workgroup partials; with `counts`, the mode counts
// [cnt_off, cnt_off + ncnt) are taken from there.
__global__ void k_bin_final(const double* __restrict__ partials, int nwg, int per_copy,
                            const double* __restrict__ counts, int cnt_off, int ncnt,
                            double* __restrict__ out) {
  int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= per_copy) return;
  if (counts && q >= cnt_off && q < cnt_off + ncnt) {
    out[q] = counts[q - cnt_off];
    return;
  }
  double t = 0.0;
  for (int w = 0; w < nwg; w++) t += partials[(int64_t)w * per_copy + q];
  out[q] = t;
}

// Chunks per row of the mirror-group form: a remainder of 16 modes or more
// takes a partial chunk (coalesced loads), a smaller one the tail pass.
static int mirror_chunks(int nzh, int ch) { return nzh / ch + (nzh % ch >= 16 ? 1 : 0); }

// FP32: two modes per lane (one 16-byte load per row and grid) once a row
// fills a 128-mode chunk; FP64: one.  MODE 2 runs 8 waves on one shared
// table.
template <typename T, bool IL, bool HB, int MODE>
static void launch_mirror(BinLaunch& L, size_t lds) {
  constexpr int M2 = sizeof(T) == 4 ? 2 : 1;
  const bool two = M2 == 2 && L.dev.nzh >= 128;
  const int threads = MODE == 2 ? 512 : 64 * L.dev.copies;
  const void* fn = two ? (const void*)k_bin_mirror<T, M2, IL, HB, MODE>
                       : (const void*)k_bin_mirror<T, 1, IL, HB, MODE>;
  if (lds > 64 * 1024) hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  const double* qinv = MODE == 2 ? L.qscale + L.dev.nk : nullptr;
  const char* name = MODE == 2 ? "bin_2d" : "bin";
  if (two)
    TIMED(L.timer, name,
          (k_bin_mirror<T, M2, IL, HB, MODE><<<L.nwg, threads, lds, L.stream>>>(
              (const T*)L.grid[0], (const T*)L.grid[1], (const T*)L.grid[2], (const T*)L.grid[3],
              L.tabs, L.dev, mirror_chunks(L.dev.nzh, 64 * M2), L.partials, qinv, L.partials2)));
  else
    TIMED(L.timer, name,
          (k_bin_mirror<T, 1, IL, HB, MODE><<<L.nwg, threads, lds, L.stream>>>(
              (const T*)L.grid[0], (const T*)L.grid[1], (const T*)L.grid[2], (const T*)L.grid[3],
              L.tabs, L.dev, mirror_chunks(L.dev.nzh, 64), L.partials, qinv, L.partials2)));
}

template <typename T, int MODE>
static void launch_mirror_any(BinLaunch& L, size_t lds) {
  const bool il = L.dev.interlace != 0, hb = L.dev.hasB != 0;
  if (il && hb) launch_mirror<T, true, true, MODE>(L, lds);
  else if (il) launch_mirror<T, true, false, MODE>(L, lds);
  else if (hb) launch_mirror<T, false, true, MODE>(L, lds);
  else launch_mirror<T, false, false, MODE>(L, lds);
}

template <typename T, int NUF>
static void launch_rows(BinLaunch& L, size_t lds) {
  // FP64 rows hold at most 8 slots in registers (16 would spill)
  constexpr int NU = (sizeof(T) == 8 && NUF > 8) ? 8 : NUF;
  constexpr bool PREF = sizeof(T) == 4 && NU <= 8;
  // the branch-free path covers linear k bins without 2-D bins
  const bool fast = !L.dev.logbins && !L.dev.out2d;
  const void* fn = fast ? (const void*)k_bin_rows<T, NU, PREF, true> : (const void*)k_bin_rows<T, NU, PREF, false>;
  if (lds > 64 * 1024) hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (fast)
    TIMED(L.timer, "bin",
          (k_bin_rows<T, NU, PREF, true><<<L.nwg, 64 * L.dev.copies, lds, L.stream>>>(
              (const T*)L.grid[0], (const T*)L.grid[1], (const T*)L.grid[2], (const T*)L.grid[3],
              L.tabs, L.dev, L.partials)));
  else
    TIMED(L.timer, "bin",
          (k_bin_rows<T, NU, PREF, false><<<L.nwg, 64 * L.dev.copies, lds, L.stream>>>(
              (const T*)L.grid[0], (const T*)L.grid[1], (const T*)L.grid[2], (const T*)L.grid[3],
              L.tabs, L.dev, L.partials)));
}

// Mode counts of the branch-free forms, once per context (they depend on
// the geometry and the bins only).
__global__ void k_fill_zero(double* __restrict__ p, int n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0.0;
}

int run_bin_counts(BinLaunch& L) {
  if (*L.counts_valid) return CAPOW_OK;
  const int ncnt = L.dev.nk + (L.dev.out2d ? L.dev.nk * L.dev.nmu : 0);
  // cleared by a kernel (no memset node when a caller captures the stream)
  k_fill_zero<<<(ncnt + 255) / 256, 256, 0, L.stream>>>(L.counts, ncnt);
  const int64_t nrows = (int64_t)L.dev.nx * L.dev.ny;
  const int nblk = (int)std::min<int64_t>((nrows + 3) / 4, 4096);
  const size_t lds = (size_t)ncnt * sizeof(double);
  if (lds > 64 * 1024)
    hipFuncSetAttribute((const void*)k_bin_counts, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  k_bin_counts<<<nblk, 256, lds, L.stream>>>(L.tabs, L.dev, L.counts);
  if (hipGetLastError() != hipSuccess) return CAPOW_ERR_HIP;
  *L.counts_valid = true;
  return CAPOW_OK;
}

// Fixed-order sum of the L.nwg workgroup partials into L.out (counts from
// L.counts for the branch-free forms).
int run_bin_final(BinLaunch& L, bool fast) {
  const int nl = L.dev.odd ? 5 : 3;
  const int F = L.dev.nspec * nl;
  TIMED(L.timer, "bin_final",
        (k_bin_final<<<(L.per_copy + 255) / 256, 256, 0, L.stream>>>(
            L.partials, L.nwg, L.per_copy, fast ? L.counts : nullptr,
            F * L.dev.nk + L.dev.nspec * (L.dev.out2d ? L.dev.nk * L.dev.nmu : 0),
            L.dev.nk + (L.dev.out2d ? L.dev.nk * L.dev.nmu : 0), L.out)));
  return hipGetLastError() == hipSuccess ? CAPOW_OK : CAPOW_ERR_HIP;
}

template <typename T>
static int run_bin_t(BinLaunch& L) {
  const int per_copy = L.per_copy;
  const size_t bins = (size_t)per_copy * L.dev.copies * sizeof(double);
  const int nuf = L.dev.nzh / 64;
  // branch-free row form (linear bins) and mirror-group form (linear or log
  // bins): LDS copies of the sums only, mode counts from k_bin_counts
  const bool fast = L.row_form && !L.dev.logbins && !L.dev.out2d;
  const bool mirror = L.row_form && L.mirror && !L.dev.out2d;
  // (k, mu) bins in the mirror form: the 1-D pass also sums the per-k bounds
  const bool mirror2d = L.row_form && L.mirror2d;
  const bool sums = fast || mirror || mirror2d;
  const int nl = L.dev.odd ? 5 : 3;
  const int F = L.dev.nspec * nl;
  const size_t lcp = sums ? (size_t)F * L.dev.nk + (mirror2d ? L.dev.nk : 0) : (size_t)per_copy;
  const size_t lds = lcp * L.dev.copies * sizeof(double) + (size_t)L.dev.nzh * 4 * sizeof(double);
  if (sums) {
    const int rc = run_bin_counts(L);
    if (rc) return rc;
  }
  if (mirror2d) {
    const int n2d = L.dev.nk * L.dev.nmu;
    const size_t lds2 = (size_t)L.dev.nspec * n2d * 8 + (size_t)L.dev.nzh * 4 * sizeof(double);
    if (lds > 160 * 1024 || lds2 > 160 * 1024) return CAPOW_ERR_UNSUPPORTED;
    launch_mirror_any<T, 1>(L, lds);
    int rc = run_bin_final(L, true);
    if (rc) return rc;
    // the bounds sit where the (k, mu) sums go
    double* sums2d = L.out + (size_t)F * L.dev.nk;
    k_bin_scale<<<(L.dev.nk + 255) / 256, 256, 0, L.stream>>>(sums2d, L.dev.nk, L.qscale);
    launch_mirror_any<T, 2>(L, lds2);
    const int len = L.dev.nspec * n2d;
    TIMED(L.timer, "bin_final",
          (k_bin_final2d<<<(len + 255) / 256, 256, 0, L.stream>>>(L.partials2, L.nwg, len, n2d,
                                                                 L.dev.nmu, L.qscale, sums2d)));
    return hipGetLastError() == hipSuccess ? CAPOW_OK : CAPOW_ERR_HIP;
  }
  if (mirror) {
    // the host sizes the copies of the sums to fit beside the z tables
    if (lds > 160 * 1024) return CAPOW_ERR_UNSUPPORTED;
    launch_mirror_any<T, 0>(L, lds);
  } else if (L.row_form && lds <= 160 * 1024) {
    // the largest instantiated full-slot count not above nzh / 64
    if (nuf >= 16 && sizeof(T) == 4) launch_rows<T, 16>(L, lds);
    else if (nuf >= 8) launch_rows<T, 8>(L, lds);
    else if (nuf >= 4) launch_rows<T, 4>(L, lds);
    else if (nuf >= 2) launch_rows<T, 2>(L, lds);
    else if (nuf >= 1) launch_rows<T, 1>(L, lds);
    else launch_rows<T, 0>(L, lds);
  } else {
    constexpr int U = sizeof(T) == 4 ? 8 : 4;
    if (bins > 64 * 1024)
      hipFuncSetAttribute((const void*)k_bin<T, U>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bins);
    TIMED(L.timer, "bin",
          (k_bin<T, U><<<L.nwg, 64 * L.dev.copies, bins, L.stream>>>(
              (const T*)L.grid[0], (const T*)L.grid[1], (const T*)L.grid[2], (const T*)L.grid[3],
              L.tabs, L.dev, L.partials)));
  }
  return run_bin_final(L, sums);
}

int run_bin(BinLaunch& L, bool fp64) { return fp64 ? run_bin_t<double>(L) : run_bin_t<float>(L); }

}  // namespace capow
