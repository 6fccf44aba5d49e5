// binning_math.h -- per-mode quantities of the periodic power binning
// (reference src/main.c:329-461), shared by the forward binning and its
// adjoint so that both see the same bins, mu, NM, deconvolution and phase.
#pragma once

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

// Per-row quantities of one (i, j) row of modes: frequencies, fx^2+fy^2,
// the x*y deconvolution and interlace phase products.
struct RowVals {
  int i, j;
  double fx, fy, fxy, corxy, qr, qi;
};

__device__ __forceinline__ RowVals row_vals(int i, int j, const double* __restrict__ tabs,
                                            const BinArgsDev& a) {
  RowVals R;
  R.i = i;
  R.j = j;
  R.fx = tabs[a.ofx + i];
  R.fy = tabs[a.ofy + j];
  R.fxy = dadd(dmul(R.fx, R.fx), dmul(R.fy, R.fy));
  R.corxy = tabs[a.ocx + i] * tabs[a.ocy + j];
  R.qr = 1.0;
  R.qi = 0.0;
  if (a.interlace) {
    const double pxr = tabs[a.opx + 2 * i], pxi = tabs[a.opx + 2 * i + 1];
    const double pyr = tabs[a.opy + 2 * j], pyi = tabs[a.opy + 2 * j + 1];
    R.qr = pxr * pyr - pxi * pyi;
    R.qi = pxr * pyi + pxi * pyr;
  }
  return R;
}

// Geometry of one mode (i, j, k): bin, mu, Hermitian weight NM, window
// deconvolution cor (x0.25 with interlacing) and interlace phase (cr, ci).
struct ModeGeom {
  bool ok;       // inside the k range and below the Nyquist frequency
  int kbin;
  double mu, NM, cor, cr, ci;
  double fktot;  // 2 pi |k|
  int b2;        // 2-D bin
  bool in2d;
};

__device__ __forceinline__ ModeGeom mode_geom(const RowVals& R, int k, double fz, double corz,
                                             double pzr, double pzi, const BinArgsDev& a) {
  ModeGeom G;
  G.ok = false;
  const int NK = a.nk;
  const double twopi = 2.0 * 3.14159265358979323846;  // the reference's 2.0*M_PI
  const int i = R.i, j = R.j;
  const double fktot = dmul(twopi, sqrt(dadd(R.fxy, dmul(fz, fz))));
  int kbin;
  if (a.logbins) {
    if (!(fktot > 0)) return G;
    kbin = (int)ddiv((double)__double2float_rn(log10(fktot) - a.mink), a.binwidth);
  } else {
    kbin = kbin_of((double)__double2float_rn(fktot - a.mink), a);
  }
  if (!((kbin >= 0) && (kbin < NK) && (fktot <= a.min_nyq))) return G;
  double mu = 0.0;
  if (fktot > 0) {
    const double fl = a.los == 1 ? R.fx : (a.los == 2 ? R.fy : fz);
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
  G.ok = true;
  G.kbin = kbin;
  G.fktot = fktot;
  G.mu = mu;
  G.NM = NM;
  G.cor = R.corxy * corz;  // x0.25 folded in corz
  G.cr = 1.0;
  G.ci = 0.0;
  if (a.interlace) {
    G.cr = R.qr * pzr - R.qi * pzi;
    G.ci = R.qr * pzi + R.qi * pzr;
  }
  G.b2 = 0;
  G.in2d = false;
  if (a.out2d) {
    int mubin = (int)(((mu + 1.0) / 2.0) * (double)(float)a.nmu);
    if (mubin >= a.nmu) mubin = a.nmu - 1;
    G.b2 = kbin * a.nmu + mubin;
    G.in2d = (G.b2 >= 0) && (G.b2 < NK * a.nmu);
  }
  return G;
}

// Branch-free part of one mode for the common binning (linear k bins, no
// 2-D bins): bin index (without the near-integer exact fallback), whether
// the mode is binned, NM * cor * power of each spectrum, L_2, L_4 and mu.  Written without early exits so that the
// compiler can interleave the long FP64 chains of several modes.
struct ModeVals {
  int kbin;
  bool ok, near;
  double np[3];
  double L2, L4, mu;
  double fktot;   // 2 pi |k| (group form)
  double bound;   // group form: NM cor (|A|^2 + |B|^2) summed over the rows,
                  // >= |np[s]| for every spectrum
};

// Linear k bin of a mode by multiplication with 1/binwidth; `near` flags a
// quotient within 1e-9 of an integer, where the product may differ from the
// reference's division and kbin_exact decides.
struct KBin {
  int kbin;
  bool ok, near;
  double fktot;
};

__device__ __forceinline__ KBin kbin_fast(double fxy, double fz, const BinArgsDev& a) {
  KBin r;
  r.fktot = dmul(2.0 * 3.14159265358979323846, sqrt(dadd(fxy, dmul(fz, fz))));
  const double q = (double)__double2float_rn(r.fktot - a.mink) * a.inv_binwidth;
  const double tq = trunc(q);
  const double fr = q - tq;
  r.near = fabs(fr) < 1e-9 || fabs(fr) > 1.0 - 1e-9;
  r.kbin = (int)tq;
  r.ok = (r.kbin >= 0) && (r.kbin < a.nk) && (r.fktot <= a.min_nyq);
  return r;
}

// The reference's division (src/main.c:372, 375) for the near-integer case.
__device__ __forceinline__ KBin kbin_exact(double fxy, double fz, const BinArgsDev& a) {
  KBin r;
  r.fktot = dmul(2.0 * 3.14159265358979323846, sqrt(dadd(fxy, dmul(fz, fz))));
  r.kbin = (int)ddiv((double)__double2float_rn(r.fktot - a.mink), a.binwidth);
  r.ok = (r.fktot <= a.min_nyq) && r.kbin >= 0 && r.kbin < a.nk;
  r.near = false;
  return r;
}

// Log k bin (OutputLog, src/main.c:369-371): (int)((float)(log10|k| - Mink)
// / binwidth) with the exact division, as mode_geom computes it; k = 0 is not
// binned.
__device__ __forceinline__ KBin kbin_log(double fxy, double fz, const BinArgsDev& a) {
  KBin r;
  r.fktot = dmul(2.0 * 3.14159265358979323846, sqrt(dadd(fxy, dmul(fz, fz))));
  r.near = false;
  r.kbin = 0;
  r.ok = false;
  if (r.fktot > 0) {
    r.kbin = (int)ddiv((double)__double2float_rn(log10(r.fktot) - a.mink), a.binwidth);
    r.ok = (r.fktot <= a.min_nyq) && r.kbin >= 0 && r.kbin < a.nk;
  }
  return r;
}

// The bin of a mode in the branch-free forms: linear bins by multiplication
// (near-integer quotients flagged for kbin_exact), log bins exactly.
__device__ __forceinline__ KBin kbin_any(double fxy, double fz, const BinArgsDev& a) {
  return a.logbins ? kbin_log(fxy, fz, a) : kbin_fast(fxy, fz, a);
}

// The (k, mu) bin of a binned mode (Output2D, src/main.c:453-461): mu by the
// exact division (it feeds the mu bin), mu bin (int)((mu + 1)/2 * NMU) capped
// at NMU - 1, as mode_geom computes it.
__device__ __forceinline__ int mu_bin2d(int kbin, double fl, double fktot, const BinArgsDev& a,
                                        bool& in2d) {
  double mu = 0.0;
  if (fktot > 0) mu = ddiv(dmul(2.0 * 3.14159265358979323846, fl), fktot);
  int mubin = (int)(((mu + 1.0) / 2.0) * (double)(float)a.nmu);
  if (mubin >= a.nmu) mubin = a.nmu - 1;
  const int b2 = kbin * a.nmu + mubin;
  in2d = (b2 >= 0) && (b2 < a.nk * a.nmu);
  return b2;
}

// The reference's Hermitian weight NM (src/main.c:422-429): 2, except 1 in
// the kz = 0 plane where a mode's partner (-i, -j, 0) is stored as well
// (the rows i = 0, NX/2 of j = 0 and i = NX/2 keep 2, as there).
__device__ __forceinline__ double nm_weight(int i, int j, int k, int nx) {
  const bool nm1 = k == 0 && ((j == 0) ? ((i != 0) && (i != nx / 2)) : (i != nx / 2));
  return nm1 ? 1.0 : 2.0;
}

// Accumulates one evaluated mode into this wave's bins (the mode counts
// come from k_bin_counts).
__device__ __forceinline__ void mode_add(const ModeVals& V, const BinArgsDev& a, double* accS,
                                         int nl) {
  const int NK = a.nk;
#pragma unroll
  for (int s = 0; s < 3; s++) {
    if (s >= a.nspec) break;
    const double np = V.np[s];
    double* b = accS + s * nl * NK + V.kbin;
    atomicAdd(&b[0], np);
    atomicAdd(&b[NK], V.L2 * np);
    atomicAdd(&b[2 * NK], V.L4 * np);
    if (a.odd) {
      const double L3 = 2.5 * V.mu * V.mu * V.mu - 1.5 * V.mu;
      atomicAdd(&b[3 * NK], V.mu * np);
      atomicAdd(&b[4 * NK], L3 * np);
    }
  }
}

// ---- Mirror-group form of the branch-free binning ----
//
// With f(N - i) = -f(i) exactly in the frequency tables and the deconvolution
// even in f (both checked on the host), the rows (i, j), (NX - i, j),
// (i, NY - j) and (NX - i, NY - j) share |k|, the k bin, the window
// correction and the Hermitian weight NM of every kz (NX, NY even); mu agrees
// up to its sign, so L_2 and L_4 agree, and with the line of sight along z
// mu itself does.  Only the interlace phase differs from row to row.  A
// group's per-mode geometry is therefore evaluated once, its (up to) four
// powers are summed, and the sum takes one set of LDS adds.

// The rows of one group: canonical (i, j), 0 <= i <= NX/2, 0 <= j <= NY/2,
// and its partners where they are distinct rows (live).  Dead slots repeat
// the canonical row and are read as zero.
struct Group {
  int i, j;
  double fx, fy, fxy, corxy;
  int64_t row[4];
  bool live[4];
  double qr[4], qi[4];   // interlace phase of each row (x phase * y phase)
};

__device__ __forceinline__ void group_rows(int g, int hy1, const BinArgsDev& a, int& i, int& j,
                                           int64_t* row, bool* live) {
  i = g / hy1;
  j = g - i * hy1;
  const bool mx = i > 0 && 2 * i < a.nx, my = j > 0 && 2 * j < a.ny;
  const int i2 = mx ? a.nx - i : i, j2 = my ? a.ny - j : j;
  row[0] = (int64_t)i * a.ny + j;
  row[1] = (int64_t)i2 * a.ny + j;
  row[2] = (int64_t)i * a.ny + j2;
  row[3] = (int64_t)i2 * a.ny + j2;
  live[0] = true;
  live[1] = mx;
  live[2] = my;
  live[3] = mx && my;
}

__device__ __forceinline__ Group group_of(int g, int hy1, const double* __restrict__ tabs,
                                          const BinArgsDev& a) {
  Group G;
  group_rows(g, hy1, a, G.i, G.j, G.row, G.live);
  G.fx = tabs[a.ofx + G.i];
  G.fy = tabs[a.ofy + G.j];
  G.fxy = dadd(dmul(G.fx, G.fx), dmul(G.fy, G.fy));
  G.corxy = tabs[a.ocx + G.i] * tabs[a.ocy + G.j];
#pragma unroll
  for (int r = 0; r < 4; r++) {
    G.qr[r] = 1.0;
    G.qi[r] = 0.0;
    if (a.interlace) {
      const int ii = (int)(G.row[r] / a.ny), jj = (int)(G.row[r] - (int64_t)ii * a.ny);
      const double pxr = tabs[a.opx + 2 * ii], pxi = tabs[a.opx + 2 * ii + 1];
      const double pyr = tabs[a.opy + 2 * jj], pyi = tabs[a.opy + 2 * jj + 1];
      G.qr[r] = pxr * pyr - pxi * pyi;
      G.qi[r] = pxr * pyi + pxi * pyr;
    }
  }
  return G;
}

// One mode kz of a group: v[r][0..3] = A, A_int, B, B_int of row r (zero for
// dead rows and absent grids).  np[s] = NM cor * (sum over the rows of the
// power of spectrum s).  IL (interlacing) and HB (B grids) are the
// context's flags as template arguments; the four products |A|^2, |B|^2,
// Im(A* B), Re(A* B) are summed over the rows and mapped onto the spectra
// once per group.
// F32: the interlace combine and the four products in FP32 (FP32 grids;
// the geometry, NM cor and L_l stay in FP64): the fused x-pass binning.
template <typename C, bool IL, bool HB, bool F32 = false>
__device__ __forceinline__ ModeVals group_mode(const Group& G, int k, double fz, double corz,
                                               double pzr, double pzi, const C (&v)[4][4],
                                               const BinArgsDev& a) {
  ModeVals V;
  const double twopi = 2.0 * 3.14159265358979323846;
  const KBin kb = kbin_any(G.fxy, fz, a);
  const double fktot = kb.fktot;
  V.fktot = fktot;
  V.near = kb.near;
  V.kbin = kb.kbin;
  V.ok = kb.ok;
  const double fl = a.los == 1 ? G.fx : (a.los == 2 ? G.fy : fz);
  double rk = __builtin_amdgcn_rcp(fktot > 0 ? fktot : 1.0);
  rk = rk * (2.0 - fktot * rk);
  rk = rk * (2.0 - fktot * rk);
  const double mu = fktot > 0 ? (twopi * fl) * rk : 0.0;
  const double wc = nm_weight(G.i, G.j, k, a.nx) * (G.corxy * corz);
  double saa = 0.0, sbb = 0.0, sim = 0.0, sre = 0.0;
  if constexpr (F32) {
    float faa = 0.f, fbb = 0.f, fim = 0.f, fre = 0.f;
#pragma unroll
    for (int r = 0; r < 4; r++) {
      float ar = (float)v[r][0].x, ai = (float)v[r][0].y, br = 0.f, bi = 0.f;
      float cr = 1.f, ci = 0.f;
      if constexpr (IL) {
        cr = (float)(G.qr[r] * pzr - G.qi[r] * pzi);
        ci = (float)(G.qr[r] * pzi + G.qi[r] * pzr);
        ar += (float)v[r][1].x * cr - (float)v[r][1].y * ci;
        ai += (float)v[r][1].x * ci + (float)v[r][1].y * cr;
      }
      faa += ar * ar + ai * ai;
      if constexpr (HB) {
        br = (float)v[r][2].x;
        bi = (float)v[r][2].y;
        if constexpr (IL) {
          br += (float)v[r][3].x * cr - (float)v[r][3].y * ci;
          bi += (float)v[r][3].x * ci + (float)v[r][3].y * cr;
        }
        fbb += br * br + bi * bi;
        fim += ar * bi - ai * br;
        fre += ar * br + ai * bi;
      }
    }
    saa = faa;
    sbb = fbb;
    sim = fim;
    sre = fre;
  } else {
#pragma unroll
  for (int r = 0; r < 4; r++) {
    double ar = (double)v[r][0].x, ai = (double)v[r][0].y, br = 0, bi = 0;
    double cr = 1.0, ci = 0.0;
    if constexpr (IL) {
      cr = G.qr[r] * pzr - G.qi[r] * pzi;
      ci = G.qr[r] * pzi + G.qi[r] * pzr;
      ar += (double)v[r][1].x * cr - (double)v[r][1].y * ci;
      ai += (double)v[r][1].x * ci + (double)v[r][1].y * cr;
    }
    saa += ar * ar + ai * ai;
    if constexpr (HB) {
      br = (double)v[r][2].x;
      bi = (double)v[r][2].y;
      if constexpr (IL) {
        br += (double)v[r][3].x * cr - (double)v[r][3].y * ci;
        bi += (double)v[r][3].x * ci + (double)v[r][3].y * cr;
      }
      sbb += br * br + bi * bi;
      sim += ar * bi - ai * br;
      sre += ar * br + ai * bi;
    }
  }
  }
#pragma unroll
  for (int s = 0; s < 3; s++) {
    double p;
    switch (a.kind[s]) {
      case SPEC_AUTO_A: p = saa; break;
      case SPEC_AUTO_B: p = sbb; break;
      case SPEC_CROSS_IM: p = sim; break;
      default: p = sre; break;
    }
    V.np[s] = wc * p;
  }
  // |Im or Re(A* B)| <= |A||B| <= (|A|^2 + |B|^2) / 2 per row
  V.bound = wc * (saa + sbb);
  const double mu2 = mu * mu;
  V.mu = mu;
  V.L2 = 1.5 * mu2 - 0.5;
  V.L4 = 4.375 * mu2 * mu2 - 3.75 * mu2 + 0.375;
  return V;
}

}  // namespace capow
