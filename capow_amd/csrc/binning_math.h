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

}  // namespace capow
