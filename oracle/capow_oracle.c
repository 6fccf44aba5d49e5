/*
 * capow_oracle.c -- FP64 CPU oracle of the periodic P_l(k) path (NTask = 1).
 *
 * TEST INFRASTRUCTURE ONLY (see capow_oracle.h).  Built with
 * -ffp-contract=off so that every double expression rounds exactly as the
 * reference built by gcc for baseline x86-64 (no FMA instructions).
 */
#include "capow_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* Padded row length 2(NZ/2+1) and flat index, reference src/read_data.c:536
 * (iz + 2*(NZ/2+1)*(iy + NY*ix)). */
#define ROWLEN(p) (2ULL * ((unsigned long long)(p)->nz / 2ULL + 1ULL))
#define IDX(p, ix, iy, iz)                                                   \
  ((unsigned long long)(long long)(iz) +                                     \
   ROWLEN(p) * ((unsigned long long)(long long)(iy) +                        \
                (unsigned long long)(p)->ny * (unsigned long long)(long long)(ix)))

long long orc_grid_size(const orc_params *p) {
  /* Total_size = 2*alloc_local + InterpOrder*alloc_slice, with
   * alloc_local = NX*NY*(NZ/2+1) at one rank (src/main.c:1131-1135). */
  unsigned long long slice = 2ULL * (unsigned long long)p->ny *
                             ((unsigned long long)p->nz / 2ULL + 1ULL);
  return (long long)(((unsigned long long)p->nx + (unsigned long long)p->interp_order) * slice);
}

static double dxof(const orc_params *p) { return (p->xmax - p->xmin) / (double)(unsigned long long)p->nx; }
static double dyof(const orc_params *p) { return (p->ymax - p->ymin) / (double)(unsigned long long)p->ny; }
static double dzof(const orc_params *p) { return (p->zmax - p->zmin) / (double)(unsigned long long)p->nz; }

/* Reference src/read_data.c:514-745, specialised to NTask = 1, ThisTask = 0,
 * Periodic = 1: xmin = XMIN, xmax = XMAX (src/read_data.c:37-38 with
 * Local_x_start = 0, Local_nx = NX) and nx = Local_nxtra = NX + InterpOrder
 * (src/main.c:1134).  The statement order and guards are kept as in the
 * reference because the edge behaviour (dropped low taps at index 0, the
 * interlaced upper-edge rejections) is part of what is being pinned. */
double orc_add_to_grid(const orc_params *p, double x, double y, double z,
                       double w, double vr, double *density) {
  const unsigned long long NX = (unsigned long long)p->nx;
  const unsigned long long NY = (unsigned long long)p->ny;
  const unsigned long long NZ = (unsigned long long)p->nz;
  const double dx = dxof(p), dy = dyof(p), dz = dzof(p);
  const double xmin = p->xmin, xmax = p->xmax;
  const int nx = (int)(NX + (unsigned long long)p->interp_order);
  const int LOS = p->los;
  double wvr = w * vr;

  if (p->interp_order == 1) {
    /* src/read_data.c:519-541; ThisTask == NTask-1 branch. */
    if ((x < xmin) || (x >= xmax)) return 0.0;
    if ((y < p->ymin) || (y >= p->ymax)) return 0.0;
    if ((z < p->zmin) || (z >= p->zmax)) return 0.0;
    int ix = (int)(((double)x - xmin) / dx);
    int iy = (int)(((double)y - p->ymin) / dy);
    int iz = (int)(((double)z - p->zmin) / dz);
    if (ix == nx) ix = nx - 1;
    if ((unsigned long long)iy == NY) iy = (int)NY - 1;
    if ((unsigned long long)iz == NZ) iz = (int)NZ - 1;
    unsigned long long ind = IDX(p, ix, iy, iz);
    if (ind >= (unsigned long long)orc_grid_size(p)) return -1.0; /* FatalError 133 */
    density[ind] += wvr;

  } else if (p->interp_order == 2) {
    /* src/read_data.c:544-605 */
    if ((x < xmin) || (x >= xmax)) return 0.0;
    if ((y < p->ymin) || (y >= p->ymax)) return 0.0;
    if ((z < p->zmin) || (z >= p->zmax)) return 0.0;
    double scalex = (x - xmin) / dx;
    double scaley = (y - p->ymin) / dy;
    double scalez = (z - p->zmin) / dz;
    int ix = (int)scalex;
    int iy = (int)scaley;
    int iz = (int)scalez;
    double idx = scalex - (double)ix;
    double idy = scaley - (double)iy;
    double idz = scalez - (double)iz;
    double itx = 1.0 - idx;
    double ity = 1.0 - idy;
    double itz = 1.0 - idz;
    int ixneigh = ix + 1;
    int iyneigh = iy + 1;
    int izneigh = iz + 1;
    if (ixneigh == nx) ixneigh = nx - 1;
    if (LOS == 0) {
      if ((unsigned long long)iyneigh == NY) iyneigh = 0;
      if ((unsigned long long)izneigh == NZ) izneigh = 0;
    }
    density[IDX(p, ix, iy, iz)] += wvr * itx * ity * itz;
    density[IDX(p, ixneigh, iy, iz)] += wvr * idx * ity * itz;
    if ((unsigned long long)iyneigh < NY) {
      density[IDX(p, ix, iyneigh, iz)] += wvr * itx * idy * itz;
      density[IDX(p, ixneigh, iyneigh, iz)] += wvr * idx * idy * itz;
      if ((unsigned long long)izneigh < NZ) {
        density[IDX(p, ix, iyneigh, izneigh)] += wvr * itx * idy * idz;
        density[IDX(p, ixneigh, iyneigh, izneigh)] += wvr * idx * idy * idz;
      }
    }
    if ((unsigned long long)izneigh < NZ) {
      density[IDX(p, ix, iy, izneigh)] += wvr * itx * ity * idz;
      density[IDX(p, ixneigh, iy, izneigh)] += wvr * idx * ity * idz;
    }

  } else if (p->interp_order == 3) {
    /* src/read_data.c:608-743 */
    if (LOS == 0) {
      if (x - p->xmin < dx) x += p->xmax - p->xmin;
      if ((x < xmin + dx) || (x >= xmax + dx)) return 0.0;
    } else {
      if ((x < xmin) || (x >= xmax + dx)) return 0.0;   /* ThisTask == 0 */
    }
    if ((y < p->ymin) || (y >= p->ymax)) return 0.0;
    if ((z < p->zmin) || (z >= p->zmax)) return 0.0;

    double scalex = (x - xmin) / dx;
    double scaley = (y - p->ymin) / dy;
    double scalez = (z - p->zmin) / dz;
    int ix = (int)(scalex + 0.5);
    int iy = (int)(scaley + 0.5);
    int iz = (int)(scalez + 0.5);
    double idx = scalex - (double)ix;
    double idy = scaley - (double)iy;
    double idz = scalez - (double)iz;
    double itx = 0.75 - idx * idx;
    double ity = 0.75 - idy * idy;
    double itz = 0.75 - idz * idz;
    double isx = 0.5 * (0.5 - idx) * (0.5 - idx);
    double isy = 0.5 * (0.5 - idy) * (0.5 - idy);
    double isz = 0.5 * (0.5 - idz) * (0.5 - idz);
    idx = 0.5 * (0.5 + idx) * (0.5 + idx);
    idy = 0.5 * (0.5 + idy) * (0.5 + idy);
    idz = 0.5 * (0.5 + idz) * (0.5 + idz);

    int ixl = ix - 1, iyl = iy - 1, izl = iz - 1;
    int ixh = ix + 1, iyh = iy + 1, izh = iz + 1;
    const int ny = (int)NY, nz = (int)NZ;

    if (ixh == nx) ixh = nx - 1;
    if (LOS == 0) {
      if (iyl < 0) iyl = ny - 1;
      if (izl < 0) izl = nz - 1;
      if (iyh >= ny) {
        iyh -= ny;
        if (iy >= ny) iy = 0;
        if (iyh > 1) iyh = 1;
      }
      if (izh >= nz) {
        izh -= nz;
        if (iz >= nz) iz = 0;
        if (izh > 1) izh = 1;
      }
    }

    /* The 27 taps with the reference's guards (src/read_data.c:690-742). */
#define T(a, b, c, f1, f2, f3) density[IDX(p, a, b, c)] += wvr * f1 * f2 * f3
    if ((iy < ny) && (iz < nz)) {
      T(ix, iy, iz, itx, ity, itz);
      if (ixl > 0) T(ixl, iy, iz, isx, ity, itz);
      if (ixh < nx) T(ixh, iy, iz, idx, ity, itz);
    }
    if (iyl > 0) {
      if (iz < nz) {
        T(ix, iyl, iz, itx, isy, itz);
        if (ixl > 0) T(ixl, iyl, iz, isx, isy, itz);
        if (ixh < nx) T(ixh, iyl, iz, idx, isy, itz);
      }
      if (izl > 0) {
        T(ix, iyl, izl, itx, isy, isz);
        if (ixl > 0) T(ixl, iyl, izl, isx, isy, isz);
        if (ixh < nx) T(ixh, iyl, izl, idx, isy, isz);
      }
      if (izh < nz) {
        T(ix, iyl, izh, itx, isy, idz);
        if (ixl > 0) T(ixl, iyl, izh, isx, isy, idz);
        if (ixh < nx) T(ixh, iyl, izh, idx, isy, idz);
      }
    }
    if (iyh < ny) {
      if (iz < nz) {
        T(ix, iyh, iz, itx, idy, itz);
        if (ixl > 0) T(ixl, iyh, iz, isx, idy, itz);
        if (ixh < nx) T(ixh, iyh, iz, idx, idy, itz);
      }
      if (izl > 0) {
        T(ix, iyh, izl, itx, idy, isz);
        if (ixl > 0) T(ixl, iyh, izl, isx, idy, isz);
        if (ixh < nx) T(ixh, iyh, izl, idx, idy, isz);
      }
      if (izh < nz) {
        T(ix, iyh, izh, itx, idy, idz);
        if (ixl > 0) T(ixl, iyh, izh, isx, idy, idz);
        if (ixh < nx) T(ixh, iyh, izh, idx, idy, idz);
      }
    }
    if (izl > 0) {
      if (iy < ny) {
        T(ix, iy, izl, itx, ity, isz);
        if (ixl > 0) T(ixl, iy, izl, isx, ity, isz);
        if (ixh < nx) T(ixh, iy, izl, idx, ity, isz);
      }
    }
    if (izh < nz) {
      if (iy < ny) {
        T(ix, iy, izh, itx, ity, idz);
        if (ixl > 0) T(ixl, iy, izh, isx, ity, idz);
        if (ixh < nx) T(ixh, iy, izh, idx, ity, idz);
      }
    }
#undef T
  }
  return w;
}

/* Halo add at NTask = 1 (src/read_data.c:152-175 with the neighbour rule of
 * src/main.c:1146-1189): for LOS = 0 the rank is its own left and right
 * neighbour, so the InterpOrder buffer slices are added onto slices
 * 0..InterpOrder-1; for LOS != 0 the neighbour is MPI_PROC_NULL and the
 * buffer is dropped. */
static void halo_add(const orc_params *p, double *g) {
  if (!g || p->los != 0) return;
  unsigned long long slice = 2ULL * (unsigned long long)p->ny * ((unsigned long long)p->nz / 2ULL + 1ULL);
  unsigned long long last = (unsigned long long)p->nx * slice;
  unsigned long long cnt = (unsigned long long)p->interp_order * slice;
  double *tmp = (double *)malloc(cnt * sizeof(double));
  memcpy(tmp, g + last, cnt * sizeof(double));
  for (unsigned long long i = 0; i < cnt; i++) g[i] += tmp[i];
  free(tmp);
}

int orc_assign(const orc_params *p, const double *pos, const double *vel,
               const double *wts, long long n, double **grids, orc_stats *st) {
  const double dx = dxof(p), dy = dyof(p), dz = dzof(p);
  double *ddg = grids[0], *ddgi = grids[1], *mom = grids[2], *momi = grids[3];
  const int M = p->momentum, I = p->do_interlacing;
  memset(st, 0, sizeof(*st));
  /* src/read_data.c:56-139 */
  for (long long q = 0; q < n; q++) {
    double w = wts ? wts[q] : 1.0;
    double tx = pos[3 * q], ty = pos[3 * q + 1], tz = pos[3 * q + 2];
    double tvx = vel[3 * q], tvy = vel[3 * q + 1], tvz = vel[3 * q + 2], tvr;
    if (p->los == 0) {
      tvr = tvz;
    } else if (p->los == 1) {
      tvr = tvx;
      tx += tvr * (1.0 + p->redshift) / (100.0 * sqrt(p->omega_m * pow((1.0 + p->redshift), 3.0) + 1.0 - p->omega_m));
    } else if (p->los == 2) {
      tvr = tvy;
      ty += tvr * (1.0 + p->redshift) / (100.0 * sqrt(p->omega_m * pow((1.0 + p->redshift), 3.0) + 1.0 - p->omega_m));
    } else {
      tvr = tvz;
      tz += tvr * (1.0 + p->redshift) / (100.0 * sqrt(p->omega_m * pow((1.0 + p->redshift), 3.0) + 1.0 - p->omega_m));
    }
    st->nread += 1;
    if ((tx < p->xmin) || (tx > p->xmax) || (ty < p->ymin) || (ty > p->ymax) ||
        (tz < p->zmin) || (tz > p->zmax)) {
      if (p->los == 0) return 110;
      st->nout += 1;
    }
    double xi = tx + dx / 2.0, yi = ty + dy / 2.0, zi = tz + dz / 2.0;
    /* Field values per Momentum (src/read_data.c:91-130). */
    double fa, fb = 0.0;
    int has_b;
    double v2 = tvr * tvr;
    switch (M) {
      case 0: fa = 1.0; has_b = 0; break;
      case 1: fa = 1.0; fb = tvr; has_b = 1; break;
      case 2: fa = tvr; has_b = 0; break;
      case 3: fa = 1.0; fb = v2; has_b = 1; break;
      case 4: fa = tvr; fb = v2; has_b = 1; break;
      case 5: fa = v2; has_b = 0; break;
      case 6: fa = 1.0; fb = tvr * tvr * tvr; has_b = 1; break;
      case 7: fa = tvr; fb = tvr * tvr * tvr; has_b = 1; break;
      default: fa = 1.0; fb = tvr * tvr * tvr * tvr; has_b = 1; break;
    }
    double acc = orc_add_to_grid(p, tx, ty, tz, w, fa, ddg);
    if (acc < 0) return 133;
    st->ngrid += acc;
    if (I) { if (orc_add_to_grid(p, xi, yi, zi, w, fa, ddgi) < 0) return 133; }
    if (has_b) {
      double a2 = orc_add_to_grid(p, tx, ty, tz, w, fb, mom);
      if (a2 < 0) return 133;
      st->npv += a2;
      if (I) { if (orc_add_to_grid(p, xi, yi, zi, w, fb, momi) < 0) return 133; }
    }
    if (M) {
      st->nsq += w;
      st->vr += w * tvr;
      st->vr2 += w * tvr * tvr;
      st->vr3 += w * tvr * tvr * tvr;
      st->vr4 += w * tvr * tvr * tvr * tvr;
    }
  }
  /* src/read_data.c:152-175 */
  halo_add(p, ddg);
  if (I) halo_add(p, ddgi);
  if ((M != 0) && (M != 2) && (M != 5)) {
    halo_add(p, mom);
    if (I) halo_add(p, momi);
  }
  return 0;
}

void orc_scalars(const orc_params *p, const orc_stats *st, double *alpha,
                 double *shot, double *norm) {
  /* src/main.c:106-148 at NTask = 1 */
  double N = st->ngrid;
  double NSQ = 0, VR = 0, VR2 = 0, VR3 = 0, VR4 = 0;
  if (p->momentum) {
    NSQ = st->nsq; VR = st->vr; VR2 = st->vr2; VR3 = st->vr3; VR4 = st->vr4;
  }
  *alpha = N / ((double)(unsigned long long)p->nx * (double)(unsigned long long)p->ny *
                (double)(unsigned long long)p->nz);
  double nbar = N / ((p->xmax - p->xmin) * (p->ymax - p->ymin) * (p->zmax - p->zmin));
  *shot = N;
  *norm = nbar * N;
  if (p->momentum) {
    NSQ /= 1.0;
    VR /= (NSQ * 1.0);
    VR2 /= (NSQ * 1.0);
    VR3 /= (NSQ * 1.0);
    VR4 /= (NSQ * 1.0);
  }
  switch (p->momentum) {
    case 1: *shot = VR * N; break;
    case 2: *shot = VR2 * N; break;
    case 3: *shot = VR2 * N; break;
    case 4: *shot = VR3 * N; break;
    case 5: *shot = VR4 * N; break;
    case 6: *shot = VR3 * N; break;
    case 7: *shot = VR4 * N; break;
    case 8: *shot = VR4 * N; break;
    default: break;
  }
}

void orc_mean_subtract(const orc_params *p, double alpha, double **grids) {
  /* src/main.c:214-225 */
  int M = p->momentum;
  if (!((M == 0) | (M == 1) | (M == 3) | (M == 6) | (M == 8))) return;
  for (unsigned long long i = 0; i < (unsigned long long)p->nx; i++)
    for (unsigned long long j = 0; j < (unsigned long long)p->ny; j++)
      for (unsigned long long k = 0; k < (unsigned long long)p->nz; k++) {
        unsigned long long ind = k + ROWLEN(p) * (j + (unsigned long long)p->ny * i);
        grids[0][ind] -= alpha;
        if (p->do_interlacing) grids[1][ind] -= alpha;
      }
}

/* src/main.c:302-468 at Local_x_start = 0, Local_nx = NX. */
void orc_compute_power(const orc_params *p, double **grids, orc_bins *b) {
  const unsigned long long NX = (unsigned long long)p->nx;
  const unsigned long long NY = (unsigned long long)p->ny;
  const unsigned long long NZ = (unsigned long long)p->nz;
  const double dx = dxof(p), dy = dyof(p), dz = dzof(p);
  const int M = p->momentum, LOS = p->los, NK = p->nk, NMU = p->nmu;
  const double *ddg = grids[0], *ddgi = grids[1], *mom = grids[2], *momi = grids[3];
  double binwidth = (p->maxk - p->mink) / (double)NK;

  double ny_x = (M_PI / dx);
  double ny_y = (M_PI / dy);
  double ny_z = (M_PI / dz);
  double min_nyquist = ny_x;
  if (ny_y < min_nyquist) min_nyquist = ny_y;
  if (ny_z < min_nyquist) min_nyquist = ny_z;

  int i, j, k;
  double sx = 1.0 / ((float)NX * dx), sy = 1.0 / ((float)NY * dy), sz = 1.0 / ((float)NZ * dz);
  for (i = 0; (unsigned long long)i < NX; i++) {
    for (j = 0; (unsigned long long)j < NY; j++) {
      for (k = 0; (unsigned long long)k <= NZ / 2; k++) {
        double fx;
        if ((unsigned long long)i <= NX / 2) {
          fx = (float)i * sx;
        } else {
          fx = ((float)i - (float)NX) * sx;
        }
        double fy;
        if ((unsigned long long)j <= NY / 2) {
          fy = (float)j * sy;
        } else {
          fy = ((float)j - (float)NY) * sy;
        }
        double fz = (float)k * sz;

        double fktot = 2.0 * M_PI * sqrt(fx * fx + fy * fy + fz * fz);
        double mu;
        if (fktot > 0) {
          if (LOS == 1) {
            mu = 2.0 * M_PI * fx / fktot;
          } else if (LOS == 2) {
            mu = 2.0 * M_PI * fy / fktot;
          } else {
            mu = 2.0 * M_PI * fz / fktot;
          }
        } else {
          mu = 0.0;
        }
        int kbin;
        if (p->output_log) {
          double lk = log10(fktot);
          /* (int) of -inf is undefined in C; the reference's x86 build
           * yields INT_MIN, i.e. rejected.  Reject explicitly. */
          if (!(lk > -1e308)) continue;
          kbin = (int)((float)(lk - p->mink) / binwidth);
        } else {
          kbin = (int)((float)(fktot - p->mink) / binwidth);
        }

        if ((kbin >= 0) && (kbin < NK) && (fktot <= min_nyquist)) {
          double sinc_x = 1.0, sinc_y = 1.0, sinc_z = 1.0;
          double ax = M_PI * fx * dx;
          double ay = M_PI * fy * dy;
          double az = M_PI * fz * dz;
          if (fx != 0.0) sinc_x = sin(ax) / ax;
          if (fy != 0.0) sinc_y = sin(ay) / ay;
          if (fz != 0.0) sinc_z = sin(az) / az;
          double grid_cor = pow(1.0 / (sinc_x * sinc_y * sinc_z), 2 * p->interp_order);

          unsigned long long base = ROWLEN(p) * ((unsigned long long)j + NY * (unsigned long long)i);
          double dkr = ddg[2 * k + base];
          double dki = ddg[2 * k + 1 + base];
          if (p->do_interlacing) {
            grid_cor *= 0.25;
            double kh = ax + ay + az;
            double sink = sin(kh);
            double cosk = cos(kh);
            double dkr_i = ddgi[2 * k + base];
            double dki_i = ddgi[2 * k + 1 + base];
            dkr += dkr_i * cosk - dki_i * sink;
            dki += dkr_i * sink + dki_i * cosk;
          }
          double dkr_mom = 0.0, dki_mom = 0.0;
          if ((M == 1) | (M == 3) | (M == 4) | (M >= 6)) {
            dkr_mom = mom[2 * k + base];
            dki_mom = mom[2 * k + 1 + base];
            if (p->do_interlacing) {
              double kh = ax + ay + az;
              double sink = sin(kh);
              double cosk = cos(kh);
              double r_i = momi[2 * k + base];
              double i_i = momi[2 * k + 1 + base];
              dkr_mom += r_i * cosk - i_i * sink;
              dki_mom += r_i * sink + i_i * cosk;
            }
          }
          int NM = 2;
          if (k == 0) {
            if (j == 0) {
              if ((i != 0) && ((unsigned long long)i != NX / 2)) NM = 1;
            } else {
              if ((unsigned long long)i != NX / 2) NM = 1;
            }
          }
          double power;
          double L2 = 1.5 * mu * mu - 0.5;
          double L4 = 4.375 * mu * mu * mu * mu - 3.75 * mu * mu + 0.375;
          if ((M == 0) | (M == 2) | (M == 5)) {
            power = (dkr * dkr + dki * dki) * grid_cor;
          } else if ((M == 1) | (M == 4) | (M == 6)) {
            power = (dkr * dki_mom - dki * dkr_mom) * grid_cor;
          } else {
            power = (dkr * dkr_mom + dki * dki_mom) * grid_cor;
          }
          b->pk0[kbin] += NM * power;
          b->pk2[kbin] += NM * L2 * power;
          b->pk4[kbin] += NM * L4 * power;
          if (p->odd_multipoles) {
            double L3 = 2.5 * mu * mu * mu - 1.5 * mu;
            b->pk1[kbin] += NM * mu * power;
            b->pk3[kbin] += NM * L3 * power;
          }
          b->nmodes[kbin] += NM;
          if (p->output_2d) {
            int mubin = (int)(((mu + 1.0) / 2.0) * (float)NMU);
            if (mubin >= NMU) mubin = NMU - 1;
            int bin = kbin * NMU + mubin;
            if ((bin >= 0) && (bin < NK * NMU)) {
              b->pk2d[bin] += NM * power;
              b->nmodes2d[bin] += NM;
            }
          }
        }
      }
    }
  }
}

int orc_finalize(const orc_params *p, double shot, double norm, orc_bins *b) {
  /* src/main.c:1014-1056 */
  const int NK = p->nk;
  int zeroed = -1;
  int *fill = (int *)calloc((size_t)NK, sizeof(int));
  for (int i = 0; i < NK; i++) {
    if (b->nmodes[i] > 0.0 && fill[i] == 0) {
      b->pk0[i] -= b->nmodes[i] * shot;
      b->pk0[i] /= b->nmodes[i] * norm;
      b->pk2[i] *= 5.0 / (b->nmodes[i] * norm);
      b->pk4[i] *= 9.0 / (b->nmodes[i] * norm);
      if (p->odd_multipoles) {
        b->pk1[i] *= 3.0 / (b->nmodes[i] * norm);
        b->pk3[i] *= 7.0 / (b->nmodes[i] * norm);
      }
      fill[i] = 1;
    }
  }
  if (p->output_2d) {
    for (int i = 0; i < NK * p->nmu; i++)
      if (b->nmodes2d[i] > 0.0) b->pk2d[i] /= b->nmodes2d[i] * norm;
  }
  for (int i = NK - 1; i >= 0; i--) {
    if (fill[i] == 1) {
      if (p->odd_multipoles) { b->pk1[i] = 0.0; b->pk3[i] = 0.0; }
      fill[i] = 0; b->pk0[i] = 0.0; b->pk2[i] = 0.0; b->pk4[i] = 0.0; b->nmodes[i] = 0;
      zeroed = i;
      break;
    }
  }
  free(fill);
  return zeroed;
}

int orc_write_output(const orc_params *p, double shot, double norm,
                     const orc_bins *b, const char *path, int precise) {
  /* src/main.c:1058-1101 */
  const char *g = precise ? "%.17g" : "%g";
  char fmt[256];
  double binwidth = (p->maxk - p->mink) / (double)p->nk;
  FILE *f = fopen(path, "w");
  if (!f) return -1;
  snprintf(fmt, sizeof fmt, "# shot = %s\n", g); fprintf(f, fmt, shot / norm);
  snprintf(fmt, sizeof fmt, "# norm = %s\n", g); fprintf(f, fmt, norm);
  if (p->odd_multipoles) fprintf(f, "# k  pk0  pk1  pk2  pk3  pk4  Nmodes\n");
  else fprintf(f, "# k  pk0  pk2  pk4  Nmodes\n");
  for (int i = 0; i < p->nk; i++) {
    double kp;
    if (p->output_log) kp = pow(10.0, p->mink + ((float)i + 0.5) * binwidth);
    else kp = p->mink + ((float)i + 0.5) * binwidth;
    if (p->odd_multipoles) {
      snprintf(fmt, sizeof fmt, "%s %s %s %s %s %s %%d\n", g, g, g, g, g, g);
      fprintf(f, fmt, kp, b->pk0[i], b->pk1[i], b->pk2[i], b->pk3[i], b->pk4[i], (int)b->nmodes[i]);
    } else {
      snprintf(fmt, sizeof fmt, "%s %s %s %s %%d\n", g, g, g, g);
      fprintf(f, fmt, kp, b->pk0[i], b->pk2[i], b->pk4[i], (int)b->nmodes[i]);
    }
  }
  fclose(f);
  if (p->output_2d) {
    char p2[4096];
    snprintf(p2, sizeof p2, "%s_2D", path);
    f = fopen(p2, "w");
    if (!f) return -1;
    snprintf(fmt, sizeof fmt, "# shot = %s\n", g); fprintf(f, fmt, shot / norm);
    snprintf(fmt, sizeof fmt, "# norm = %s\n", g); fprintf(f, fmt, norm);
    fprintf(f, "# k, mu, pk, Nmodes\n");
    snprintf(fmt, sizeof fmt, "%s %s %s %%d\n", g, g, g);
    for (int i = 0; i < p->nk; i++) {
      double kp;
      if (p->output_log) kp = pow(10.0, p->mink + ((float)i + 0.5) * binwidth);
      else kp = p->mink + ((float)i + 0.5) * binwidth;
      for (int j = 0; j < p->nmu; j++) {
        double mup = 2.0 * ((float)j + 0.5) / p->nmu - 1.0;
        fprintf(f, fmt, kp, mup, b->pk2d[i * p->nmu + j], (int)b->nmodes2d[i * p->nmu + j]);
      }
    }
    fclose(f);
  }
  return 0;
}

int orc_run(const orc_params *p, const double *pos, const double *vel,
            const double *w, long long n, orc_bins *b, orc_stats *st,
            double *alpha, double *shot, double *norm) {
  long long gs = orc_grid_size(p);
  int M = p->momentum, I = p->do_interlacing;
  int has_b = (M != 0) && (M != 2) && (M != 5);
  double *g[4] = {0, 0, 0, 0};
  g[0] = (double *)calloc((size_t)gs, sizeof(double));
  if (I) g[1] = (double *)calloc((size_t)gs, sizeof(double));
  if (has_b) g[2] = (double *)calloc((size_t)gs, sizeof(double));
  if (has_b && I) g[3] = (double *)calloc((size_t)gs, sizeof(double));
  int rc = orc_assign(p, pos, vel, w, n, g, st);
  if (rc == 0) {
    orc_scalars(p, st, alpha, shot, norm);
    orc_mean_subtract(p, *alpha, g);
    for (int q = 0; q < 4; q++)
      if (g[q]) orc_fft_r2c_3d(g[q], p->nx, p->ny, p->nz);
    orc_compute_power(p, g, b);
  }
  for (int q = 0; q < 4; q++) free(g[q]);
  return rc;
}

/* ---------------------------------------------------------------------------
 * Survey (Yamamoto) multipole estimator: reference src/main.c:472-776
 * (compute_survey_power) and src/main.c:781-967 (assign_survey_power), at
 * NTask = 1.  The grids hold the real-space fields before any transform; they
 * are not modified.  Work copies are weighted by products of the unit vector
 * r^ = (cell position - origin) / |.| and transformed; the monopole pass keeps
 * its k-space field (F0, F0_mom) for the cross products of the later passes.
 * --------------------------------------------------------------------------- */

/* One weighted copy of a real-space grid: dst = src * (product of the r^
 * components in idx[0..n-1]) on the real cells, as src/main.c:551-560
 * (dipole), 603-612 (quadrupole), 663-672 (octopole) and 734-743
 * (hexadecapole, whose first index appears twice). */
static void survey_weight(const orc_params *p, const double *origin, const double *src,
                          double *dst, int n, const int *idx) {
  const unsigned long long NX = (unsigned long long)p->nx, NY = (unsigned long long)p->ny,
                           NZ = (unsigned long long)p->nz;
  const double dx = dxof(p), dy = dyof(p), dz = dzof(p);
  for (unsigned long long ix = 0; ix < NX; ix++)
    for (unsigned long long iy = 0; iy < NY; iy++)
      for (unsigned long long iz = 0; iz < NZ; iz++) {
        const unsigned long long ind = iz + ROWLEN(p) * (iy + NY * ix);
        if (n == 0) {
          dst[ind] = src[ind];
          continue;
        }
        double vec[3] = {(double)ix * dx + p->xmin - origin[0], (double)iy * dy + p->ymin - origin[1],
                         (double)iz * dz + p->zmin - origin[2]};
        double r2 = vec[0] * vec[0] + vec[1] * vec[1] + vec[2] * vec[2];
        if (r2 == 0.0) r2 = 1.0;
        double v;
        if (n == 1) v = src[ind] * vec[idx[0]] / sqrt(r2);
        else if (n == 2) v = src[ind] * vec[idx[0]] * vec[idx[1]] / r2;
        else if (n == 3) v = src[ind] * vec[idx[0]] * vec[idx[1]] * vec[idx[2]] / (r2 * sqrt(r2));
        else v = src[ind] * vec[idx[0]] * vec[idx[0]] * vec[idx[1]] * vec[idx[2]] / (r2 * r2);
        dst[ind] = v;
      }
}

/* assign_survey_power (src/main.c:781-967) for one weighted, transformed
 * field.  w[0..3] = ddg_2, ddg_interlace_2, ddg_mom_2, ddg_mom_interlace_2. */
static void survey_assign(const orc_params *p, int multipole, int ii, int jj, int kk, double **w,
                          double *F0, double *F0m, orc_bins *b) {
  const unsigned long long NX = (unsigned long long)p->nx, NY = (unsigned long long)p->ny,
                           NZ = (unsigned long long)p->nz;
  const double dx = dxof(p), dy = dyof(p), dz = dzof(p);
  const int M = p->momentum, NK = p->nk;
  const int mom = (M != 0) && (M != 1);
  const double binwidth = (p->maxk - p->mink) / (double)NK;
  double min_nyquist = M_PI / dx;
  if (M_PI / dy < min_nyquist) min_nyquist = M_PI / dy;
  if (M_PI / dz < min_nyquist) min_nyquist = M_PI / dz;
  const double sx = 1.0 / ((float)NX * dx), sy = 1.0 / ((float)NY * dy), sz = 1.0 / ((float)NZ * dz);
  for (unsigned long long i = 0; i < NX; i++)
    for (unsigned long long j = 0; j < NY; j++)
      for (unsigned long long k = 0; k <= NZ / 2; k++) {
        const double fx = i <= NX / 2 ? (float)i * sx : ((float)i - (float)NX) * sx;
        const double fy = j <= NY / 2 ? (float)j * sy : ((float)j - (float)NY) * sy;
        const double fz = (float)k * sz;
        const double fktot = 2.0 * M_PI * sqrt(fx * fx + fy * fy + fz * fz);
        int kbin;
        if (p->output_log) {
          const double lk = log10(fktot);
          if (!(lk > -1e308)) continue;   /* (int) of -inf: rejected */
          kbin = (int)((float)(lk - p->mink) / binwidth);
        } else {
          kbin = (int)((float)(fktot - p->mink) / binwidth);
        }
        if (!((kbin >= 0) && (kbin < NK) && (fktot <= min_nyquist))) continue;
        double sinc_x = 1.0, sinc_y = 1.0, sinc_z = 1.0;
        const double ax = M_PI * fx * dx, ay = M_PI * fy * dy, az = M_PI * fz * dz;
        if (fx != 0.0) sinc_x = sin(ax) / ax;
        if (fy != 0.0) sinc_y = sin(ay) / ay;
        if (fz != 0.0) sinc_z = sin(az) / az;
        double grid_cor = pow(1.0 / (sinc_x * sinc_y * sinc_z), 2 * p->interp_order);
        const double kvec[3] = {2. * M_PI * fx, 2. * M_PI * fy, 2. * M_PI * fz};
        const unsigned long long ind = 2 * k + ROWLEN(p) * (j + NY * i);
        double dkr = w[0][ind], dki = w[0][ind + 1];
        const double kh = ax + ay + az, sink = sin(kh), cosk = cos(kh);
        if (p->do_interlacing) {
          grid_cor *= 0.25;
          dkr += w[1][ind] * cosk - w[1][ind + 1] * sink;
          dki += w[1][ind] * sink + w[1][ind + 1] * cosk;
        }
        double dkr_mom = 0.0, dki_mom = 0.0;
        if (mom) {
          dkr_mom = w[2][ind];
          dki_mom = w[2][ind + 1];
          if (p->do_interlacing) {
            dkr_mom += w[3][ind] * cosk - w[3][ind + 1] * sink;
            dki_mom += w[3][ind] * sink + w[3][ind + 1] * cosk;
          }
        }
        double power;
        if (!mom) {
          if (multipole == 0) {
            F0[ind] = dkr;
            F0[ind + 1] = dki;
            power = (dkr * dkr + dki * dki) * grid_cor;
          } else if (multipole % 2) {
            power = (dkr * F0[ind + 1] - dki * F0[ind]) * grid_cor;
          } else {
            power = (dkr * F0[ind] + dki * F0[ind + 1]) * grid_cor;
          }
        } else {
          if (multipole == 0) {
            F0[ind] = dkr;
            F0[ind + 1] = dki;
            F0m[ind] = dkr_mom;
            F0m[ind + 1] = dki_mom;
            power = (dkr * dki_mom - dki * dkr_mom) * grid_cor;
          } else if (multipole % 2) {
            power = 0.5 * (dkr * F0m[ind] + dki * F0m[ind + 1] + dkr_mom * F0[ind] +
                           dki_mom * F0[ind + 1]) * grid_cor;
          } else {
            power = 0.5 * (dkr * F0m[ind + 1] - dki * F0m[ind] + dkr_mom * F0[ind + 1] -
                           dki_mom * F0[ind]) * grid_cor;
          }
        }
        int NM = 2;
        if (k == 0) {
          if (j == 0) {
            if ((i != 0) && (i != NX / 2)) NM = 1;
          } else {
            if (i != NX / 2) NM = 1;
          }
        }
        power *= NM;
        if (multipole == 0) {
          b->pk0[kbin] += power;
          b->pk2[kbin] -= 0.5 * power;
          b->pk4[kbin] += 0.375 * power;
          b->nmodes[kbin] += NM;
        } else if (!(fktot > 0.0)) {
          continue;
        } else if (multipole == 1) {
          power *= kvec[ii] / fktot;
          b->pk1[kbin] += power;
          b->pk3[kbin] -= 1.5 * power;
        } else if (multipole == 2) {
          const double pre = ii != jj ? 2.0 : 1.0;
          power *= pre * kvec[ii] * kvec[jj] / (fktot * fktot);
          b->pk2[kbin] += 1.5 * power;
          b->pk4[kbin] -= 3.75 * power;
        } else if (multipole == 3) {
          const double pre = ii != kk ? (ii != jj ? 6.0 : 3.0) : 1.0;
          b->pk3[kbin] += 2.5 * pre * kvec[ii] * kvec[jj] * kvec[kk] / (fktot * fktot * fktot) * power;
        } else {
          const double pre = ii == jj ? (ii != kk ? 4.0 : 1.0) : (jj == kk ? 6.0 : 12.0);
          b->pk4[kbin] += 4.375 * pre * kvec[ii] * kvec[ii] * kvec[jj] * kvec[kk] /
                          (fktot * fktot * fktot * fktot) * power;
        }
      }
}

/* The index sets of the reference's loops: quadrupole (ii <= jj),
 * octopole (src/main.c:649-657) and hexadecapole (src/main.c:709-726). */
int orc_survey_combos(int multipole, int *idx) {
  int n = 0;
  if (multipole == 0) return 1;
  if (multipole == 1) {
    for (int ii = 0; ii < 3; ii++) idx[3 * n++] = ii;
    return n;
  }
  if (multipole == 2) {
    for (int ii = 0; ii < 3; ii++)
      for (int jj = ii; jj < 3; jj++) {
        idx[3 * n] = ii;
        idx[3 * n + 1] = jj;
        n++;
      }
    return n;
  }
  if (multipole == 3) {
    for (int ii = 0; ii < 3; ii++) {
      const int jjhigh = ii == 0 ? ii + 2 : ii + 1;
      for (int jj = ii; jj < jjhigh; jj++)
        for (int kk = jj != ii ? 2 : 0; kk < 3; kk++) {
          idx[3 * n] = ii;
          idx[3 * n + 1] = jj;
          idx[3 * n + 2] = kk;
          n++;
        }
    }
    return n;
  }
  for (int ii = 0; ii < 3; ii++)
    for (int jj = 0; jj < 3; jj++) {
      int kklow = 0, kkhigh = 3;
      if (jj != ii) {
        if (ii == 1) {
          kklow = 2;
        } else if (ii == 2) {
          if (jj == 1) continue;
          kklow = 1;
          kkhigh = 2;
        } else {
          kklow = jj;
        }
      }
      for (int kk = kklow; kk < kkhigh; kk++) {
        idx[3 * n] = ii;
        idx[3 * n + 1] = jj;
        idx[3 * n + 2] = kk;
        n++;
      }
    }
  return n;
}

int orc_survey_power(const orc_params *p, const double *origin, double **grids, orc_bins *b) {
  const int M = p->momentum;
  if (M == 2 || M == 5) return -1;   /* no ddg_mom grid (src/main.c:1198) */
  const int mom = (M != 0) && (M != 1);
  const int il = p->do_interlacing != 0;
  const long long sz = orc_grid_size(p);
  double *w[4] = {NULL, NULL, NULL, NULL};
  const int used[4] = {1, il, mom, mom && il};
  for (int s = 0; s < 4; s++)
    if (used[s]) w[s] = (double *)calloc((size_t)sz, sizeof(double));
  double *F0 = (double *)calloc((size_t)sz, sizeof(double));
  double *F0m = mom ? (double *)calloc((size_t)sz, sizeof(double)) : NULL;
  const int step = p->odd_multipoles ? 1 : 2;
  int idx[3 * 15];
  for (int m = 0; m <= 4; m += step) {
    const int nc = orc_survey_combos(m, idx);
    for (int c = 0; c < nc; c++) {
      const int *ix = idx + 3 * c;
      for (int s = 0; s < 4; s++)
        if (used[s]) {
          survey_weight(p, origin, grids[s], w[s], m, ix);
          orc_fft_r2c_3d(w[s], p->nx, p->ny, p->nz);
        }
      survey_assign(p, m, ix[0], ix[1], ix[2], w, F0, F0m, b);
    }
  }
  for (int s = 0; s < 4; s++) free(w[s]);
  free(F0);
  free(F0m);
  return 0;
}
