/* capow_host.c -- host-side output math of include/capow.h (no GPU).
 *
 * Shared by the capow host executable and the Python checking layer. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "capow.h"

/* src/main.c:103-148 at one MPI rank (NTask = 1), for the spectrum whose
 * reference Momentum value is `code`. */
int capow_scalars(const capow_params *p, const capow_assign_stats *st, int32_t code,
                  double *shot, double *norm) {
  if (!p || !st || !shot || !norm || code < 0 || code > 8) return CAPOW_ERR_ARG;
  double N = st->ngrid;
  double nbar = N / ((p->xmax - p->xmin) * (p->ymax - p->ymin) * (p->zmax - p->zmin));
  *norm = nbar * N;
  *shot = N;
  if (code) {
    double NSQ = st->nsq;
    double VR = st->vr / NSQ, VR2 = st->vr2 / NSQ, VR3 = st->vr3 / NSQ, VR4 = st->vr4 / NSQ;
    switch (code) {
      case 1: *shot = VR * N; break;
      case 2: case 3: *shot = VR2 * N; break;
      case 4: case 6: *shot = VR3 * N; break;
      default: *shot = VR4 * N; break;
    }
  }
  return CAPOW_OK;
}

/* src/main.c:1014-1056 */
int capow_finalize(const capow_params *p, double shot, double norm, double *pk5,
                   int64_t *nmodes, double *pk2d, int64_t *nmodes2d, int32_t *last_bin) {
  if (!p || !pk5 || !nmodes) return CAPOW_ERR_ARG;
  const int NK = p->nk;
  double *P0 = pk5, *P1 = pk5 + NK, *P2 = pk5 + 2 * NK, *P3 = pk5 + 3 * NK, *P4 = pk5 + 4 * NK;
  int *fill = (int *)calloc((size_t)NK, sizeof(int));
  if (!fill) return CAPOW_ERR_NOMEM;
  for (int i = 0; i < NK; i++) {
    if (nmodes[i] > 0.0 && fill[i] == 0) {
      double nm = (double)nmodes[i];
      P0[i] -= nm * shot;
      P0[i] /= nm * norm;
      P2[i] *= 5.0 / (nm * norm);
      P4[i] *= 9.0 / (nm * norm);
      if (p->odd_multipoles) {
        P1[i] *= 3.0 / (nm * norm);
        P3[i] *= 7.0 / (nm * norm);
      }
      fill[i] = 1;
    }
  }
  if (p->output_2d && pk2d && nmodes2d) {
    for (int i = 0; i < NK * p->nmu; i++)
      if (nmodes2d[i] > 0.0) pk2d[i] /= (double)nmodes2d[i] * norm;
  }
  int zeroed = -1;
  for (int i = NK - 1; i >= 0; i--) {
    if (fill[i] == 1) {
      if (p->odd_multipoles) { P1[i] = 0.0; P3[i] = 0.0; }
      fill[i] = 0; P0[i] = 0.0; P2[i] = 0.0; P4[i] = 0.0; nmodes[i] = 0;
      zeroed = i;
      break;
    }
  }
  free(fill);
  if (last_bin) *last_bin = zeroed;
  return CAPOW_OK;
}

/* src/main.c:1058-1101 */
int capow_write_output(const capow_params *p, double shot, double norm, const double *pk5,
                       const int64_t *nmodes, const double *pk2d, const int64_t *nmodes2d,
                       const char *path, int32_t precise) {
  if (!p || !pk5 || !nmodes || !path) return CAPOW_ERR_ARG;
  const int NK = p->nk;
  const double *P0 = pk5, *P1 = pk5 + NK, *P2 = pk5 + 2 * NK, *P3 = pk5 + 3 * NK, *P4 = pk5 + 4 * NK;
  double binwidth = (p->maxk - p->mink) / (double)NK;
  FILE *f = fopen(path, "w");
  if (!f) return CAPOW_ERR_IO;
  if (precise) {
    fprintf(f, "# shot = %.17g\n", shot / norm);
    fprintf(f, "# norm = %.17g\n", norm);
  } else {
    fprintf(f, "# shot = %g\n", shot / norm);
    fprintf(f, "# norm = %g\n", norm);
  }
  if (p->odd_multipoles) fprintf(f, "# k  pk0  pk1  pk2  pk3  pk4  Nmodes\n");
  else fprintf(f, "# k  pk0  pk2  pk4  Nmodes\n");
  for (int i = 0; i < NK; i++) {
    double kp;
    if (p->output_log) kp = pow(10.0, p->mink + ((float)i + 0.5) * binwidth);
    else kp = p->mink + ((float)i + 0.5) * binwidth;
    if (p->odd_multipoles) {
      if (precise)
        fprintf(f, "%.17g %.17g %.17g %.17g %.17g %.17g %d\n", kp, P0[i], P1[i], P2[i], P3[i], P4[i], (int)nmodes[i]);
      else
        fprintf(f, "%g %g %g %g %g %g %d\n", kp, P0[i], P1[i], P2[i], P3[i], P4[i], (int)nmodes[i]);
    } else {
      if (precise)
        fprintf(f, "%.17g %.17g %.17g %.17g %d\n", kp, P0[i], P2[i], P4[i], (int)nmodes[i]);
      else
        fprintf(f, "%g %g %g %g %d\n", kp, P0[i], P2[i], P4[i], (int)nmodes[i]);
    }
  }
  if (fclose(f) != 0) return CAPOW_ERR_IO;
  if (p->output_2d && pk2d && nmodes2d) {
    size_t L = strlen(path);
    char *p2 = (char *)malloc(L + 4);
    if (!p2) return CAPOW_ERR_NOMEM;
    memcpy(p2, path, L);
    memcpy(p2 + L, "_2D", 4);
    f = fopen(p2, "w");
    free(p2);
    if (!f) return CAPOW_ERR_IO;
    if (precise) {
      fprintf(f, "# shot = %.17g\n# norm = %.17g\n", shot / norm, norm);
    } else {
      fprintf(f, "# shot = %g\n# norm = %g\n", shot / norm, norm);
    }
    fprintf(f, "# k, mu, pk, Nmodes\n");
    for (int i = 0; i < NK; i++) {
      double kp;
      if (p->output_log) kp = pow(10.0, p->mink + ((float)i + 0.5) * binwidth);
      else kp = p->mink + ((float)i + 0.5) * binwidth;
      for (int j = 0; j < p->nmu; j++) {
        double mup = 2.0 * ((float)j + 0.5) / p->nmu - 1.0;
        if (precise)
          fprintf(f, "%.17g %.17g %.17g %d\n", kp, mup, pk2d[i * p->nmu + j], (int)nmodes2d[i * p->nmu + j]);
        else
          fprintf(f, "%g %g %g %d\n", kp, mup, pk2d[i * p->nmu + j], (int)nmodes2d[i * p->nmu + j]);
      }
    }
    if (fclose(f) != 0) return CAPOW_ERR_IO;
  }
  return CAPOW_OK;
}
