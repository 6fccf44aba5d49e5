/*
 * oracle_fft.c -- FP64 mixed-radix FFT used by the CPU oracle.
 *
 * TEST INFRASTRUCTURE ONLY.  Stands in for FFTW 3.3.x's forward r2c
 * transform (reference src/main.c:1200-1217 plans, executed at
 * src/main.c:266-271): out[k] = sum_n in[n] exp(-2 pi i k n / N),
 * unnormalised.  Recursive decimation in time over the prime factors of N
 * with a direct O(p^2) butterfly per factor; twiddles are cos/sin of
 * 2 pi j / N evaluated directly in double.  Pinned by the DFT definition,
 * numpy.fft and Parseval (tests/test_oracle_fft.py).
 */
#include "capow_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef struct { double re, im; } cpx;

static int factorize(long long n, long long *f) {
  int nf = 0;
  static const long long pref[] = {4, 2, 3, 5};
  for (int q = 0; q < 4; q++) {
    while (n % pref[q] == 0 && n > 1) { f[nf++] = pref[q]; n /= pref[q]; }
  }
  for (long long d = 7; d * d <= n; d += 2)
    while (n % d == 0) { f[nf++] = d; n /= d; }
  if (n > 1) f[nf++] = n;
  return nf;
}

/* out[0..n) = DFT of in[0, s, 2s, ...]; tw[j] = exp(-2 pi i j / N),
 * tstep = N / n. */
static void rec(const cpx *in, cpx *out, long long n, long long s,
                const long long *f, const cpx *tw, long long tstep, long long N) {
  if (n == 1) { out[0] = in[0]; return; }
  long long p = f[0], m = n / p;
  for (long long q = 0; q < p; q++) rec(in + q * s, out + q * m, m, s * p, f + 1, tw, tstep * p, N);
  cpx tstack[64];
  cpx *t = p <= 64 ? tstack : (cpx *)malloc((size_t)p * sizeof(cpx));
  for (long long k = 0; k < m; k++) {
    for (long long u = 0; u < p; u++) {
      cpx a = out[u * m + k];
      long long e = (u * k * tstep) % N;
      cpx w = tw[e];
      t[u].re = a.re * w.re - a.im * w.im;
      t[u].im = a.re * w.im + a.im * w.re;
    }
    for (long long v = 0; v < p; v++) {
      double sr = 0.0, si = 0.0;
      for (long long u = 0; u < p; u++) {
        long long e = ((u * v) % p) * (N / p);
        cpx w = tw[e];
        sr += t[u].re * w.re - t[u].im * w.im;
        si += t[u].re * w.im + t[u].im * w.re;
      }
      out[v * m + k].re = sr;
      out[v * m + k].im = si;
    }
  }
  if (t != tstack) free(t);
}

static cpx *make_tw(long long N) {
  cpx *tw = (cpx *)malloc((size_t)N * sizeof(cpx));
  for (long long j = 0; j < N; j++) {
    double a = -2.0 * M_PI * (double)j / (double)N;
    tw[j].re = cos(a);
    tw[j].im = sin(a);
  }
  return tw;
}

typedef struct {
  long long n, f[64];
  cpx *tw, *tmp;
} plan1d;

static void plan_init(plan1d *pl, long long n) {
  pl->n = n;
  memset(pl->f, 0, sizeof pl->f);
  factorize(n, pl->f);
  pl->tw = make_tw(n);
  pl->tmp = (cpx *)malloc((size_t)n * sizeof(cpx));
}

static void plan_free(plan1d *pl) { free(pl->tw); free(pl->tmp); }

/* In place on data[0..n) */
static void plan_exec(plan1d *pl, cpx *data) {
  rec(data, pl->tmp, pl->n, 1, pl->f, pl->tw, 1, pl->n);
  memcpy(data, pl->tmp, (size_t)pl->n * sizeof(cpx));
}

void orc_fft_c2c_1d(double *data, long long n) {
  plan1d pl;
  plan_init(&pl, n);
  plan_exec(&pl, (cpx *)data);
  plan_free(&pl);
}

void orc_fft_r2c_3d(double *grid, long long nx, long long ny, long long nz) {
  const long long nzh = nz / 2 + 1, row = 2 * nzh;
  long long mx = nx > ny ? nx : ny;
  if (nz > mx) mx = nz;
  cpx *buf = (cpx *)malloc((size_t)mx * sizeof(cpx));
  plan1d pz, py, px;
  plan_init(&pz, nz);
  plan_init(&py, ny);
  plan_init(&px, nx);
  /* z: real rows -> nz/2+1 complex, in the padded row */
  for (long long i = 0; i < nx; i++)
    for (long long j = 0; j < ny; j++) {
      double *r = grid + row * (j + ny * i);
      for (long long k = 0; k < nz; k++) { buf[k].re = r[k]; buf[k].im = 0.0; }
      plan_exec(&pz, buf);
      for (long long k = 0; k < nzh; k++) { r[2 * k] = buf[k].re; r[2 * k + 1] = buf[k].im; }
    }
  /* y */
  for (long long i = 0; i < nx; i++)
    for (long long k = 0; k < nzh; k++) {
      for (long long j = 0; j < ny; j++) {
        double *c = grid + row * (j + ny * i) + 2 * k;
        buf[j].re = c[0]; buf[j].im = c[1];
      }
      plan_exec(&py, buf);
      for (long long j = 0; j < ny; j++) {
        double *c = grid + row * (j + ny * i) + 2 * k;
        c[0] = buf[j].re; c[1] = buf[j].im;
      }
    }
  /* x */
  for (long long j = 0; j < ny; j++)
    for (long long k = 0; k < nzh; k++) {
      for (long long i = 0; i < nx; i++) {
        double *c = grid + row * (j + ny * i) + 2 * k;
        buf[i].re = c[0]; buf[i].im = c[1];
      }
      plan_exec(&px, buf);
      for (long long i = 0; i < nx; i++) {
        double *c = grid + row * (j + ny * i) + 2 * k;
        c[0] = buf[i].re; c[1] = buf[i].im;
      }
    }
  plan_free(&pz); plan_free(&py); plan_free(&px);
  free(buf);
}
