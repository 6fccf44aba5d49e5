/* Survey ingest on the host: the redshift-distance table, natural cubic
 * splines, the survey catalogue reader, the n(z) estimate, FKP weights, the
 * survey sums (alpha, shot noise, normalisation) and the packing of data and
 * randoms into the rows the survey deposit takes.
 *
 * Reference: src/read_data.c:184-494 (read_survey_serial_ascii,
 * comoving_distance, compute_nbar, compute_fkp, assign_survey_data),
 * src/main.c:77-94 (call order), 149-204 (sums), 1233-1248 (z-r table).
 * The reference uses GSL for the integral (gsl_integration_qags) and the
 * splines (gsl_interp_cspline); here both are written from the mathematics:
 * Gauss-Legendre quadrature on each table interval, accumulated, and the
 * natural cubic spline through its second derivatives (tridiagonal solve). */
#define _GNU_SOURCE
#include <errno.h>
#include <fcntl.h>
#include <math.h>
#include <pthread.h>
#include <sched.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include "capow.h"

#define LIGHT_SPEED 299792.458 /* km/s, src/vars.h:101 */

/* ---- natural cubic spline ----------------------------------------------- */

int capow_cspline_init(int32_t n, const double *x, const double *y, double *m) {
  if (n < 3 || !x || !y || !m) return CAPOW_ERR_ARG;
  for (int32_t i = 1; i < n; i++)
    if (!(x[i] > x[i - 1])) return CAPOW_ERR_ARG;
  /* Thomas algorithm on the interior second derivatives, m[0] = m[n-1] = 0:
   * h[i-1] m[i-1] + 2 (h[i-1] + h[i]) m[i] + h[i] m[i+1] = 6 (s[i] - s[i-1]) */
  double *c = (double *)malloc((size_t)n * sizeof(double));
  if (!c) return CAPOW_ERR_NOMEM;
  m[0] = 0.0;
  m[n - 1] = 0.0;
  c[0] = 0.0;
  double prev = 0.0; /* m'-elimination carried in m[] as the right-hand side */
  for (int32_t i = 1; i < n - 1; i++) {
    const double h0 = x[i] - x[i - 1], h1 = x[i + 1] - x[i];
    const double rhs = 6.0 * ((y[i + 1] - y[i]) / h1 - (y[i] - y[i - 1]) / h0);
    const double diag = 2.0 * (h0 + h1) - h0 * c[i - 1];
    c[i] = h1 / diag;
    m[i] = (rhs - h0 * prev) / diag;
    prev = m[i];
  }
  for (int32_t i = n - 2; i >= 1; i--) m[i] -= c[i] * m[i + 1];
  free(c);
  return CAPOW_OK;
}

int capow_cspline_eval(int32_t n, const double *x, const double *y, const double *m, double xq,
                       double *out) {
  if (n < 3 || !out) return CAPOW_ERR_ARG;
  /* gsl_spline_eval reports xq outside [x[0], x[n-1]] as a domain error */
  if (!(xq >= x[0] && xq <= x[n - 1])) return CAPOW_ERR_ARG;
  int32_t lo = 0, hi = n - 1;
  while (hi - lo > 1) {
    const int32_t mid = (lo + hi) / 2;
    if (x[mid] > xq) hi = mid;
    else lo = mid;
  }
  const double h = x[hi] - x[lo], a = x[hi] - xq, b = xq - x[lo];
  *out = (m[lo] * a * a * a + m[hi] * b * b * b) / (6.0 * h) + (y[lo] / h - m[lo] * h / 6.0) * a +
         (y[hi] / h - m[hi] * h / 6.0) * b;
  return CAPOW_OK;
}

/* ---- redshift <-> comoving distance ------------------------------------- */

struct capow_cosmo {
  int32_t n;
  double *z, *r, *m_r_of_z, *m_z_of_r;
};

/* 1/E(z) for flat LCDM, src/read_data.c:375-378 */
static double inv_e(double om, double z) {
  const double a = 1.0 + z;
  return 1.0 / sqrt(om * a * a * a + (1.0 - om));
}

/* 10-point Gauss-Legendre nodes and weights on [-1, 1] (symmetric half). */
static const double GL_X[5] = {0.1488743389816312108848260, 0.4333953941292471907992659,
                               0.6794095682990244062343274, 0.8650633666889845107320967,
                               0.9739065285171717200779640};
static const double GL_W[5] = {0.2955242247147528701738930, 0.2692667193099963550912269,
                               0.2190863625159820439955349, 0.1494513491505805931457763,
                               0.0666713443086881375935688};

static double gl_interval(double om, double a, double b) {
  const double c = 0.5 * (a + b), h = 0.5 * (b - a);
  double s = 0.0;
  for (int k = 0; k < 5; k++) s += GL_W[k] * (inv_e(om, c - h * GL_X[k]) + inv_e(om, c + h * GL_X[k]));
  return s * h;
}

int capow_zr_table(double omega_m, int32_t n, double zmax, double *z, double *r) {
  if (n < 2 || !z || !r || !(zmax > 0.0) || !(omega_m >= 0.0)) return CAPOW_ERR_ARG;
  /* src/main.c:1233-1242: n nodes z_i = i zmax/(n-1), r(0) = 0,
   * r(z) = c/100 int_0^z dz'/E(z') (Mpc/h).  The integral is accumulated
   * interval by interval, each split in four 10-point Gauss-Legendre
   * panels: on the table's 0.001-wide intervals this is exact to rounding,
   * where the reference's qags stops at 1e-7 relative. */
  const double w = zmax / (double)(n - 1);
  double acc = 0.0;
  z[0] = 0.0;
  r[0] = 0.0;
  for (int32_t i = 1; i < n; i++) {
    const double z0 = (double)(i - 1) * w, z1 = (double)i * w;
    for (int p = 0; p < 4; p++) acc += gl_interval(omega_m, z0 + (z1 - z0) * p / 4.0, z0 + (z1 - z0) * (p + 1) / 4.0);
    z[i] = z1;
    r[i] = LIGHT_SPEED * acc / 100.0;
  }
  return CAPOW_OK;
}

int capow_cosmo_create(double omega_m, capow_cosmo **out) {
  if (!out) return CAPOW_ERR_ARG;
  *out = NULL;
  const int32_t n = 10000; /* src/main.c:1232 */
  capow_cosmo *c = (capow_cosmo *)calloc(1, sizeof *c);
  if (!c) return CAPOW_ERR_NOMEM;
  c->n = n;
  c->z = (double *)malloc((size_t)n * sizeof(double));
  c->r = (double *)malloc((size_t)n * sizeof(double));
  c->m_r_of_z = (double *)malloc((size_t)n * sizeof(double));
  c->m_z_of_r = (double *)malloc((size_t)n * sizeof(double));
  if (!c->z || !c->r || !c->m_r_of_z || !c->m_z_of_r) {
    capow_cosmo_destroy(c);
    return CAPOW_ERR_NOMEM;
  }
  int rc = capow_zr_table(omega_m, n, 10.0, c->z, c->r);
  if (!rc) rc = capow_cspline_init(n, c->z, c->r, c->m_r_of_z);
  if (!rc) rc = capow_cspline_init(n, c->r, c->z, c->m_z_of_r);
  if (rc) {
    capow_cosmo_destroy(c);
    return rc;
  }
  *out = c;
  return CAPOW_OK;
}

int capow_cosmo_destroy(capow_cosmo *c) {
  if (!c) return CAPOW_OK;
  free(c->z);
  free(c->r);
  free(c->m_r_of_z);
  free(c->m_z_of_r);
  free(c);
  return CAPOW_OK;
}

int capow_cosmo_r_of_z(const capow_cosmo *c, double z, double *r) {
  if (!c) return CAPOW_ERR_ARG;
  return capow_cspline_eval(c->n, c->z, c->r, c->m_r_of_z, z, r);
}

int capow_cosmo_z_of_r(const capow_cosmo *c, double r, double *z) {
  if (!c) return CAPOW_ERR_ARG;
  return capow_cspline_eval(c->n, c->r, c->z, c->m_z_of_r, r, z);
}

/* ---- survey catalogue reader -------------------------------------------- */

/* Numbers at the start of a line, as sscanf's "%lf %lf ..." reads them:
 * conversion stops at the first field that is not a number. */
static int scan_numbers(const char *s, double *v, int maxn) {
  int k = 0;
  while (k < maxn) {
    char *e;
    errno = 0;
    const double d = strtod(s, &e);
    if (e == s) break;
    v[k++] = d;
    s = e;
  }
  return k;
}

static void bad_copy(char *dst, int32_t len, const char *src) {
  if (!dst || len <= 0) return;
  strncpy(dst, src, (size_t)len - 1);
  dst[len - 1] = '\0';
}

/* One line (without its newline) of a data or randoms file.  Returns 0 =
 * kept, 1 = skipped, or a CAPOW_ERR_* status. */
static int survey_line(const char *line, int randoms, const capow_params *g,
                       const capow_survey_params *sp, const capow_cosmo *cosmo, const double lo[3],
                       const double hi[3], capow_survey_obj *o, double bad[4]) {
  double v[13];
  double tx, ty, tz, tw, tnbar, tlogdist = 0.0, tlogdist_err, tred;
  if (randoms) {
    /* "x y z w nbar_pv nbar logdist_err pv_err": eight numbers, more ignored */
    if (scan_numbers(line, v, 8) != 8) return CAPOW_ERR_PARSE;
    tx = v[0]; ty = v[1]; tz = v[2]; tw = v[3]; tnbar = v[4]; tlogdist_err = v[6];
  } else {
    /* twelve numbers exactly: the reference's format string holds thirteen
     * conversions and accepts a count of twelve (src/read_data.c:249) */
    if (scan_numbers(line, v, 13) != 12) return CAPOW_ERR_PARSE;
    tx = v[0]; ty = v[1]; tz = v[2]; tw = v[3]; tnbar = v[4];
    tlogdist = v[6]; tlogdist_err = v[7];
  }
  (void)g;
  if (sp->coord_type == 0) {
    if (capow_cosmo_z_of_r(cosmo, sqrt(tx * tx + ty * ty + tz * tz), &tred)) {
      bad[0] = tx; bad[1] = ty; bad[2] = tz; bad[3] = NAN;
      return CAPOW_ERR_ARG;
    }
    if (tred < sp->redmin_in || tred > sp->redmax_in) return 1;
  } else {
    tred = tz;
    if (tred < sp->redmin_in || tred > sp->redmax_in) return 1;
    if (tw == 0) return 1;
    double dist;
    if (capow_cosmo_r_of_z(cosmo, tred, &dist)) {
      bad[0] = tx; bad[1] = ty; bad[2] = tz; bad[3] = tred;
      return CAPOW_ERR_ARG;
    }
    double ra = tx, dec = ty;
    if (sp->coord_type == 1) {
      ra *= (M_PI / 180.0);
      dec *= (M_PI / 180.0);
    }
    tx = dist * cos(dec) * cos(ra);
    ty = dist * cos(dec) * sin(ra);
    tz = dist * sin(dec);
  }
  if (tx < lo[0] || tx >= hi[0] || ty < lo[1] || ty >= hi[1] || tz < lo[2] || tz >= hi[2]) {
    bad[0] = tx; bad[1] = ty; bad[2] = tz; bad[3] = tred;
    return CAPOW_ERR_OUT_OF_BOX;
  }
  o->x = tx; o->y = ty; o->z = tz;
  o->redshift = tred;
  o->weight = tw;
  o->weight_pv = tw;
  o->nbar = (sp->nbar_column > 0) ? tnbar : 0.0;
  o->pv = 0.0;
  o->pverr = 0.0;
  if (sp->momentum) {
    /* src/read_data.c:305-316: the low-order z(1+z) expansion with
     * q = 3/2 Om - 1, and pv = c zmod ln10 logdist / (1 + zmod) */
    const double q = 1.5 * sp->omega_m - 1.0;
    const double zmod = tred * (1.0 + 0.5 * (1.0 - q) * tred - 1.0 / 6.0 * (2.0 - q - 3.0 * q * q) * tred * tred);
    o->pv = randoms ? 0.0 : LIGHT_SPEED * zmod * log(10.0) * tlogdist / (1.0 + zmod);
    o->pverr = LIGHT_SPEED * zmod * log(10.0) * tlogdist_err / (1.0 + zmod);
  }
  return 0;
}

/* One thread's share of the file: whole lines in [beg, end). */
typedef struct {
  const char *base;
  size_t beg, end;
  int randoms;
  const capow_params *g;
  const capow_survey_params *sp;
  const capow_cosmo *cosmo;
  const double *lo, *hi;
  capow_survey_obj *objs;
  int64_t nkeep, cap, nread;
  int err;               /* first error in the range (file order), 0 if none */
  size_t err_line;       /* its offset */
  double bad[4];
  double zr[2], xl[3], xh[3];
} survey_chunk;

static void *survey_range(void *arg) {
  survey_chunk *c = (survey_chunk *)arg;
  char buf[2048];
  char *big = NULL;
  c->zr[0] = 1.0e30; c->zr[1] = -1.0e30;
  for (int a = 0; a < 3; a++) { c->xl[a] = 1.0e-30; c->xh[a] = -1.0e30; }
  size_t i = c->beg;
  while (i < c->end) {
    const char *nl = (const char *)memchr(c->base + i, '\n', c->end - i);
    if (!nl) break;                      /* the unterminated last piece is not read */
    const size_t len = (size_t)(nl - (c->base + i));
    if (len == 0) { i++; continue; }     /* empty lines: strtok skips them */
    char *line = buf;
    if (len + 1 > sizeof buf) {
      free(big);
      big = (char *)malloc(len + 1);
      if (!big) { c->err = CAPOW_ERR_NOMEM; c->err_line = i; break; }
      line = big;
    }
    memcpy(line, c->base + i, len);
    line[len] = '\0';
    capow_survey_obj o;
    const int st = survey_line(line, c->randoms, c->g, c->sp, c->cosmo, c->lo, c->hi, &o, c->bad);
    c->nread++;
    if (st != 0 && st != 1) { c->err = st; c->err_line = i; break; }
    i += len + 1;
    if (st == 1) continue;
    if (c->nkeep == c->cap) {
      const int64_t nc = c->cap ? 2 * c->cap : 4096;
      capow_survey_obj *t = (capow_survey_obj *)realloc(c->objs, (size_t)nc * sizeof *t);
      if (!t) { c->err = CAPOW_ERR_NOMEM; c->err_line = i; break; }
      c->objs = t;
      c->cap = nc;
    }
    c->objs[c->nkeep++] = o;
    if (o.redshift < c->zr[0]) c->zr[0] = o.redshift;
    if (o.redshift > c->zr[1]) c->zr[1] = o.redshift;
    const double p[3] = {o.x, o.y, o.z};
    for (int a = 0; a < 3; a++) {
      if (p[a] < c->xl[a]) c->xl[a] = p[a];
      if (p[a] > c->xh[a]) c->xh[a] = p[a];
    }
  }
  free(big);
  return NULL;
}

static int survey_threads(size_t size) {
  /* OMP_NUM_THREADS, else the CPUs this process may run on; at most 16,
   * and at least 1 MB of text per thread */
  int nt;
  const char *e = getenv("OMP_NUM_THREADS");
  if (e && atoi(e) > 0) {
    nt = atoi(e);
  } else {
    cpu_set_t set;
    CPU_ZERO(&set);
    nt = sched_getaffinity(0, sizeof set, &set) == 0 ? CPU_COUNT(&set) : 1;
  }
  if (nt > 16) nt = 16;
  const size_t max_nt = size / (1u << 20) + 1;
  if ((size_t)nt > max_nt) nt = (int)max_nt;
  return nt < 1 ? 1 : nt;
}

int capow_read_survey(const char *path, int32_t randoms, const capow_params *g,
                      const capow_survey_params *sp, const capow_cosmo *cosmo, int64_t nobj_max,
                      capow_survey_obj **rows, int64_t *n_keep, int64_t *n_read, double zrange[2],
                      double xyz_range[6], char *bad_line, int32_t bad_len, double bad[4]) {
  if (!path || !g || !sp || !cosmo || !rows || !n_keep || !zrange || nobj_max < 0)
    return CAPOW_ERR_ARG;
  *rows = NULL;
  *n_keep = 0;
  if (n_read) *n_read = 0;
  /* src/read_data.c:196-205: the box shrunk by the interpolation reach */
  const double d[3] = {(g->xmax - g->xmin) / (double)g->nx, (g->ymax - g->ymin) / (double)g->ny,
                       (g->zmax - g->zmin) / (double)g->nz};
  const double mn[3] = {g->xmin, g->ymin, g->zmin}, mx[3] = {g->xmax, g->ymax, g->zmax};
  double lo[3], hi[3];
  for (int a = 0; a < 3; a++) {
    lo[a] = mn[a] + (g->interp_order - 1) * d[a];
    hi[a] = mx[a] - (g->interp_order - 1) * d[a];
    if (g->do_interlacing) hi[a] -= 0.5 * d[a];
  }
  const int fd = open(path, O_RDONLY);
  if (fd < 0) return CAPOW_ERR_IO;
  struct stat sb;
  if (fstat(fd, &sb) != 0) { close(fd); return CAPOW_ERR_IO; }
  const size_t size = (size_t)sb.st_size;
  const char *base = NULL;
  if (size > 0) {
    base = (const char *)mmap(NULL, size, PROT_READ, MAP_PRIVATE, fd, 0);
    if (base == MAP_FAILED) { close(fd); return CAPOW_ERR_IO; }
    madvise((void *)base, size, MADV_SEQUENTIAL);
  }
  close(fd);
  /* Lines are the newline-terminated runs of the file; empty ones are
   * skipped (strtok on "\n", src/read_data.c:237); the '#' lines before the
   * first other line are skipped (src/read_data.c:215-217); a last line
   * without a newline is not read (the reference's chunk loop stops before
   * it, src/read_data.c:238, 335). */
  size_t start = 0;
  while (start < size && base[start] == '#') {
    const char *nl = (const char *)memchr(base + start, '\n', size - start);
    start = nl ? (size_t)(nl - base) + 1 : size;
  }
  const int nt = survey_threads(size - start);
  survey_chunk *ch = (survey_chunk *)calloc((size_t)nt, sizeof *ch);
  pthread_t *th = (pthread_t *)calloc((size_t)nt, sizeof *th);
  if (!ch || !th) {
    free(ch); free(th);
    if (base) munmap((void *)base, size);
    return CAPOW_ERR_NOMEM;
  }
  size_t prev = start;
  for (int t = 0; t < nt; t++) {
    size_t e = (t == nt - 1) ? size : start + (size_t)((unsigned __int128)(size - start) * (t + 1) / nt);
    if (e < prev) e = prev;
    while (e < size && e > 0 && base[e - 1] != '\n') e++;
    survey_chunk *c = &ch[t];
    c->base = base; c->beg = prev; c->end = e; c->randoms = randoms;
    c->g = g; c->sp = sp; c->cosmo = cosmo; c->lo = lo; c->hi = hi;
    prev = e;
  }
  int started = 0;
  for (int t = 1; t < nt; t++) {
    if (pthread_create(&th[t], NULL, survey_range, &ch[t]) != 0) break;
    started = t;
  }
  survey_range(&ch[0]);
  for (int t = started + 1; t < nt; t++) survey_range(&ch[t]);
  for (int t = 1; t <= started; t++) pthread_join(th[t], NULL);

  /* the first error in file order; the objects kept before it decide
   * whether the sequential reader would have passed NOBJ_Max first */
  int rc = CAPOW_OK, ferr = -1;
  int64_t kept = 0, nread = 0;
  for (int t = 0; t < nt; t++) {
    kept += ch[t].nkeep;
    nread += ch[t].nread;
    if (ch[t].err) { ferr = t; break; }
  }
  if (n_read) *n_read = nread;
  double xl[3] = {1.0e-30, 1.0e-30, 1.0e-30}, xh[3] = {-1.0e30, -1.0e30, -1.0e30};
  const int tlast = ferr >= 0 ? ferr : nt - 1;
  for (int t = 0; t <= tlast; t++) {
    if (ch[t].nkeep == 0) continue;
    if (ch[t].zr[0] < zrange[0]) zrange[0] = ch[t].zr[0];
    if (ch[t].zr[1] > zrange[1]) zrange[1] = ch[t].zr[1];
    for (int a = 0; a < 3; a++) {
      if (ch[t].xl[a] < xl[a]) xl[a] = ch[t].xl[a];
      if (ch[t].xh[a] > xh[a]) xh[a] = ch[t].xh[a];
    }
  }
  if (xyz_range)
    for (int a = 0; a < 3; a++) { xyz_range[2 * a] = xl[a]; xyz_range[2 * a + 1] = xh[a]; }
  capow_survey_obj *out = NULL;
  if (kept > nobj_max) {
    rc = CAPOW_ERR_NOMEM;
    bad_copy(bad_line, bad_len, "more objects than NOBJ_Max");
  } else if (ferr >= 0) {
    const survey_chunk *c = &ch[ferr];
    rc = c->err;
    if (bad) memcpy(bad, c->bad, sizeof c->bad);
    if (bad_line && bad_len > 0) {
      const char *l = base + c->err_line;
      const char *nl = (const char *)memchr(l, '\n', size - c->err_line);
      size_t n = nl ? (size_t)(nl - l) : size - c->err_line;
      if (n > (size_t)bad_len - 1) n = (size_t)bad_len - 1;
      memcpy(bad_line, l, n);
      bad_line[n] = '\0';
    }
  } else {
    out = (capow_survey_obj *)malloc((size_t)(kept > 0 ? kept : 1) * sizeof *out);
    if (!out) {
      rc = CAPOW_ERR_NOMEM;
    } else {
      int64_t k = 0;
      for (int t = 0; t < nt; t++) {
        if (ch[t].nkeep) memcpy(out + k, ch[t].objs, (size_t)ch[t].nkeep * sizeof *out);
        k += ch[t].nkeep;
      }
    }
  }
  for (int t = 0; t < nt; t++) free(ch[t].objs);
  free(ch);
  free(th);
  if (base) munmap((void *)base, size);
  if (rc) {
    free(out);
    return rc;
  }
  *rows = out;
  *n_keep = kept;
  return CAPOW_OK;
}

void capow_free_survey(capow_survey_obj *rows) { free(rows); }

/* ---- n(z), FKP weights, sums -------------------------------------------- */

/* n(z) spline evaluations for objects [i0, i1) of data followed by randoms. */
typedef struct {
  int32_t n;
  const double *x, *y, *m;
  capow_survey_obj *data;
  int64_t ndata;
  capow_survey_obj *randoms;
  int64_t nrand, i0, i1;
  int rc;
} nbar_job;

static void *nbar_eval(void *arg) {
  nbar_job *j = (nbar_job *)arg;
  for (int64_t i = j->i0; i < j->i1 && !j->rc; i++) {
    capow_survey_obj *o = i < j->ndata ? &j->data[i] : &j->randoms[i - j->ndata];
    j->rc = capow_cspline_eval(j->n, j->x, j->y, j->m, o->redshift, &o->nbar);
  }
  return NULL;
}

int32_t capow_survey_nbar_bins(const capow_survey_params *sp) {
  if (!sp) return -1;
  return (int32_t)ceil((sp->redmax_in - sp->redmin_in) / 0.002); /* src/read_data.c:387 */
}

int capow_survey_nbar(const capow_survey_params *sp, const capow_cosmo *cosmo, const double zrange[2],
                      capow_survey_obj *data, int64_t ndata, capow_survey_obj *randoms, int64_t nrand,
                      double *znbar, double *nbar, int32_t cap) {
  if (!sp || !cosmo || !zrange || !znbar || !nbar || ndata < 0 || nrand < 0) return CAPOW_ERR_ARG;
  const int32_t nbins = capow_survey_nbar_bins(sp);
  if (nbins < 1 || cap < nbins + 2) return CAPOW_ERR_ARG;
  /* the histogram reads the first NDATA randoms (src/read_data.c:402-406) */
  if (nrand < ndata) return CAPOW_ERR_ARG;
  const double zmin = zrange[0], zmax = zrange[1];
  const double w = (zmax - zmin) / (float)nbins;
  memset(znbar, 0, (size_t)(nbins + 2) * sizeof(double));
  memset(nbar, 0, (size_t)(nbins + 2) * sizeof(double));
  for (int64_t i = 0; i < ndata; i++) {
    int b = (int)floor((randoms[i].redshift - zmin) / w) + 1;
    if (b == nbins + 1) b--;
    if (b < 0 || b > nbins + 1) return CAPOW_ERR_INTERNAL;
    nbar[b] += 1.0;
  }
  for (int32_t i = 0; i < nbins; i++) {
    double ro, ri;
    if (capow_cosmo_r_of_z(cosmo, (i + 1) * w + zmin, &ro) || capow_cosmo_r_of_z(cosmo, i * w + zmin, &ri))
      return CAPOW_ERR_ARG;
    const double vol = sp->sky_area / (180.0 * 180.0) * M_PI * M_PI * (ro * ro * ro - ri * ri * ri) / 3.0;
    znbar[i + 1] = (i + 0.5) * w + zmin;
    nbar[i + 1] /= vol;
  }
  znbar[0] = zmin - 0.5 * w;
  znbar[nbins + 1] = zmax + 0.5 * w;
  double *m = (double *)malloc((size_t)(nbins + 2) * sizeof(double));
  if (!m) return CAPOW_ERR_NOMEM;
  int rc = capow_cspline_init(nbins + 2, znbar, nbar, m);
  if (!rc) {
    /* the evaluations are independent: split data and randoms over threads */
    const int nt = survey_threads((size_t)((ndata + nrand) / 65536) << 20);
    nbar_job job[16];
    pthread_t th[16];
    for (int t = 0; t < nt; t++) {
      job[t] = (nbar_job){nbins + 2, znbar, nbar, m, data, ndata, randoms, nrand,
                          (ndata + nrand) * t / nt, (ndata + nrand) * (t + 1) / nt, 0};
    }
    int started = 0;
    for (int t = 1; t < nt; t++) {
      if (pthread_create(&th[t], NULL, nbar_eval, &job[t]) != 0) break;
      started = t;
    }
    nbar_eval(&job[0]);
    for (int t = started + 1; t < nt; t++) nbar_eval(&job[t]);
    for (int t = 1; t <= started; t++) pthread_join(th[t], NULL);
    for (int t = 0; t < nt && !rc; t++) rc = job[t].rc;
  }
  free(m);
  return rc;
}

int capow_survey_fkp(const capow_survey_params *sp, capow_survey_obj *o, int64_t n) {
  if (!sp || (n > 0 && !o) || n < 0) return CAPOW_ERR_ARG;
  /* src/read_data.c:446-462 */
  for (int64_t i = 0; i < n; i++) {
    const double vv = o[i].pverr * o[i].pverr + 300.0 * 300.0;
    if (sp->momentum == 0) {
      o[i].weight *= 1.0 / (1.0 + o[i].nbar * sp->fkp_pk);
    } else if (sp->momentum == 1) {
      o[i].weight *= 1.0 / (vv + o[i].nbar * sp->fkp_pk_mom);
    } else {
      const double tw = 1.0 / (1.0 + o[i].nbar * sp->fkp_pk);
      const double tp = 1.0 / (vv + o[i].nbar * sp->fkp_pk_mom);
      const double tc = 1.0 / sqrt(1.0 + tw * tp * o[i].nbar * o[i].nbar * sp->fkp_pk_cross * sp->fkp_pk_cross);
      o[i].weight_pv = o[i].weight * tp * tc;
      o[i].weight *= tw * tc;
    }
  }
  return CAPOW_OK;
}

int capow_survey_sums(const capow_survey_params *sp, const capow_survey_obj *data, int64_t ndata,
                      const capow_survey_obj *randoms, int64_t nrand, capow_survey_totals *t) {
  if (!sp || !t || ndata < 0 || nrand < 0) return CAPOW_ERR_ARG;
  /* src/main.c:149-204 */
  memset(t, 0, sizeof *t);
  for (int64_t i = 0; i < ndata; i++) {
    const capow_survey_obj *o = &data[i];
    t->data_nbw += o->weight;
    t->data_nbwsq += o->weight * o->weight;
    t->data_nbsqwsq += o->weight * o->weight * o->nbar;
    if (sp->momentum) {
      t->data_vr += o->weight * o->weight_pv * o->pv;
      t->data_vrsq += o->weight * o->weight * o->pv * o->pv;
      t->data_nbsqwsq_pv += o->weight_pv * o->weight_pv * o->nbar;
    }
  }
  for (int64_t i = 0; i < nrand; i++) {
    const capow_survey_obj *o = &randoms[i];
    t->rand_nbw += o->weight;
    t->rand_nbwsq += o->weight * o->weight;
    t->rand_nbsqwsq += o->weight * o->weight * o->nbar;
    if (sp->momentum == 2) t->rand_nbsqwsq_pv += o->weight_pv * o->weight_pv * o->nbar;
  }
  t->alpha = t->data_nbw / t->rand_nbw;
  t->rand_nbwsq *= t->alpha * t->alpha;
  t->rand_nbsqwsq *= t->alpha;
  if (sp->momentum == 0) {
    t->shot = t->data_nbwsq + t->rand_nbwsq;
    t->norm = t->rand_nbsqwsq;
  } else if (sp->momentum == 1) {
    t->shot = t->data_vrsq;
    t->norm = t->rand_nbsqwsq;
  } else if (sp->momentum == 2) {
    t->shot = t->data_vr;
    t->norm = sqrt(t->rand_nbsqwsq * t->rand_nbsqwsq_pv);
  } else if (sp->momentum == 3) {
    t->shot = t->data_vrsq;
    t->norm = sqrt(t->rand_nbsqwsq * t->rand_nbsqwsq_pv);
  }
  return CAPOW_OK;
}

int capow_survey_pack(const capow_survey_params *sp, const capow_survey_obj *data, int64_t ndata,
                      const capow_survey_obj *randoms, int64_t nrand, double alpha, int32_t dtype,
                      void *xyzv, void *w, int64_t *n_out) {
  if (!sp || !xyzv || !w || !n_out || ndata < 0 || nrand < 0) return CAPOW_ERR_ARG;
  if (dtype != CAPOW_DTYPE_F32 && dtype != CAPOW_DTYPE_F64) return CAPOW_ERR_ARG;
  if (sp->momentum != 0 && sp->momentum != 1 && sp->momentum != 3) return CAPOW_ERR_UNSUPPORTED;
  /* assign_survey_data (src/read_data.c:464-486): data with prefactor 1,
   * then randoms with prefactor -alpha (src/main.c:228), except for
   * Momentum 1, which deposits the data alone. */
  const int64_t nr = (sp->momentum == 1) ? 0 : nrand;
  int64_t k = 0;
  for (int pass = 0; pass < 2; pass++) {
    const capow_survey_obj *src = pass ? randoms : data;
    const int64_t n = pass ? nr : ndata;
    for (int64_t i = 0; i < n; i++, k++) {
      const capow_survey_obj *o = &src[i];
      double v = 0.0, wt = pass ? -alpha * o->weight : o->weight;
      if (!pass && sp->momentum == 1) v = o->pv;
      if (!pass && sp->momentum == 3) {
        /* B = weight_pv pv is carried as w v.  A zero-weight object (kept
         * for Coord_Type 0, src/read_data.c:259-266) has weight_pv = 0
         * too (tw, or weight tp tc after FKP, src/read_data.c:460), so it
         * deposits nothing into either field, as in the reference.  A
         * nonzero weight_pv pv on a zero weight cannot be carried. */
        if (o->weight == 0.0) {
          if (o->weight_pv * o->pv != 0.0) return CAPOW_ERR_UNSUPPORTED;
          v = 0.0;
        } else {
          v = o->weight_pv * o->pv / o->weight;
        }
      }
      const double row[4] = {o->x, o->y, o->z, v};
      if (dtype == CAPOW_DTYPE_F64) {
        memcpy((double *)xyzv + 4 * k, row, sizeof row);
        ((double *)w)[k] = wt;
      } else {
        for (int c = 0; c < 4; c++) ((float *)xyzv)[4 * k + c] = (float)row[c];
        ((float *)w)[k] = (float)wt;
      }
    }
  }
  *n_out = k;
  return CAPOW_OK;
}
