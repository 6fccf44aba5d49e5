/* Host-side driver for the sanitizer builds (tests/test_sanitizers.py): the
 * library's host C -- the threaded periodic catalogue reader, the survey
 * reader, n(z), FKP weights, sums and packing, the output math and writer
 * -- compiled with gcc -fsanitize=address,undefined or -fsanitize=thread
 * together with this driver and run on generated files with 1 and 8
 * threads, including their failure paths.  No GPU code is involved.
 *
 * usage: sanitize_driver <scratch directory>
 * Prints "ok" and exits 0 when every check passes; a sanitizer report makes
 * the process exit nonzero (halt_on_error / exitcode options). */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#include "capow.h"

#define CHECK(c)                                                      \
  do {                                                                \
    if (!(c)) {                                                       \
      fprintf(stderr, "check failed at line %d: %s\n", __LINE__, #c); \
      exit(3);                                                        \
    }                                                                 \
  } while (0)

static unsigned long long rng = 88172645463325252ull;
static double urand(void) {
  rng ^= rng << 13;
  rng ^= rng >> 7;
  rng ^= rng << 17;
  return (double)(rng >> 11) * (1.0 / 9007199254740992.0);
}

static void periodic_file(const char *path, int n, int bad_at, int oob_at) {
  FILE *f = fopen(path, "w");
  CHECK(f);
  fprintf(f, "# x y z vx vy vz\n");
  for (int i = 0; i < n; i++) {
    if (i == bad_at) {
      fprintf(f, "1.0 2.0 oops\n");
      continue;
    }
    double x = urand() * 1000.0, y = urand() * 1000.0, z = urand() * 1000.0;
    if (i == oob_at) x = 1200.0;
    fprintf(f, "%.9g %.9g %.9g %.4f %.4f %.4f extra\n", x, y, z, urand() - 0.5, urand() - 0.5,
            urand() - 0.5);
  }
  fclose(f);
}

static void periodic_reader(const char *dir) {
  char path[4096];
  snprintf(path, sizeof path, "%s/cat.txt", dir);
  const double lo[3] = {0, 0, 0}, hi[3] = {1000, 1000, 1000};
  /* enough lines for several 1 MB thread ranges */
  periodic_file(path, 60000, -1, -1);
  void *rows1 = NULL, *rows8 = NULL;
  int64_t n1 = 0, n8 = 0;
  char bad[256];
  double bxyz[3];
  CHECK(capow_read_ascii(path, 0, lo, hi, 1, 1, &rows1, &n1, bad, sizeof bad, bxyz) == CAPOW_OK);
  CHECK(capow_read_ascii(path, 0, lo, hi, 1, 8, &rows8, &n8, bad, sizeof bad, bxyz) == CAPOW_OK);
  CHECK(n1 == 60000 && n8 == n1 && memcmp(rows1, rows8, (size_t)n1 * 32) == 0);
  capow_free_rows(rows1);
  capow_free_rows(rows8);
  /* FP32 rows, LOS 1 (v_los = vx) */
  CHECK(capow_read_ascii(path, 1, lo, hi, 0, 8, &rows8, &n8, bad, sizeof bad, bxyz) == CAPOW_OK);
  capow_free_rows(rows8);
  /* the first failure in file order, whatever the thread count */
  periodic_file(path, 60000, 41000, 20000);
  for (int nt = 1; nt <= 8; nt += 7) {
    void *r = NULL;
    int64_t n = 0;
    CHECK(capow_read_ascii(path, 0, lo, hi, 1, nt, &r, &n, bad, sizeof bad, bxyz) ==
          CAPOW_ERR_OUT_OF_BOX);
    CHECK(fabs(bxyz[0] - 1200.0) < 1e-9);
  }
  periodic_file(path, 60000, 1000, 50000);
  CHECK(capow_read_ascii(path, 0, lo, hi, 1, 8, &rows8, &n8, bad, sizeof bad, bxyz) ==
        CAPOW_ERR_PARSE);
  CHECK(strstr(bad, "oops") != NULL);
  snprintf(path, sizeof path, "%s/missing.txt", dir);
  CHECK(capow_read_ascii(path, 0, lo, hi, 1, 4, &rows8, &n8, NULL, 0, NULL) == CAPOW_ERR_IO);
}

/* The binary sidecar: written from the text by several threads, read back
 * by several threads into the text reader's rows; the first out-of-box
 * particle; a truncated sidecar. */
static void periodic_sidecar(const char *dir) {
  char path[4096], side[4200];
  snprintf(path, sizeof path, "%s/cat.txt", dir);
  snprintf(side, sizeof side, "%s/cat.capowbin", dir);
  const double lo[3] = {0, 0, 0}, hi[3] = {1000, 1000, 1000};
  periodic_file(path, 60000, -1, -1);
  char bad[256];
  double bxyz[3];
  void *rt = NULL, *rs = NULL;
  int64_t nt = 0, ns = 0;
  CHECK(capow_write_sidecar(path, side, 8, bad, sizeof bad) == CAPOW_OK);
  CHECK(capow_read_ascii(path, 2, lo, hi, 0, 8, &rt, &nt, bad, sizeof bad, bxyz) == CAPOW_OK);
  CHECK(capow_read_sidecar(side, path, 2, lo, hi, 0, 8, &rs, &ns, bxyz) == CAPOW_OK);
  CHECK(nt == 60000 && ns == nt && memcmp(rt, rs, (size_t)nt * 16) == 0);
  capow_free_rows(rt);
  capow_free_rows(rs);
  periodic_file(path, 60000, -1, 20000);
  CHECK(capow_write_sidecar(path, side, 8, bad, sizeof bad) == CAPOW_OK);
  CHECK(capow_read_sidecar(side, path, 0, lo, hi, 1, 8, &rs, &ns, bxyz) == CAPOW_ERR_OUT_OF_BOX);
  CHECK(fabs(bxyz[0] - 1200.0) < 1e-9);
  /* truncated: not a whole sidecar */
  FILE *f = fopen(side, "r+b");
  CHECK(f != NULL && ftruncate(fileno(f), 64 + 48 * 100 - 8) == 0);
  fclose(f);
  CHECK(capow_read_sidecar(side, NULL, 0, lo, hi, 1, 8, &rs, &ns, bxyz) == CAPOW_ERR_PARSE);
}

static void survey_file(const char *path, int n, int randoms) {
  FILE *f = fopen(path, "w");
  CHECK(f);
  fprintf(f, "# ra dec z w nbar ...\n");
  for (int i = 0; i < n; i++) {
    const double ra = urand() * 360.0, dec = urand() * 160.0 - 80.0, z = 0.05 + 0.4 * urand();
    const double w = i % 97 == 0 ? 0.0 : 0.5 + urand();
    if (randoms)
      fprintf(f, "%.9g %.9g %.9g %.6g %.6g 0 %.6g %.6g\n", ra, dec, z, w, 1e-4, 0.01 * urand(),
              0.05);
    else
      fprintf(f, "%.9g %.9g %.9g %.6g %.6g 0 %.6g %.6g 0 0 0 0\n", ra, dec, z, w, 1e-4,
              0.02 * (urand() - 0.5), 0.05);
  }
  fclose(f);
}

static void survey(const char *dir) {
  char dpath[4096], rpath[4096];
  snprintf(dpath, sizeof dpath, "%s/data.txt", dir);
  snprintf(rpath, sizeof rpath, "%s/rand.txt", dir);
  survey_file(dpath, 70000, 0);
  survey_file(rpath, 140000, 1);
  capow_params g;
  memset(&g, 0, sizeof g);
  g.nx = g.ny = g.nz = 64;
  g.xmin = g.ymin = g.zmin = -2000.0;
  g.xmax = g.ymax = g.zmax = 2000.0;
  g.interp_order = 3;
  g.do_interlacing = 1;
  g.nk = 20;
  g.maxk = 0.2;
  capow_survey_params sp;
  memset(&sp, 0, sizeof sp);
  sp.coord_type = 1;
  sp.momentum = 3;
  sp.nbar_column = -1;
  sp.fkp_column = -1;
  sp.omega_m = 0.31;
  sp.redmin_in = 0.05;
  sp.redmax_in = 0.45;
  sp.sky_area = 41253.0;
  sp.fkp_pk = 1e4;
  sp.fkp_pk_mom = 5e9;
  sp.fkp_pk_cross = 1e6;
  capow_cosmo *cosmo = NULL;
  CHECK(capow_cosmo_create(sp.omega_m, &cosmo) == CAPOW_OK);
  capow_survey_obj *d = NULL, *r = NULL;
  int64_t nd = 0, nr = 0, nread = 0;
  double zr[2] = {1e30, -1e30}, xr[6], bad[4];
  char line[2000];
  CHECK(capow_read_survey(dpath, 0, &g, &sp, cosmo, 1 << 20, &d, &nd, &nread, zr, xr, line,
                          sizeof line, bad) == CAPOW_OK);
  CHECK(capow_read_survey(rpath, 1, &g, &sp, cosmo, 1 << 20, &r, &nr, &nread, zr, xr, line,
                          sizeof line, bad) == CAPOW_OK);
  CHECK(nd > 60000 && nr > 120000);
  const int32_t nb = capow_survey_nbar_bins(&sp);
  double *zn = (double *)malloc((size_t)(nb + 2) * sizeof(double));
  double *nbar = (double *)malloc((size_t)(nb + 2) * sizeof(double));
  CHECK(capow_survey_nbar(&sp, cosmo, zr, d, nd, r, nr, zn, nbar, nb + 2) == CAPOW_OK);
  CHECK(capow_survey_fkp(&sp, d, nd) == CAPOW_OK && capow_survey_fkp(&sp, r, nr) == CAPOW_OK);
  capow_survey_totals t;
  CHECK(capow_survey_sums(&sp, d, nd, r, nr, &t) == CAPOW_OK && t.alpha > 0);
  float *xyzv = (float *)malloc((size_t)(nd + nr) * 4 * sizeof(float));
  float *w = (float *)malloc((size_t)(nd + nr) * sizeof(float));
  int64_t np = 0;
  CHECK(capow_survey_pack(&sp, d, nd, r, nr, t.alpha, CAPOW_DTYPE_F32, xyzv, w, &np) == CAPOW_OK);
  CHECK(np == nd + nr);
  /* a shrunk grid: an object outside it is fatal */
  g.xmin = g.ymin = g.zmin = -300.0;
  g.xmax = g.ymax = g.zmax = 300.0;
  capow_survey_obj *d2 = NULL;
  zr[0] = 1e30;
  zr[1] = -1e30;
  CHECK(capow_read_survey(dpath, 0, &g, &sp, cosmo, 1 << 20, &d2, &nd, &nread, zr, xr, line,
                          sizeof line, bad) == CAPOW_ERR_OUT_OF_BOX);
  free(zn);
  free(nbar);
  free(xyzv);
  free(w);
  capow_free_survey(d);
  capow_free_survey(r);
  capow_cosmo_destroy(cosmo);
}

static void output(const char *dir) {
  capow_params p;
  memset(&p, 0, sizeof p);
  p.nx = p.ny = p.nz = 32;
  p.xmax = p.ymax = p.zmax = 1000.0;
  p.interp_order = 2;
  p.nk = 16;
  p.maxk = 0.16;
  p.odd_multipoles = 1;
  p.output_2d = 1;
  p.nmu = 5;
  double pk[5 * 16], pk2d[16 * 5];
  int64_t nm[16], nm2d[16 * 5];
  for (int i = 0; i < 5 * 16; i++) pk[i] = urand();
  for (int i = 0; i < 16; i++) nm[i] = i < 12 ? 10 + i : 0;
  for (int i = 0; i < 16 * 5; i++) {
    pk2d[i] = urand();
    nm2d[i] = i % 3;
  }
  capow_assign_stats st;
  memset(&st, 0, sizeof st);
  st.ngrid = 1000;
  st.nsq = 1000;
  st.vr2 = 5.0;
  double shot, norm;
  CHECK(capow_scalars(&p, &st, 2, &shot, &norm) == CAPOW_OK);
  int32_t last = -2;
  CHECK(capow_finalize(&p, shot, norm, pk, nm, pk2d, nm2d, &last) == CAPOW_OK && last == 11);
  char path[4096];
  snprintf(path, sizeof path, "%s/pk.txt", dir);
  CHECK(capow_write_output(&p, shot, norm, pk, nm, pk2d, nm2d, path, 0) == CAPOW_OK);
  CHECK(capow_write_output(&p, shot, norm, pk, nm, pk2d, nm2d, path, 1) == CAPOW_OK);
  snprintf(path, sizeof path, "%s/no/such/dir/pk.txt", dir);
  CHECK(capow_write_output(&p, shot, norm, pk, nm, NULL, NULL, path, 0) == CAPOW_ERR_IO);
}

int main(int argc, char **argv) {
  if (argc < 2) {
    fprintf(stderr, "usage: %s <scratch directory>\n", argv[0]);
    return 2;
  }
  periodic_reader(argv[1]);
  periodic_sidecar(argv[1]);
  survey(argv[1]);
  output(argv[1]);
  printf("ok\n");
  return 0;
}
