/*
 * capow_main.c -- host program of the gfx950 periodic P_l(k) path.
 *
 * Same command line, parameter file, catalogue format, stdout banners and
 * output files as the reference's main() (src/main.c:28-298) on its
 * periodic branch, at one process: read_parameterfile + check_inputs
 * (src/read_params.c:30-428), read_periodic_serial_ascii's parsing
 * (src/read_data.c:32-88), then the C ABI of include/capow.h for the
 * assignment, FFT and binning, and output_power (src/main.c:971-1119).
 *
 *   capow <ParameterFile> <DataFile> <RandomFile> <OutputFile>
 *
 * Environment (not in the reference): CAPOW_PRECISION=fp32|fp64 (default
 * fp64, the reference's arithmetic), CAPOW_EDGE=reference|periodic
 * (default reference), CAPOW_PRECISE_OUTPUT=1 also writes
 * <OutputFile>.precise with %.17g values.
 */
#include <ctype.h>
#include <errno.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "capow.h"

/* ---- reference globals (src/vars.h:70-134), periodic subset + the rest
 * of the 41 registered tags so that the same file parses ------------------ */
static int NumFilesInParallel, DoInterlacing, InterpOrder, NK, NMU, Output2D, OutputLog;
static int Periodic, LOS, Survey, Momentum, NOBJ_Max, Coord_Type, x_Column, y_Column, z_Column;
static int NBAR_Column, FKP_Column, Odd_Multipoles;
static long long NX, NY, NZ;
static double XMIN, XMAX, YMIN, YMAX, ZMIN, ZMAX, Mink, Maxk, Omega_m, Redshift;
static double X_Origin, Y_Origin, Z_Origin, REDMININ, REDMAXIN, SkyArea, FKP_Pk, FKP_Pk_mom;
static double FKP_Pk_cross;

/* src/main.c:1298-1303 */
static void FatalError(const char *routine, int line) {
  printf("Fatal Error at line %d in routine %s\n", line, routine);
  fflush(stdout);
  exit(1);
}

#define T_FLOAT 1
#define T_INT 3
#define T_LL 4

/* src/read_params.c:30-290.  The reference registers 41 tags; each must
 * appear exactly once.  Five tags of its own example file (InputDir,
 * OutputDir, FileBase, RandFileBase, InputFormat) are commented out of the
 * reference's registry, so its shipped example is rejected there; this
 * host accepts them with a warning instead. */
static void read_parameterfile(const char *fname) {
  struct { const char *name; void *addr; int type; int seen; } tags[] = {
      {"NumFilesInParallel", &NumFilesInParallel, T_INT, 0},
      {"NX", &NX, T_LL, 0}, {"NY", &NY, T_LL, 0}, {"NZ", &NZ, T_LL, 0},
      {"XMIN", &XMIN, T_FLOAT, 0}, {"XMAX", &XMAX, T_FLOAT, 0},
      {"YMIN", &YMIN, T_FLOAT, 0}, {"YMAX", &YMAX, T_FLOAT, 0},
      {"ZMIN", &ZMIN, T_FLOAT, 0}, {"ZMAX", &ZMAX, T_FLOAT, 0},
      {"DoInterlacing", &DoInterlacing, T_INT, 0}, {"InterpOrder", &InterpOrder, T_INT, 0},
      {"NK", &NK, T_INT, 0}, {"NMU", &NMU, T_INT, 0}, {"Output2D", &Output2D, T_INT, 0},
      {"OutputLog", &OutputLog, T_INT, 0}, {"Mink", &Mink, T_FLOAT, 0},
      {"Maxk", &Maxk, T_FLOAT, 0}, {"Omega_m", &Omega_m, T_FLOAT, 0},
      {"Redshift", &Redshift, T_FLOAT, 0}, {"Periodic", &Periodic, T_INT, 0},
      {"LOS", &LOS, T_INT, 0}, {"Survey", &Survey, T_INT, 0},
      {"Momentum", &Momentum, T_INT, 0}, {"X_Origin", &X_Origin, T_FLOAT, 0},
      {"Y_Origin", &Y_Origin, T_FLOAT, 0}, {"Z_Origin", &Z_Origin, T_FLOAT, 0},
      {"REDMININ", &REDMININ, T_FLOAT, 0}, {"REDMAXIN", &REDMAXIN, T_FLOAT, 0},
      {"NOBJ_Max", &NOBJ_Max, T_INT, 0}, {"Coord_Type", &Coord_Type, T_INT, 0},
      {"x_Column", &x_Column, T_INT, 0}, {"y_Column", &y_Column, T_INT, 0},
      {"z_Column", &z_Column, T_INT, 0}, {"NBAR_Column", &NBAR_Column, T_INT, 0},
      {"SkyArea", &SkyArea, T_FLOAT, 0}, {"FKP_Column", &FKP_Column, T_INT, 0},
      {"FKP_Pk", &FKP_Pk, T_FLOAT, 0}, {"FKP_Pk_mom", &FKP_Pk_mom, T_FLOAT, 0},
      {"FKP_Pk_cross", &FKP_Pk_cross, T_FLOAT, 0},
      {"Odd_Multipoles", &Odd_Multipoles, T_INT, 0},
  };
  static const char *legacy[] = {"InputDir", "OutputDir", "FileBase", "RandFileBase",
                                 "InputFormat"};
  const int nt = (int)(sizeof tags / sizeof tags[0]);
  int legacy_seen[5] = {0, 0, 0, 0, 0};
  int errorFlag = 0;
  char buf[500], buf1[500], buf2[500], buf3[500];
  FILE *fd = fopen(fname, "r");
  if (!fd) {
    fprintf(stdout, "\nERROR: Parameter file '%s' not found.\n\n", fname);
    FatalError("read_param.c", 330);
  }
  while (!feof(fd)) {
    buf[0] = 0;
    if (!fgets(buf, 500, fd)) buf[0] = 0;
    if (sscanf(buf, "%499s%499s%499s", buf1, buf2, buf3) < 2) continue;
    if (buf1[0] == '%') continue;
    int j = -1;
    for (int i = 0; i < nt; i++)
      if (!tags[i].seen && strcmp(buf1, tags[i].name) == 0) { j = i; break; }
    if (j >= 0) {
      tags[j].seen = 1;
      if (tags[j].type == T_FLOAT) *(double *)tags[j].addr = atof(buf2);
      else if (tags[j].type == T_INT) *(int *)tags[j].addr = atoi(buf2);
      else *(long long *)tags[j].addr = (long long)atoi(buf2);
      continue;
    }
    int lg = -1;
    for (int i = 0; i < 5; i++)
      if (strcmp(buf1, legacy[i]) == 0) lg = i;
    if (lg >= 0 && !legacy_seen[lg]) {
      legacy_seen[lg] = 1;
      fprintf(stdout, "WARNING: tag '%s' is not used by the periodic path; ignored.\n", buf1);
      continue;
    }
    fprintf(stdout, "\nERROR: In file %s:  Tag '%s' not allowed or multiple defined.\n\n", fname, buf1);
    errorFlag = 1;
  }
  fclose(fd);
  for (int i = 0; i < nt; i++) {
    if (!tags[i].seen) {
      fprintf(stdout, "\nERROR: I miss a value for tag '%s' in parameter file '%s'.\n\n", tags[i].name, fname);
      errorFlag = 1;
    }
  }
  if (errorFlag) {
    fflush(stdout);
    exit(1);
  }
}

/* src/read_params.c:294-428 */
static void check_inputs(void) {
  if (NumFilesInParallel < 1) {
    printf("\nERROR: `NumFilesInParallel' is less than 1 so no processors will be writing out the data.\n");
    printf("        Please set NumFileInParallel > 0'.\n\n");
    FatalError("read_param.c", 297);
  }
  if (1 < NumFilesInParallel) {
    printf("\nWARNING: Number of processors smaller than `NumFilesInParallel'.\n");
    printf("         Setting NumFilesInParallel = Number of processors'.\n\n");
  }
  if (Periodic) {
    if ((LOS != 0) && (LOS != 1) && (LOS != 2) && (LOS != 3)) {
      printf("\nERROR: `LOS' is neither 0 (real-space), 1, 2 nor 3.\n");
      printf("        Please choose a suitable value for `LOS'.\n\n");
      FatalError("read_param.c", 313);
    }
  }
  if ((InterpOrder != 1) && (InterpOrder != 2) && (InterpOrder != 3)) {
    printf("\nERROR: `InterpOrder' is neither 1, 2 nor 3.\n");
    printf("        Please choose a suitable value for `InterpOrder'.\n\n");
    FatalError("read_param.c", 329);
  }
  if ((DoInterlacing != 0) && (DoInterlacing != 1)) {
    printf("\nERROR: `DoInterlacing' is neither 0 nor 1.\n");
    printf("        Please choose a suitable value for `DoInterlacing'.\n\n");
    FatalError("read_param.c", 337);
  }
  if ((Output2D != 0) && (Output2D != 1)) {
    printf("\nERROR: `Output2D' is neither 0 nor 1.\n");
    printf("        Please choose a suitable value for `Output2D'.\n\n");
    FatalError("read_param.c", 345);
  }
  if (Output2D && NMU <= 0) {
    printf("\nERROR: `NMU' is less than 1.  if Output2D is 1, we must choose some mu-bins to output.\n");
    printf("        Please set 'NMU' > 0.\n\n");
    FatalError("read_param.c", 354);
  }
  if ((OutputLog != 0) && (OutputLog != 1)) {
    printf("\nERROR: `OutputLog' is neither 0 nor 1.\n");
    printf("        Please choose a suitable value for `OutputLog'.\n\n");
    FatalError("read_param.c", 363);
  }
  if (NX <= 0) {
    printf("\nERROR: `NX' is less than 1. We must generate a mesh.\n");
    printf("        Please set 'NX' > 0.\n\n");
    FatalError("read_param.c", 371);
  }
  if (NY <= 0) {
    printf("\nERROR: `NY' is less than 1. We must generate a mesh.\n");
    printf("        Please set 'NY' > 0.\n\n");
    FatalError("read_param.c", 379);
  }
  if (NZ <= 0) {
    printf("\nERROR: `NZ' is less than 1. We must generate a mesh.\n");
    printf("        Please set 'NZ' > 0.\n\n");
    FatalError("read_param.c", 387);
  }
  if (NK <= 0) {
    printf("\nERROR: `NK' is less than 1. We must choose some k-bins to output.\n");
    printf("        Please set 'NK' > 0.\n\n");
    FatalError("read_param.c", 395);
  }
  if (XMAX < XMIN) {
    printf("\nERROR: `XMAX' is less than `XMIN'.\n");
    printf("        Please set `XMAX' > `XMIN'.\n\n");
    FatalError("read_param.c", 403);
  }
  if (YMAX < YMIN) {
    printf("\nERROR: `XMAX' is less than `XMIN'.\n");
    printf("        Please set `XMAX' > `XMIN'.\n\n");
    FatalError("read_param.c", 411);
  }
  if (ZMAX < ZMIN) {
    printf("\nERROR: `XMAX' is less than `XMIN'.\n");
    printf("        Please set `XMAX' > `XMIN'.\n\n");
    FatalError("read_param.c", 419);
  }
  if (Maxk < Mink) {
    printf("\nERROR: `Maxk' is less than `Mink'.\n");
    printf("        Please set `Maxk' > `Mink'.\n\n");
    FatalError("read_param.c", 427);
  }
}

static void banner(const char *t) {
  size_t n = strlen(t);
  printf("\n%s\n", t);
  for (size_t i = 0; i < n; i++) putchar('=');
  putchar('\n');
  fflush(stdout);
}

/* src/read_data.c:32-88: read the catalogue, pick v_los by LOS, check the
 * box for LOS = 0 (capow_read_ascii: the same line semantics, parsed over
 * up to 16 threads).  Returns rows of (x, y, z, v_los) in the working
 * precision.  The RSD shift and the deposit run on the GPU. */
/* CAPOW_SIDECAR=1: the rows come from the binary sidecar <path>.capowbin
 * when it is current (written from this text file, unchanged since);
 * otherwise the text is parsed and the sidecar written for the next run.
 * Both give the same rows and the same errors. */
static void *read_catalogue(const char *path, int fp64, long long *n_out) {
  const double lo[3] = {XMIN, YMIN, ZMIN}, hi[3] = {XMAX, YMAX, ZMAX};
  void *rows = NULL;
  int64_t n = 0;
  char bad[2000];
  double xyz[3] = {0, 0, 0};
  const char *sc = getenv("CAPOW_SIDECAR");
  const int use_side = sc && atoi(sc) != 0;
  char side[4200];
  int rc = -1;
  if (use_side && snprintf(side, sizeof side, "%s.capowbin", path) < (int)sizeof side) {
    rc = capow_read_sidecar(side, path, LOS, lo, hi, fp64, 0, &rows, &n, xyz);
    if (rc != CAPOW_OK && rc != CAPOW_ERR_OUT_OF_BOX) {
      /* absent, stale or damaged: parse the text and write a new one */
      rc = capow_write_sidecar(path, side, 0, bad, (int32_t)sizeof bad);
      if (rc == CAPOW_OK) rc = capow_read_sidecar(side, path, LOS, lo, hi, fp64, 0, &rows, &n, xyz);
      else if (rc != CAPOW_ERR_PARSE) rc = -1;   /* cannot write it: the text reader */
    }
  }
  if (rc == -1)
    rc = capow_read_ascii(path, LOS, lo, hi, fp64, 0, &rows, &n, bad, (int32_t)sizeof bad, xyz);
  if (rc == CAPOW_ERR_IO) {
    printf("Task %d cannot open input file\n", 0);
    FatalError("read_data", 96);
  }
  if (rc == CAPOW_ERR_PARSE) {
    printf("Task %d has error reading file: %s\n", 0, bad);
    FatalError("read_data", 102);
  }
  if (rc == CAPOW_ERR_OUT_OF_BOX) {
    printf("Task %d has particle out of grid bounds: x=%lf, y=%lf, z=%lf\n", 0, xyz[0], xyz[1], xyz[2]);
    FatalError("read_data", 110);
  }
  if (rc != CAPOW_OK) {
    printf("Task 0 unable to allocate memory for the catalogue\n");
    FatalError("read_data", 96);
  }
  *n_out = (long long)n;
  return rows;
}

static void check(capow_ctx *ctx, int rc, const char *routine, int line) {
  if (rc == CAPOW_OK) return;
  printf("%s\n", capow_last_error(ctx));
  FatalError(routine, line);
}

/* Finalisation and output shared by both branches (src/main.c:971-1119). */
static void write_spectra(const capow_params *p, double shot, double norm, double *pk,
                          int64_t *nmodes, double *pk2d, int64_t *nmodes2d, const char *outfile) {
  banner("Outputting power spectra");
  printf("Writing multipoles to file: %s\n", outfile);
  if (Output2D) printf("Writing 2D power spectrum to file:%s_2D\n", outfile);
  fflush(stdout);
  int32_t last = -1;
  capow_finalize(p, shot, norm, pk, nmodes, Output2D ? pk2d : NULL, Output2D ? nmodes2d : NULL, &last);
  if (last >= 0) {
    double bw = (Maxk - Mink) / (double)NK;
    double kmax = OutputLog ? pow(10.0, Mink + ((float)last + 0.5) * bw) : Mink + ((float)last + 0.5) * bw;
    printf("Last full bin: %d (k=%g)\n", last, kmax);
    fflush(stdout);
  }
  int rc = capow_write_output(p, shot, norm, pk, nmodes, Output2D ? pk2d : NULL,
                              Output2D ? nmodes2d : NULL, outfile, 0);
  if (rc) {
    printf("cannot open output file: %s\n", outfile);
    FatalError("read_data", 142);
  }
  const char *pr = getenv("CAPOW_PRECISE_OUTPUT");
  if (pr && pr[0] == '1') {
    size_t L = strlen(outfile);
    char *q = (char *)malloc(L + 16);
    snprintf(q, L + 16, "%s.precise", outfile);
    capow_write_output(p, shot, norm, pk, nmodes, Output2D ? pk2d : NULL,
                       Output2D ? nmodes2d : NULL, q, 1);
    free(q);
  }
}

/* read_survey_serial_ascii for one file (src/read_data.c:184-361), with
 * the reference's messages for its fatal cases. */
static capow_survey_obj *read_survey_file(const char *path, int randoms, const capow_params *p,
                                          const capow_survey_params *sp, const capow_cosmo *cosmo,
                                          double zr[2], long long *n_out) {
  printf("Reading: %s\n", path);
  fflush(stdout);
  capow_survey_obj *rows = NULL;
  int64_t n = 0, nread = 0;
  double xr[6], bad[4] = {0, 0, 0, 0};
  char line[2000];
  line[0] = 0;
  const int rc = capow_read_survey(path, randoms, p, sp, cosmo, NOBJ_Max, &rows, &n, &nread, zr, xr,
                                   line, (int32_t)sizeof line, bad);
  if (rc == CAPOW_ERR_IO) {
    printf("Task %d cannot open input file\n", 0);
    FatalError("read_data", 113);
  }
  if (rc == CAPOW_ERR_PARSE) {
    printf("Task %d has error reading file: %s\n", 0, line);
    FatalError("read_data", 102);
  }
  if (rc == CAPOW_ERR_OUT_OF_BOX) {
    printf("Task %d has object out of grid bounds for chosen InterpOrder: x=%lf, y=%lf, z=%lf, red=%lf\n",
           0, bad[0], bad[1], bad[2], bad[3]);
    FatalError("read_data", 155);
  }
  if (rc == CAPOW_ERR_NOMEM) {
    printf("Task %d has more objects than NOBJ_Max (%d)\n", 0, NOBJ_Max);
    FatalError("main", 72);
  }
  if (rc) {
    printf("Task %d has an object outside the redshift-distance table: %s\n", 0, line);
    FatalError("read_data", 261);
  }
  printf("Have %llu objects\n", (unsigned long long)n);
  printf("%12.6lf <   X  < %12.6lf\n", xr[0], xr[1]);
  printf("%12.6lf <   Y  < %12.6lf\n", xr[2], xr[3]);
  printf("%12.6lf <   Z  < %12.6lf\n", xr[4], xr[5]);
  fflush(stdout);
  *n_out = (long long)n;
  return rows;
}

/* The survey branch of src/main.c:77-94, 149-204, 226-228 and 279-280:
 * data and randoms read, n(z) and FKP weights on the host, the data and the
 * -alpha randoms deposited on the GPU, the Yamamoto estimator, and the
 * shared output. */
static int run_survey(const char *datafile, const char *randfile, const char *outfile, int fp64) {
  capow_params p;
  memset(&p, 0, sizeof p);
  p.nx = NX; p.ny = NY; p.nz = NZ;
  p.xmin = XMIN; p.xmax = XMAX; p.ymin = YMIN; p.ymax = YMAX; p.zmin = ZMIN; p.zmax = ZMAX;
  p.interp_order = InterpOrder; p.do_interlacing = DoInterlacing; p.los = 0;
  p.momentum = Momentum;
  p.nk = NK; p.nmu = NMU; p.output_2d = Output2D; p.output_log = OutputLog;
  p.odd_multipoles = Odd_Multipoles; p.mink = Mink; p.maxk = Maxk;
  p.omega_m = Omega_m; p.redshift = Redshift;
  p.precision = fp64 ? CAPOW_FP64 : CAPOW_FP32;
  p.edge_mode = CAPOW_EDGE_SURVEY;
  capow_survey_params sp;
  memset(&sp, 0, sizeof sp);
  sp.coord_type = Coord_Type; sp.momentum = Momentum; sp.nbar_column = NBAR_Column;
  sp.fkp_column = FKP_Column; sp.omega_m = Omega_m; sp.redmin_in = REDMININ;
  sp.redmax_in = REDMAXIN; sp.sky_area = SkyArea; sp.fkp_pk = FKP_Pk;
  sp.fkp_pk_mom = FKP_Pk_mom; sp.fkp_pk_cross = FKP_Pk_cross;
  if (Momentum != 0 && Momentum != 1 && Momentum != 3) {
    printf("\nERROR: survey `Momentum' must be 0, 1 or 3 (2 has no ddg_mom grid).\n\n");
    FatalError("main", 1197);
  }

  banner("Allocating grid");
  printf("Creating redshift-distance lookup table...\n");
  fflush(stdout);
  capow_cosmo *cosmo = NULL;
  if (capow_cosmo_create(Omega_m, &cosmo)) FatalError("main", 1242);

  banner("Reading input data and assigning to grid");
  double zr[2] = {1.0e30, -1.0e30};
  long long nd = 0, nr = 0;
  capow_survey_obj *data = read_survey_file(datafile, 0, &p, &sp, cosmo, zr, &nd);
  capow_survey_obj *rand = read_survey_file(randfile, 1, &p, &sp, cosmo, zr, &nr);
  if (NBAR_Column < 0) {
    printf("Calculating Number Density...\n");
    const int32_t nb = capow_survey_nbar_bins(&sp);
    double *zn = (double *)calloc((size_t)nb + 2, sizeof(double));
    double *nbv = (double *)calloc((size_t)nb + 2, sizeof(double));
    if (capow_survey_nbar(&sp, cosmo, zr, data, nd, rand, nr, zn, nbv, nb + 2)) {
      printf("Task 0 cannot compute the number density (fewer randoms than data, or a bin outside the table)\n");
      FatalError("read_data", 384);
    }
    for (int32_t i = 1; i <= nb; i++) printf("%12.6lf, %12.6lf\n", zn[i], nbv[i]);
    free(zn);
    free(nbv);
  }
  if (FKP_Column < 0) {
    printf("Computing FKP Weights...\n");
    capow_survey_fkp(&sp, data, nd);
    capow_survey_fkp(&sp, rand, nr);
  }

  banner("Computing shot-noise and normalisation");
  capow_survey_totals t;
  capow_survey_sums(&sp, data, nd, rand, nr, &t);
  if (Momentum == 0) {
    printf("alpha = %g\n", t.alpha);
    printf("shot-noise [data, randoms]    = %g, %g\n", t.data_nbwsq, t.rand_nbwsq);
    printf("normalisation [data, randoms] = %g, %g\n", t.data_nbsqwsq, t.rand_nbsqwsq);
  } else if (Momentum == 1) {
    printf("shot-noise [data]    = %g \n", t.data_vrsq);
    printf("normalisation [data, randoms] = %g, %g\n", t.data_nbsqwsq, t.rand_nbsqwsq);
  } else {
    printf("alpha = %g\n", t.alpha);
    printf("shot-noise    = %g\n", t.data_vr);
    printf("normalisation = %g\n", t.norm);
  }
  fflush(stdout);

  banner("Subtracting off expected density");
  const int64_t ntot = (int64_t)nd + (int64_t)nr;
  const size_t es = fp64 ? sizeof(double) : sizeof(float);
  void *xyzv = malloc((size_t)(ntot > 0 ? ntot : 1) * 4 * es);
  void *w = malloc((size_t)(ntot > 0 ? ntot : 1) * es);
  int64_t np = 0;
  if (!xyzv || !w) FatalError("main", 72);
  if (capow_survey_pack(&sp, data, nd, rand, nr, t.alpha, fp64 ? CAPOW_DTYPE_F64 : CAPOW_DTYPE_F32,
                        xyzv, w, &np)) {
    printf("Task 0 has a data object with zero weight and a velocity\n");
    FatalError("read_data", 480);
  }
  capow_free_survey(data);
  capow_free_survey(rand);
  capow_cosmo_destroy(cosmo);
  capow_ctx *ctx = NULL;
  check(NULL, capow_create(&p, &ctx), "main", 1123);
  check(ctx, capow_set_particles(ctx, xyzv, w, np, fp64 ? CAPOW_DTYPE_F64 : CAPOW_DTYPE_F32, CAPOW_HOST),
        "read_data", 471);
  capow_assign_stats st;
  check(ctx, capow_assign(ctx, &st), "read_data", 471);
  free(xyzv);
  free(w);
  printf("Task %d has %llu objects, gridded %lf\n", 0, (unsigned long long)nd, (double)nd);
  fflush(stdout);

  banner("Fourier transforming overdensity field");
  banner("Iterating over grid cells");
  const int nk = NK, nmu = Output2D ? NMU : 1;
  double *pk = (double *)calloc((size_t)5 * nk, sizeof(double));
  int64_t *nmodes = (int64_t *)calloc((size_t)nk, sizeof(int64_t));
  double *pk2d = (double *)calloc((size_t)nk * nmu, sizeof(double));
  int64_t *nmodes2d = (int64_t *)calloc((size_t)nk * nmu, sizeof(int64_t));
  const double origin[3] = {X_Origin, Y_Origin, Z_Origin};
  check(ctx, capow_survey_power(ctx, origin, pk, nmodes), "main", 280);
  write_spectra(&p, t.shot, t.norm, pk, nmodes, pk2d, nmodes2d, outfile);
  banner("Cleaning up");
  capow_destroy(ctx);
  free(pk); free(nmodes); free(pk2d); free(nmodes2d);
  return 0;
}

int main(int argc, char **argv) {
  if (argc < 5) {
    fprintf(stdout, "Input parameters not found\n");
    fprintf(stdout, "Call with <ParameterFile, DataFile, RandomFile, OutputFile>\n");
    fflush(stdout);
    exit(1);
  }
  banner("Reading Input Parameters and setting up CAPOW");
  read_parameterfile(argv[1]);
  check_inputs();
  const char *datafile = argv[2], *outfile = argv[4];
  const char *pe = getenv("CAPOW_PRECISION");
  const int fp64 = !(pe && (strcmp(pe, "fp32") == 0 || strcmp(pe, "FP32") == 0));
  if (!Periodic) {
    if (Survey) return run_survey(datafile, argv[3], outfile, fp64);
    printf("\nERROR: neither `Periodic' nor `Survey' is set.\n\n");
    FatalError("main", 75);
  }
  const char *ee = getenv("CAPOW_EDGE");
  const int edge = (ee && strcmp(ee, "periodic") == 0) ? CAPOW_EDGE_PERIODIC : CAPOW_EDGE_REFERENCE;

  capow_params p;
  memset(&p, 0, sizeof p);
  p.nx = NX; p.ny = NY; p.nz = NZ;
  p.xmin = XMIN; p.xmax = XMAX; p.ymin = YMIN; p.ymax = YMAX; p.zmin = ZMIN; p.zmax = ZMAX;
  p.interp_order = InterpOrder; p.do_interlacing = DoInterlacing; p.los = LOS;
  p.momentum = (Momentum >= 0 && Momentum <= 8) ? Momentum : 8;  /* src/read_data.c:124 else */
  p.nk = NK; p.nmu = NMU; p.output_2d = Output2D; p.output_log = OutputLog;
  p.odd_multipoles = Odd_Multipoles; p.mink = Mink; p.maxk = Maxk;
  p.omega_m = Omega_m; p.redshift = Redshift;
  p.precision = fp64 ? CAPOW_FP64 : CAPOW_FP32;
  p.edge_mode = edge;

  banner("Allocating grid");
  banner("Reading input data and assigning to grid");
  printf("\nReading: %s\n", datafile);
  fflush(stdout);
  long long n = 0;
  void *rows = read_catalogue(datafile, fp64, &n);
  /* the device grids are allocated once the catalogue has parsed */
  capow_ctx *ctx = NULL;
  check(NULL, capow_create(&p, &ctx), "main", 1123);
  check(ctx, capow_set_particles(ctx, rows, NULL, n, fp64 ? CAPOW_DTYPE_F64 : CAPOW_DTYPE_F32,
                                 CAPOW_HOST), "read_data", 96);
  capow_assign_stats st;
  int rc = capow_assign(ctx, &st);
  if (rc == CAPOW_ERR_OUT_OF_BOX) {
    printf("Task 0 has particle out of grid bounds\n");
    FatalError("read_data", 110);
  }
  check(ctx, rc, "read_data", 133);
  capow_free_rows(rows);
  if (LOS == 0) printf("Task %d read %lf particles, gridded %lf\n", 0, st.nread, st.ngrid);
  else printf("Task %d read %lf particles, gridded %lf (%lf out of bounds due to RSD)\n", 0, st.nread, st.ngrid, st.nout);

  banner("Computing shot-noise and normalisation");
  double shot, norm;
  capow_scalars(&p, &st, p.momentum, &shot, &norm);
  double nbar = st.ngrid / ((XMAX - XMIN) * (YMAX - YMIN) * (ZMAX - ZMIN));
  printf("mean          = %g\n", st.alpha);
  printf("nbar          = %g\n", nbar);
  printf("shot-noise    = %g\n", st.ngrid);
  printf("normalisation = %g\n", norm);
  if (Momentum) {
    printf("vr_ave = %g\n", st.vr / st.nsq);
    printf("vr2_ave = %g\n", st.vr2 / st.nsq);
    printf("vr3_ave = %g\n", st.vr3 / st.nsq);
    printf("vr4_ave = %g\n", st.vr4 / st.nsq);
  }
  banner("Subtracting off expected density");
  banner("Fourier transforming overdensity field");
  check(ctx, capow_fft(ctx), "main", 266);
  banner("Iterating over grid cells");
  const int nk = NK, nmu = Output2D ? NMU : 1;
  double *pk = (double *)calloc((size_t)3 * 5 * nk, sizeof(double));
  int64_t *nmodes = (int64_t *)calloc((size_t)nk, sizeof(int64_t));
  double *pk2d = (double *)calloc((size_t)3 * nk * nmu, sizeof(double));
  int64_t *nmodes2d = (int64_t *)calloc((size_t)nk * nmu, sizeof(int64_t));
  capow_bins b;
  memset(&b, 0, sizeof b);
  b.pk = pk; b.nmodes = nmodes; b.pk2d = pk2d; b.nmodes2d = nmodes2d;
  check(ctx, capow_bin(ctx, &b), "main", 302);

  write_spectra(&p, shot, norm, pk, nmodes, pk2d, nmodes2d, outfile);
  banner("Cleaning up");
  capow_destroy(ctx);
  free(pk); free(nmodes); free(pk2d); free(nmodes2d);
  return 0;
}
