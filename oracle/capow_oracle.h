/*
 * capow_oracle.h -- CPU oracle for the periodic-box P_l(k) path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C FP64 restatement of the
 * reference CAPOW periodic pipeline at one MPI rank (NTask = 1), used as the
 * checker for the HIP product path.  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load it; the product library never
 * links or calls it.
 *
 * Every function cites the reference file:line it restates.  The FFT is this
 * project's own mixed-radix FP64 transform standing in for FFTW 3.3.x (the
 * reference links FFTW; FFTW is not in this image), pinned against the DFT
 * definition, numpy.fft and Parseval in tests/test_oracle_fft.py.
 *
 * Pinning status per function (what checks each one, see tests/):
 *   orc_add_to_grid        closed-form single-particle weights for every edge
 *                          rule, mass conservation away from edges
 *                          (tests/test_oracle_deposit.py)
 *   orc_assign             golden single-particle grids in tests/golden/ and
 *                          moments against numpy sums
 *   orc_fft_r2c_3d         brute-force DFT (<= 12^3), numpy.fft.rfftn,
 *                          Parseval (tests/test_oracle_fft.py)
 *   orc_compute_power      mode counts against an independent numpy
 *                          enumeration, plane-wave spike, L2-shaped field
 *                          (tests/test_oracle_power.py)
 *   orc_output_power       byte comparison with hand-built %g rows
 *                          (tests/test_oracle_output.py)
 *   orc_survey_power       closed forms for one and two point sources at
 *                          generic directions (sums of NM cor L_l(k.n)
 *                          and of their sin/cos cross terms), the combos'
 *                          multiplicities against the multinomial
 *                          coefficients (tests/test_oracle_survey.py)
 */
#ifndef CAPOW_ORACLE_H
#define CAPOW_ORACLE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Run parameters of the periodic path (reference src/vars.h:70-134). */
typedef struct {
  long long nx, ny, nz;           /* NX NY NZ                      */
  double xmin, xmax, ymin, ymax, zmin, zmax;
  int interp_order;               /* InterpOrder 1=NGP 2=CIC 3=TSC */
  int do_interlacing;             /* DoInterlacing                 */
  int los;                        /* LOS 0..3                      */
  int momentum;                   /* Momentum 0..8                 */
  int nk;                         /* NK                            */
  int nmu;                        /* NMU                           */
  int output_2d;                  /* Output2D                      */
  int output_log;                 /* OutputLog                     */
  int odd_multipoles;             /* Odd_Multipoles                */
  double mink, maxk;              /* Mink Maxk                     */
  double omega_m, redshift;       /* Omega_m Redshift              */
} orc_params;

/* Global sums gathered while reading (reference src/read_data.c:32-179). */
typedef struct {
  double nread, nout, ngrid, npv;
  double nsq, vr, vr2, vr3, vr4;  /* sum w, sum w v^n (n = 1..4)   */
} orc_stats;

/* Raw per-bin sums (before normalisation), reference src/main.c:302-468. */
typedef struct {
  double *pk0, *pk1, *pk2, *pk3, *pk4;  /* NK each (pk1/pk3 may be NULL) */
  long long *nmodes;                    /* NK                            */
  double *pk2d;                         /* NK*NMU or NULL                */
  long long *nmodes2d;                  /* NK*NMU or NULL                */
} orc_bins;

/* Number of doubles in one grid buffer at NTask = 1, including the
 * InterpOrder buffer slices (reference src/main.c:1131-1135). */
long long orc_grid_size(const orc_params *p);

/* Reference src/read_data.c:514-745 at NTask = 1 (xmin = XMIN,
 * xmax = XMAX, nx = NX + InterpOrder).  Returns w if accepted else 0. */
double orc_add_to_grid(const orc_params *p, double x, double y, double z,
                       double w, double vr, double *density);

/* Reference src/read_data.c:56-175: per-particle LOS selection, RSD shift,
 * bounds check, the Momentum-dependent deposits, moments, then the NTask = 1
 * halo add.  pos and vel are n x 3 (x y z / vx vy vz); w may be NULL (unit
 * weights, as the reference hard-codes).  grids[0..3] = ddg, ddg_interlace,
 * ddg_mom, ddg_mom_interlace (entries may be NULL when unused); each holds
 * orc_grid_size() doubles and must be zeroed by the caller.
 * Returns 0, or 110 for the LOS = 0 out-of-box fatal (read_data.c:81-84). */
int orc_assign(const orc_params *p, const double *pos, const double *vel,
               const double *w, long long n, double **grids, orc_stats *st);

/* Reference src/main.c:103-148: alpha, shot, norm from the read sums. */
void orc_scalars(const orc_params *p, const orc_stats *st, double *alpha,
                 double *shot, double *norm);

/* Reference src/main.c:211-225: ddg -= alpha on real cells. */
void orc_mean_subtract(const orc_params *p, double alpha, double **grids);

/* In-place forward r2c 3-D FFT of one padded grid, unnormalised, e^{-ikx}
 * (stands in for fftw_execute of the plans at src/main.c:1200-1217). */
void orc_fft_r2c_3d(double *grid, long long nx, long long ny, long long nz);

/* Forward complex 1-D FFT, exposed for the FFT tests. */
void orc_fft_c2c_1d(double *data, long long n);

/* Reference src/main.c:302-468.  Accumulates into b (zeroed by caller). */
void orc_compute_power(const orc_params *p, double **grids, orc_bins *b);

/* Reference src/main.c:1014-1056: shot subtraction, (2l+1)/(N norm),
 * last-filled-bin zeroing.  Works in place on b.  Returns the index of the
 * zeroed bin or -1. */
int orc_finalize(const orc_params *p, double shot, double norm, orc_bins *b);

/* Reference src/main.c:1058-1101: writes the text file(s).  precise != 0
 * writes %.17g instead of %g (a parity side dump).  Returns 0 or -1. */
int orc_write_output(const orc_params *p, double shot, double norm,
                     const orc_bins *b, const char *path, int precise);

/* Survey (Yamamoto) estimator, reference src/main.c:472-967 at NTask = 1:
 * grids[0..3] are the real-space fields (ddg, ddg_interlace, ddg_mom,
 * ddg_mom_interlace; unused entries may be NULL), untouched; origin =
 * X_Origin, Y_Origin, Z_Origin.  Accumulates Pk0..Pk4 (pk1/pk3 when
 * Odd_Multipoles) and Nmodes into b (zeroed by the caller).  Returns 0, or -1
 * for Momentum 2 or 5, where the reference has no ddg_mom grid. */
int orc_survey_power(const orc_params *p, const double *origin, double **grids, orc_bins *b);

/* The (ii, jj, kk) index triples of the reference's multipole loops
 * (src/main.c:487-776) for one multipole, in its order; returns their count
 * (1, 3, 6, 10, 15). */
int orc_survey_combos(int multipole, int *idx);

/* Whole periodic pipeline (reference src/main.c:28-298) on arrays. */
int orc_run(const orc_params *p, const double *pos, const double *vel,
            const double *w, long long n, orc_bins *b, orc_stats *st,
            double *alpha, double *shot, double *norm);

#ifdef __cplusplus
}
#endif
#endif
