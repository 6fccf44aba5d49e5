/*
 * capow.h -- C ABI of the MI355X (gfx950) periodic-box power-spectrum path.
 *
 * The reference (CAPOW) has no library boundary: its periodic path is a set
 * of internal C calls on globals,
 *   read_parameterfile()        src/read_params.c:30      (parameters)
 *   create_grids()              src/main.c:1123           (grids, FFT plans)
 *   read_periodic_serial_ascii  src/read_data.c:32        (ingest + deposit)
 *   add_to_grid()               src/read_data.c:514       (NGP/CIC/TSC)
 *   mean subtraction            src/main.c:211-225
 *   fftw_execute(plan...)       src/main.c:266-271        (r2c 3-D FFT)
 *   compute_periodic_power()    src/main.c:302            (binning)
 *   output_power()              src/main.c:971            (normalise, write)
 * and it aborts through FatalError (src/main.c:1298-1303) on any error.
 * This header replaces that sequence with an explicit context: every entry
 * point returns a status (0 = CAPOW_OK) and never aborts; the host maps a
 * nonzero status onto the reference's "Fatal Error" + exit(1).
 *
 * Conventions
 *   - Plain C types only.  Host arrays belong to the caller; device memory
 *     (grids, sorted particles, bins) belongs to the context.
 *   - A context is not thread-safe: one context per host thread.
 *   - All GPU work of a context runs on one HIP stream (the legacy default
 *     stream unless capow_set_stream is called); calls that return host
 *     results synchronise that stream, the others do not.
 */
#ifndef CAPOW_H
#define CAPOW_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CAPOW_ABI_VERSION 1

/* Status codes.  CAPOW_ERR_OUT_OF_BOX mirrors the reference's fatal exit for
 * a LOS = 0 particle outside the box (src/read_data.c:81-84). */
enum {
  CAPOW_OK = 0,
  CAPOW_ERR_ARG = 1,        /* invalid argument / parameter value          */
  CAPOW_ERR_HIP = 2,        /* a HIP runtime call failed                   */
  CAPOW_ERR_NOMEM = 3,      /* device or host allocation failed            */
  CAPOW_ERR_STATE = 4,      /* call out of order (e.g. bin before fft)     */
  CAPOW_ERR_UNSUPPORTED = 5,/* grid shape the kernels do not handle        */
  CAPOW_ERR_INTERNAL = 6,   /* a kernel self-check failed                  */
  CAPOW_ERR_IO = 7,         /* file could not be opened / written          */
  CAPOW_ERR_PARSE = 8,      /* catalogue line without six numbers          */
  CAPOW_ERR_STALE = 9,      /* binary sidecar older than its text catalogue */
  CAPOW_ERR_OUT_OF_BOX = 110
};

/* Arithmetic form.  FP64 is the plain form: every stage in double, as the
 * reference.  FP32: positions, deposit weights, grids and FFT in float;
 * twiddles generated in double; bin accumulation, mode counts,
 * normalisation and shot noise in double. */
enum { CAPOW_FP32 = 0, CAPOW_FP64 = 1 };

/* Edge rules of the mass assignment.  REFERENCE reproduces the reference's
 * rules at one MPI rank exactly (TSC drops low taps landing on index 0,
 * interlaced CIC/NGP reject shifted coordinates >= MAX, LOS != 0 is
 * non-periodic: src/read_data.c:514-745 with the halo add of
 * src/read_data.c:152-175).  PERIODIC is textbook periodic assignment. */
enum { CAPOW_EDGE_REFERENCE = 0, CAPOW_EDGE_PERIODIC = 1, CAPOW_EDGE_SURVEY = 2 };
/* CAPOW_EDGE_SURVEY deposits a survey catalogue (data with weight w, randoms
 * with weight -alpha w) the way assign_survey_data does
 * (src/read_data.c:461-486, src/main.c:94, 228): `momentum` is the survey
 * code (0 density: A = w; 1 velocity: A = w v; 3 density + velocity: A = w,
 * B = w v, where the caller passes v = weight_pv pv / weight so that B holds
 * weight_pv pv), no mean subtraction, no RSD shift, non-periodic edges.  The
 * survey reader admits only objects whose every tap lies inside the grid
 * (src/read_data.c:196-205, 287-290), so the reference's survey edge clamps
 * (src/read_data.c:577-589) never apply.  Survey codes 2 and 4..8 are
 * CAPOW_ERR_UNSUPPORTED (the reference allocates no ddg_mom for 2,
 * src/main.c:1197, and has no survey field for the rest).  capow_fft,
 * capow_bin and capow_backward reject a survey context; capow_survey_power
 * is its estimator. */

/* Element type of arrays handed to capow_set_particles. */
enum { CAPOW_DTYPE_F32 = 0, CAPOW_DTYPE_F64 = 1 };

/* Where a pointer lives. */
enum { CAPOW_HOST = 0, CAPOW_DEVICE = 1 };

/* Extra spectra computed from the same grids in the same binning pass.  The
 * deposited fields are fixed by `momentum` exactly as in the reference
 * (src/read_data.c:91-130): field A (ddg) = w*v^a, field B (ddg_mom) =
 * w*v^b.  The primary spectrum is always the reference's `momentum`
 * spectrum.  CAPOW_SPEC_AUTO_A adds |A|^2 (the reference spectrum with
 * Momentum = P_aa: 0, 2 or 5) and CAPOW_SPEC_AUTO_B adds |B|^2 (P_bb: 2 or
 * 5; only when b <= 2). */
enum { CAPOW_SPEC_AUTO_A = 1, CAPOW_SPEC_AUTO_B = 2 };

/* Run parameters: the periodic subset of the reference's 41 tags
 * (src/read_params.c:65-227, stored in src/vars.h:70-134) plus the
 * arithmetic form and edge rules. */
typedef struct capow_params {
  int64_t nx, ny, nz;                  /* NX NY NZ                        */
  double xmin, xmax, ymin, ymax, zmin, zmax;
  int32_t interp_order;                /* InterpOrder: 1 NGP, 2 CIC, 3 TSC */
  int32_t do_interlacing;              /* DoInterlacing 0/1               */
  int32_t los;                         /* LOS 0 (real space, mu along z), 1..3 */
  int32_t momentum;                    /* Momentum 0..8                   */
  int32_t nk;                          /* NK                              */
  int32_t nmu;                         /* NMU (used when output_2d)       */
  int32_t output_2d;                   /* Output2D 0/1                    */
  int32_t output_log;                  /* OutputLog 0/1                   */
  int32_t odd_multipoles;              /* Odd_Multipoles 0/1              */
  double mink, maxk;                   /* Mink Maxk                       */
  double omega_m, redshift;            /* Omega_m Redshift (RSD shift)    */
  int32_t precision;                   /* CAPOW_FP32 | CAPOW_FP64         */
  int32_t edge_mode;                   /* CAPOW_EDGE_*                    */
  int32_t extra_spectra;               /* OR of CAPOW_SPEC_*              */
  int32_t reserved;
} capow_params;

/* Global sums of the assignment (reference src/read_data.c:132-138 and the
 * NGRID/NOUT counters of src/read_data.c:37-88). */
typedef struct capow_assign_stats {
  double nread;        /* particles handed in                         */
  double nout;         /* LOS != 0: outside the box after RSD         */
  double ngrid;        /* sum of w over main-grid-accepted particles  */
  double nsq;          /* sum w over all particles (Momentum != 0)     */
  double vr, vr2, vr3, vr4;  /* sum w v^n over all particles           */
  double alpha;        /* ngrid / (NX NY NZ)                           */
} capow_assign_stats;

/* Raw (un-normalised) per-bin sums of one binning pass, in host memory
 * supplied by the caller.  For spectrum s (0 = primary, then AUTO_A, then
 * AUTO_B when requested) and multipole l = 0..4:
 *   pk[(s*5 + l)*nk + kbin]   (l = 1, 3 only filled with odd_multipoles)
 *   nmodes[kbin]              (int64; identical for every spectrum)
 *   pk2d[s*nk*nmu + kbin*nmu + mubin], nmodes2d[kbin*nmu + mubin]
 * as in src/main.c:432-461.  The spectrum codes (reference Momentum value
 * each spectrum equals) are written to codes[0..n_spectra). */
typedef struct capow_bins {
  int32_t n_spectra;
  int32_t codes[3];
  double *pk;          /* n_spectra*5*nk                               */
  int64_t *nmodes;     /* nk                                           */
  double *pk2d;        /* n_spectra*nk*nmu, or NULL                    */
  int64_t *nmodes2d;   /* nk*nmu, or NULL                              */
} capow_bins;

typedef struct capow_ctx capow_ctx;

/* ---- lifecycle --------------------------------------------------------- */

/* Validates p as check_inputs() does (src/read_params.c:294-428), allocates
 * the device grids in the reference's padded layout
 * [NX][NY][2(NZ/2+1)] (src/main.c:1123-1217; no buffer slices -- the
 * deposit wraps directly) and builds FFT and binning tables.
 * Replaces create_grids() (src/main.c:1123). */
int capow_create(const capow_params *p, capow_ctx **out);
int capow_destroy(capow_ctx *ctx);

/* Message of the last failure of ctx (or of capow_create when ctx = NULL). */
const char *capow_last_error(const capow_ctx *ctx);

/* Work on a caller-supplied hipStream_t (NULL = default stream). */
int capow_set_stream(capow_ctx *ctx, void *hip_stream);
int capow_synchronize(capow_ctx *ctx);

/* Record HIP events around every kernel launch (0 = off, the default). */
int capow_set_timing(capow_ctx *ctx, int enable);

/* ---- stages ------------------------------------------------------------ */

/* Particles as n rows of (x, y, z, v_los): 4 values per particle of type
 * dtype, already with the LOS velocity column selected as
 * src/read_data.c:65-76 does (vz for LOS 0 and 3, vx for 1, vy for 2).
 * w: n weights of the same dtype, or NULL for unit weights (the reference
 * hard-codes w = 1, src/read_data.c:60).  where = CAPOW_HOST copies to the
 * device; CAPOW_DEVICE uses the pointers in place (they must stay valid
 * until capow_assign returns).  dtype must match the context's precision
 * (F32 for CAPOW_FP32, F64 for CAPOW_FP64). */
int capow_set_particles(capow_ctx *ctx, const void *xyzv, const void *w,
                        int64_t n, int32_t dtype, int32_t where);

/* Mass assignment of every field the reference's `momentum` needs, with
 * interlacing, followed by the mean subtraction; replaces the deposit loop
 * of read_periodic_serial_ascii (src/read_data.c:56-175) and
 * src/main.c:103-148 + 211-225.  The RSD shift of src/read_data.c:67-76 is
 * applied on the device for LOS 1..3.  stats may be NULL. */
int capow_assign(capow_ctx *ctx, capow_assign_stats *stats);

/* In-place forward r2c 3-D FFT of every deposited grid, unnormalised,
 * e^{-ik.x}; replaces fftw_execute at src/main.c:266-271. */
int capow_fft(capow_ctx *ctx);

/* Fused deconvolution, interlace combine, power, Legendre weighting and
 * (k, mu) binning; replaces compute_periodic_power (src/main.c:302-468).
 * Fills out (host memory) with raw sums; synchronises. */
int capow_bin(capow_ctx *ctx, capow_bins *out);

/* capow_assign + capow_fft + capow_bin. */
int capow_power(capow_ctx *ctx, capow_assign_stats *stats, capow_bins *out);

/* Survey (Yamamoto) multipole estimator on the context's real-space grids
 * (filled by capow_assign or capow_set_grid, not transformed); replaces
 * compute_survey_power and assign_survey_power (src/main.c:472-967) for the
 * grids the survey ingest deposits.  origin = X_Origin, Y_Origin, Z_Origin.
 * For each multipole the grids are weighted by the r^ products of degree l,
 * transformed and paired with the monopole's k-space field times the k^
 * products.  pk (host, [5][nk]) receives the raw sums Pk0..Pk4 (Pk1, Pk3
 * zero without Odd_Multipoles), nmodes (host, nk) the mode counts; Output2D
 * bins stay empty, as in the reference's survey branch.  Momentum 2 and 5
 * are CAPOW_ERR_UNSUPPORTED (the reference has no ddg_mom grid for them);
 * Momentum 1 uses the density grids only (src/main.c:519).  The grids hold
 * the same real-space fields afterwards.  Synchronises. */
int capow_survey_power(capow_ctx *ctx, const double origin[3], double *pk, int64_t *nmodes);

/* No counterpart in the reference.  capow_bin for callers that cannot
 * allocate device memory (code under stream capture, tight serving loops,
 * harnesses that forbid hipMalloc): the same binning kernels on k-space
 * grids the caller owns, with no context.  p: the parameters (validated as
 * capow_create does; survey parameters are refused).  grids: device
 * pointers, slots as capow_get_grid's (0 A, 1 A interlaced, 2 B, 3 B
 * interlaced; the slots p does not use are ignored), each [NX][NY][row_pitch]
 * complex values of the precision of p, 256-byte aligned, holding the r2c
 * half spectrum in the first NZ/2+1 of every row (the layout capow_fft
 * leaves); row_pitch >= NZ/2+1 and a whole number of 128-byte lines (a
 * multiple of 16 in FP32, of 8 in FP64).  hip_stream: the hipStream_t the
 * work runs on (NULL: the default stream).  out: host arrays as for
 * capow_bin, filled with the same raw sums and mode counts; synchronises
 * the stream.  Every buffer the kernels need (tables, partial bins, counts)
 * lives in one device array of 64 MiB per GPU that the library holds until
 * the process ends; bins that need more are CAPOW_ERR_UNSUPPORTED.  Calls
 * that overlap on one GPU share that array and race.  Errors are reported
 * by capow_last_error(NULL). */
int capow_bin_noalloc(const capow_params *p, const void *const grids[4], int64_t row_pitch,
                      void *hip_stream, capow_bins *out);

/* Number of spectra capow_bin produces and their reference Momentum codes. */
int capow_spectra(const capow_params *p, int32_t *n_spectra, int32_t codes[3]);

/* ---- adjoint ------------------------------------------------------------ */
/* No counterpart in the reference (it computes no derivatives).  Given
 * dL_draw = dL/d(raw bins) in capow_bins.pk's layout ([n_spectra][5][nk],
 * host array; entries of multipoles not computed are ignored), recomputes
 * the assignment and FFT of the particles last set with capow_set_particles
 * and writes dL/d(x, y, z, v_los) as n x 4 rows into grad_xyzv and dL/dw
 * into grad_w (NULL: not wanted), both in the context precision and on the
 * host or the device (`where`).  The gradient covers the deposit (fields
 * w v^a, RSD shift of the line-of-sight coordinate), the mean subtraction,
 * the r2c FFT and the binning; the host-side normalisation (shot noise,
 * N^2/V) is the caller's to differentiate.  2-D bins carry no gradient.
 * Register FFT sizes only (CAPOW_ERR_UNSUPPORTED otherwise).  Afterwards the
 * grids hold gradients (capow_assign must run again before capow_fft). */
int capow_backward(capow_ctx *ctx, const double *dL_draw, void *grad_xyzv,
                   void *grad_w, int32_t where);
/* Forward kept for the backward (no counterpart in the reference): with
 * on = 1, capow_assign also records each sorted row's input index (4 bytes
 * a row), and capow_backward reuses the forward's sort and transform
 * instead of recomputing them while the grids still hold the transform
 * (capow_fft) of the assignment (capow_assign) of the particles last set:
 * capow_set_particles, capow_set_grid, capow_mark_assigned,
 * capow_mark_transformed and capow_backward itself end that.  A caller that
 * rewrites the particles in place between the forward and the backward
 * sets them again (and the backward then recomputes).  The gradients are
 * bitwise those of the recomputing backward.  Off by default. */
int capow_keep_forward(capow_ctx *ctx, int32_t on);
/* Whether the last capow_backward reused the forward (1) or recomputed it. */
int capow_backward_reused(const capow_ctx *ctx, int32_t *reused);

/* ---- batches of catalogues (mock runs) --------------------------------- */
/* No counterpart in the reference, which measures one catalogue per
 * process run (src/main.c:28-298); a batch runs many catalogues through the
 * same stages (capow_assign, capow_fft, capow_bin) without host
 * synchronisation between them: catalogue i goes to stream i % nstreams,
 * each stream owning one context per parameter set. */

typedef struct capow_batch capow_batch;

typedef struct {
  const void *xyzv;    /* device rows x y z v_los (the configs' precision) */
  const void *w;       /* device weights, or NULL for unit weights         */
  int64_t n;           /* particles                                        */
  int32_t config;      /* index of the parameter set                       */
  int32_t reserved;
} capow_mock;

/* configs: nconf parameter sets (e.g. two grid sizes); max_particles: the
 * largest catalogue (sizes the per-context sort buffers: 2 copies per
 * particle, a catalogue that needs more is re-run alone); nstreams >= 1
 * concurrent streams (default priority). */
int capow_batch_create(const capow_params *configs, int32_t nconf,
                       int64_t max_particles, int32_t nstreams,
                       capow_batch **out);
/* Runs n catalogues.  out[i] (caller-owned arrays, as for capow_bin) gets
 * mock i's raw bins; stats[i] (or NULL) its assignment sums; ms[i] (or
 * NULL) the GPU time from its first to its last stage on its stream.
 * Returns the first failing mock's status (all mocks are attempted). */
int capow_batch_run(capow_batch *b, const capow_mock *mocks, int32_t n,
                    capow_bins *out, capow_assign_stats *stats, float *ms);
/* Graph mode (no counterpart in the reference; for callers that run the
 * same catalogue list many times, e.g. a covariance loop over mocks whose
 * contents are rewritten in place): on = 1 makes capow_batch_run capture a
 * list's launches on every stream into one HIP graph after running it once
 * eagerly, and replay that graph whenever the next list is the same (the
 * same device pointers, counts and parameter sets, in the same order, and
 * the same streams in use); a different list runs eagerly and is captured
 * in turn.  Replayed runs read the catalogues' current contents; ms[i] is
 * then -1.  A catalogue re-run alone (sort buffer too small) or an
 * autotune drops the graph.  on = 0 (the default) drops it.  Synchronises. */
int capow_batch_set_graph(capow_batch *b, int32_t on);
/* Stacked mode (no counterpart in the reference): m >= 1 runs the
 * catalogues of each parameter set (and of the same weighting: all weighted
 * or all unit weights) in stacks of up to m, every stage one launch for
 * the whole stack -- the sort, the statistics, the deposit, the z and y
 * passes and the fused x pass + binning, each kernel taking the catalogue
 * from its grid's second index -- on grids laid out catalogue after
 * catalogue.  Each set's stack runs on stream (set % streams in use).  With
 * the same settings (the stack context starts from the set's stream
 * contexts'; capow_batch_autotune may re-tune it) the bins and sums are
 * bitwise those of the per-catalogue path; a different sort block count
 * changes only the rounding of the FP64 moment sums (and, for weighted
 * catalogues, of alpha).  Needs the fused
 * binning (FP32 grids, 1-D bins, register FFT lengths) for every set,
 * else CAPOW_ERR_UNSUPPORTED and the mode is unchanged; m is capped so
 * that m times max_particles stays below 2^32 - 4096 sorted rows.  ms[i]
 * is then the GPU time of mock i's stack.  m = 0 (the default) returns to
 * one catalogue per launch.  Allocates the stacks' buffers; synchronises. */
int capow_batch_set_stack(capow_batch *b, int32_t m);
int capow_batch_destroy(capow_batch *b);
const char *capow_batch_last_error(const capow_batch *b);

/* ---- autotuner ------------------------------------------------------------ */
/* No counterpart in the reference, whose only planning is FFTW_ESTIMATE
 * (src/main.c:1200-1217).  capow_autotune times the pipeline (capow_assign,
 * capow_fft, capow_bin) on the particles last set with capow_set_particles,
 * trying the candidates of one tunable at a time and keeping the fastest
 * (median of `reps` runs, <= 0: 3): the sort's workgroups per CU and its LDS
 * run aggregation, the deposit's halo rows, the FFT passes' write-run width
 * and batch order, and the x pass fused with the binning.  The choices stay
 * with the context; `report` (may be NULL) receives every timing and the
 * choice as text.  The results do not depend on the choices, except that
 * the fused binning's sums differ from the separate binning's at the FP32
 * rounding level.  Synchronises. */
int capow_autotune(capow_ctx *ctx, int32_t reps, char *report, int32_t report_len);
/* The same for a batch: each parameter set is tuned on its first catalogue
 * in `mocks` and the choice copied to every stream's context of that set;
 * then the number of streams in use (2, 4, ... up to the batch's) is timed
 * on the whole list, then stacks of 16 and 64 catalogues
 * (capow_batch_set_stack) against the per-catalogue streams; the fastest
 * mode stays set, and with stacks each set's stack context is tuned again
 * (sort blocks, LDS run aggregation, halo rows, bank order, FFT write runs
 * and batch order), starting from the library defaults and timing the
 * whole list (the sets' stacks share the GPU). */
int capow_batch_autotune(capow_batch *b, const capow_mock *mocks, int32_t n, int32_t reps,
                         char *report, int32_t report_len);

/* ---- catalogue input (no GPU) ------------------------------------------ */

/* Reads the periodic ASCII catalogue the way read_periodic_serial_ascii
 * parses it (src/read_data.c:32-88): fgets pieces of at most 1999
 * characters, '#' pieces skipped, six numbers per piece ("x y z vx vy vz",
 * extra columns ignored), v_los = vx, vy or vz for LOS 1, 2, else vz; with
 * LOS = 0 a particle outside [lo, hi] is fatal.  Parsed by nthreads threads
 * (<= 0: OMP_NUM_THREADS if set, else the CPUs the process may run on; at
 * most 16) over line-aligned ranges of the mapped file; the result and the
 * reported error do not depend on the thread count.
 * On success *rows holds n rows of (x, y, z, v_los) in float64 (fp64) or
 * float32, allocated by the library: release them with capow_free_rows.
 * Errors: CAPOW_ERR_IO (cannot open), CAPOW_ERR_PARSE (the first piece in
 * file order without six numbers, copied to bad_line), CAPOW_ERR_OUT_OF_BOX
 * (LOS = 0: the first particle outside the box, its position in bad_xyz and
 * its piece in bad_line), CAPOW_ERR_NOMEM.  bad_line / bad_xyz may be NULL. */
int capow_read_ascii(const char *path, int32_t los, const double lo[3], const double hi[3],
                     int32_t fp64, int32_t nthreads, void **rows, int64_t *n,
                     char *bad_line, int32_t bad_len, double bad_xyz[3]);
void capow_free_rows(void *rows);

/* Binary sidecar of the periodic catalogue (no counterpart in the
 * reference, which parses the text on every run: src/read_data.c:56-64).
 * capow_write_sidecar parses the text file at text_path with
 * capow_read_ascii's line rules (nthreads as there) and writes bin_path: a
 * 64-byte header ("CAPOWCAT", version 1, 6 columns, the row count, the text
 * file's size and modification time) and the six numbers of every line as
 * little-endian doubles in file order; written beside bin_path and renamed
 * over it.  Errors: CAPOW_ERR_IO, CAPOW_ERR_PARSE (bad_line as for
 * capow_read_ascii), CAPOW_ERR_NOMEM.
 * capow_read_sidecar returns what capow_read_ascii returns on the text --
 * the same rows bit for bit for any LOS, box and precision, the same first
 * particle outside the box (LOS = 0) -- from the sidecar alone.  With
 * text_path (may be NULL) it refuses a sidecar whose text file's size or
 * modification time differ from those it recorded: CAPOW_ERR_STALE.  A
 * file that is not a whole sidecar: CAPOW_ERR_PARSE.  Release the rows with
 * capow_free_rows. */
int capow_write_sidecar(const char *text_path, const char *bin_path, int32_t nthreads,
                        char *bad_line, int32_t bad_len);
int capow_read_sidecar(const char *bin_path, const char *text_path, int32_t los,
                       const double lo[3], const double hi[3], int32_t fp64, int32_t nthreads,
                       void **rows, int64_t *n, double bad_xyz[3]);

/* ---- survey ingest (no GPU) -------------------------------------------- */

/* One kept survey object (the reference's struct survey_data,
 * src/vars.h): Cartesian position after the Coord_Type conversion,
 * redshift, weight (FKP-weighted after capow_survey_fkp), weight_pv, nbar,
 * pv and pverr (both zero unless Momentum != 0; pv zero for randoms, whose
 * file holds no logdist column). */
typedef struct capow_survey_obj {
  double x, y, z, redshift, weight, weight_pv, nbar, pv, pverr;
} capow_survey_obj;

/* The survey tags of run_parameters.dat (src/read_params.c). */
typedef struct capow_survey_params {
  int32_t coord_type;      /* Coord_Type: 0 Cartesian, 1 RA/Dec degrees, else radians */
  int32_t momentum;        /* Momentum (survey code)                       */
  int32_t nbar_column;     /* NBAR_Column: > 0 keeps the file's nbar        */
  int32_t fkp_column;      /* FKP_Column: < 0 computes FKP weights          */
  double omega_m;          /* Omega_m                                       */
  double redmin_in, redmax_in; /* REDMININ REDMAXIN                          */
  double sky_area;         /* SkyArea (deg^2)                               */
  double fkp_pk, fkp_pk_mom, fkp_pk_cross;  /* FKP_Pk FKP_Pk_mom FKP_Pk_cross */
} capow_survey_params;

/* Sums of src/main.c:149-204 (randoms' sums already scaled by alpha). */
typedef struct capow_survey_totals {
  double data_nbw, data_nbwsq, data_nbsqwsq, data_nbsqwsq_pv, data_vr, data_vrsq;
  double rand_nbw, rand_nbwsq, rand_nbsqwsq, rand_nbsqwsq_pv;
  double alpha, shot, norm;
} capow_survey_totals;

/* Natural cubic spline (gsl_interp_cspline in the reference): m receives the
 * n second derivatives; eval returns CAPOW_ERR_ARG outside [x0, x(n-1)],
 * where GSL reports a domain error. */
int capow_cspline_init(int32_t n, const double *x, const double *y, double *m);
int capow_cspline_eval(int32_t n, const double *x, const double *y, const double *m, double xq,
                       double *out);

/* r(z) = c/100 int_0^z dz'/sqrt(Om (1+z')^3 + 1 - Om) in Mpc/h at n nodes
 * z_i = i zmax/(n-1); replaces comoving_distance (src/read_data.c:363-378)
 * over the table of create_grids (src/main.c:1230-1242). */
int capow_zr_table(double omega_m, int32_t n, double zmax, double *z, double *r);

/* The table at 10000 nodes on [0, 10] with its two splines, r(z) and z(r)
 * (src/main.c:1230-1245: dist_spline, red_spline). */
typedef struct capow_cosmo capow_cosmo;
int capow_cosmo_create(double omega_m, capow_cosmo **out);
int capow_cosmo_destroy(capow_cosmo *c);
int capow_cosmo_r_of_z(const capow_cosmo *c, double z, double *r);
int capow_cosmo_z_of_r(const capow_cosmo *c, double r, double *z);

/* Reads a survey data (randoms = 0: twelve numbers a line) or randoms file
 * (eight or more) the way read_survey_serial_ascii does
 * (src/read_data.c:184-361): leading '#' lines skipped, nbar from the fifth
 * column, the redshift cut [REDMININ, REDMAXIN] (Coord_Type 0 through z(r)),
 * zero-weight objects dropped for Coord_Type != 0, (RA, Dec, z) converted to
 * Cartesian, and an object outside the grid shrunk by the interpolation
 * reach fatal (CAPOW_ERR_OUT_OF_BOX; bad = x, y, z, redshift, bad_line =
 * the line).  A malformed line is CAPOW_ERR_PARSE, more than nobj_max kept
 * objects CAPOW_ERR_NOMEM, a distance or redshift outside the z-r table
 * CAPOW_ERR_ARG.  zrange (in/out) widens to the kept redshifts (the
 * reference's REDMIN, REDMAX); xyz_range (or NULL) gets the X, Y, Z
 * low/high pairs the reference prints.  *rows is released with
 * capow_free_survey. */
int capow_read_survey(const char *path, int32_t randoms, const capow_params *g,
                      const capow_survey_params *sp, const capow_cosmo *cosmo, int64_t nobj_max,
                      capow_survey_obj **rows, int64_t *n_keep, int64_t *n_read, double zrange[2],
                      double xyz_range[6], char *bad_line, int32_t bad_len, double bad[4]);
void capow_free_survey(capow_survey_obj *rows);

/* compute_nbar (src/read_data.c:383-440): a histogram of the first ndata
 * randoms' redshifts in ceil((REDMAXIN - REDMININ)/0.002) bins over zrange,
 * divided by the shell volumes over SkyArea, a natural spline through the
 * bin centres with zero end points, evaluated at every object.  znbar and
 * nbar (capacity cap >= capow_survey_nbar_bins + 2) receive the table. */
int32_t capow_survey_nbar_bins(const capow_survey_params *sp);
int capow_survey_nbar(const capow_survey_params *sp, const capow_cosmo *cosmo, const double zrange[2],
                      capow_survey_obj *data, int64_t ndata, capow_survey_obj *randoms, int64_t nrand,
                      double *znbar, double *nbar, int32_t cap);

/* compute_fkp (src/read_data.c:444-463), in place. */
int capow_survey_fkp(const capow_survey_params *sp, capow_survey_obj *o, int64_t n);

/* alpha, shot noise and normalisation (src/main.c:149-204). */
int capow_survey_sums(const capow_survey_params *sp, const capow_survey_obj *data, int64_t ndata,
                      const capow_survey_obj *randoms, int64_t nrand, capow_survey_totals *t);

/* Rows for capow_set_particles on a CAPOW_EDGE_SURVEY context: the data,
 * then (Momentum != 1) the randoms with weight -alpha w
 * (assign_survey_data, src/read_data.c:464-486, src/main.c:94, 228).
 * xyzv (4 per row) and w (1 per row) have room for ndata + nrand rows of
 * dtype. */
int capow_survey_pack(const capow_survey_params *sp, const capow_survey_obj *data, int64_t ndata,
                      const capow_survey_obj *randoms, int64_t nrand, double alpha, int32_t dtype,
                      void *xyzv, void *w, int64_t *n_out);

/* ---- host-side output math (no GPU) ------------------------------------ */

/* shot and norm for spectrum code `code` from the assignment sums
 * (src/main.c:103-148 at one rank). */
int capow_scalars(const capow_params *p, const capow_assign_stats *st,
                  int32_t code, double *shot, double *norm);

/* In place on one spectrum's arrays (pk0..pk4 = 5 consecutive nk-vectors,
 * nmodes, and the optional 2-D arrays): shot subtraction on P0,
 * (2l+1)/(Nmodes norm) scaling, and zeroing of the highest filled bin
 * (src/main.c:1014-1056).  Returns the zeroed bin index via *last_bin
 * (-1 if none). */
int capow_finalize(const capow_params *p, double shot, double norm,
                   double *pk5, int64_t *nmodes, double *pk2d,
                   int64_t *nmodes2d, int32_t *last_bin);

/* Writes the reference's text file(s) (src/main.c:1058-1101): "%g" rows,
 * and `path`_2D when output_2d.  precise != 0 uses "%.17g" (a parity side
 * dump; not the reference format). */
int capow_write_output(const capow_params *p, double shot, double norm,
                       const double *pk5, const int64_t *nmodes,
                       const double *pk2d, const int64_t *nmodes2d,
                       const char *path, int32_t precise);

/* ---- inspection (parity tests, dumps) ---------------------------------- */

/* Grid slots: 0 = A (ddg), 1 = A interlaced, 2 = B (ddg_mom),
 * 3 = B interlaced.  Element type is the context precision; the buffer is
 * the padded [NX][NY][2(NZ/2+1)] array (real before capow_fft, complex
 * after).  n = number of elements the host buffer holds. */
int capow_grid_elems(const capow_ctx *ctx, int64_t *n);
int capow_get_grid(capow_ctx *ctx, int32_t slot, void *host, int64_t n);
int capow_set_grid(capow_ctx *ctx, int32_t slot, const void *host, int64_t n);
/* Marks the grids as deposited (after capow_set_grid) with the given
 * assignment sums, so capow_fft / capow_bin may run on uploaded grids. */
int capow_mark_assigned(capow_ctx *ctx, const capow_assign_stats *st);
/* Marks the grids as transformed (after uploading k-space grids). */
int capow_mark_transformed(capow_ctx *ctx);

/* Per-kernel GPU times of the last stages run with timing on: names
 * (NUL-separated into `names`, up to names_len bytes), total ms and launch
 * counts, up to max entries; returns the number of entries via *count. */
int capow_kernel_times(const capow_ctx *ctx, char *names, int32_t names_len,
                       double *ms, int32_t *launches, int32_t max,
                       int32_t *count);
int capow_reset_kernel_times(capow_ctx *ctx);

#ifdef __cplusplus
}
#endif
#endif
