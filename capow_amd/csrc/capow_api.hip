// capow_api.hip -- the C ABI of include/capow.h: context, validation,
// device buffers, tables and stage launches.
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <stdarg.h>

#include <algorithm>
#include <chrono>
#include <string>
#include <vector>

#include "capow.h"
#include "common.h"
#include "kernels.h"

using namespace capow;

struct capow_ctx {
  capow_params p;
  Geom g;
  bool fp64 = false;
  size_t esz = 4;
  hipStream_t stream = nullptr;
  Timer timer;
  std::string err;
  int state = 0;  // 0 created, 1 particles, 2 assigned, 3 transformed
  // grids: slots 0 A, 1 A_int, 2 B, 3 B_int
  void* grid[4] = {nullptr, nullptr, nullptr, nullptr};
  // out-of-place FFT: a second buffer per grid (swapped with the grid after
  // the transform); null when the in-place passes are used
  void* scratch[4] = {nullptr, nullptr, nullptr, nullptr};
  int64_t fft_px = 0, fft_py = 0;   // row pitches of the intermediate layouts
  // one spare buffer (scratch[0]) instead of one per grid: the grids are too
  // large for a second set (e.g. four FP32 grids at 2048^3); the FFT passes
  // then run grid by grid and rotate the buffers
  bool fft_one_spare = false;
  int fft_persist = 1;              // persistent row passes (CAPOW_FFT_PERSIST=0: off)
  bool dep_halo = true;             // halo-row deposit (CAPOW_DEPOSIT_HALO=0: off)
  bool dep_bank = true;             // its bank-balanced copy order (CAPOW_DEPOSIT_BANK=0: off)
  // batch order of the row passes, bit per pass (z, y, x): outer index
  // fastest (FP32 default z and x: 128-byte write runs of neighbouring
  // outer rows land close together; measured 45.7 -> 42.7-44.9 ms at 1024^3;
  // FP64: none); CAPOW_FFT_ORDER overrides
  int fft_order = -1;
  int ncu = 256;                    // compute units of the context's device
  int fft_rows_mul = 0;             // CAPOW_FFT_ROWS_MUL=1|2 forces 64|128-byte write runs
  int64_t grid_elems = 0;
  size_t grid_bytes = 0;            // capacity of every grid and scratch buffer
  // particles
  const void* parts = nullptr;
  const void* wts = nullptr;
  int64_t n = 0;
  void* own_parts = nullptr;
  size_t own_parts_cap = 0;
  void* own_w = nullptr;
  size_t own_w_cap = 0;
  void* sorted = nullptr;
  size_t sorted_cap = 0;
  void* sorted_w = nullptr;
  size_t sorted_w_cap = 0;
  // buckets / sort
  uint32_t *counts = nullptr, *offsets = nullptr, *cursor = nullptr, *bsum = nullptr,
           *boff = nullptr;
  int count_blocks = 3072;   // bucket-kernel blocks: 12 per CU (measured best at C2)
  double* partials = nullptr;
  unsigned int* errs = nullptr;
  double* stats = nullptr;
  // fft
  FftPlan plan[3];
  void* tw[3] = {nullptr, nullptr, nullptr};
  // binning
  BinArgsDev bin;
  double* bin_tabs = nullptr;
  double* bin_partials = nullptr;
  double* bin_out = nullptr;
  double* bin_counts = nullptr;     // mode counts of the branch-free binning
  bool bin_mirror2d = false;        // (k, mu) bins in the mirror form
  long long* bin_partials2 = nullptr;   // their int64 workgroup tables
  double* bin_qscale = nullptr;     // their per-k quanta
  bool bin_counts_valid = false;
  int bin_nwg = 1024;
  int bin_per_copy = 0;
  bool bin_row_form = true;
  bool bin_mirror = false;  // mirror-group binning (set up in setup())
  bool bin_sym = false;     // symmetric x and y tables (the survey's mirror groups)
  int bin_mirror_env = -1;  // CAPOW_BIN_MIRROR: 0 off, 1 on where it applies
  // x pass fused with the binning (FP32 out-of-place FFT, mirror form, 1-D
  // bins): capow_fft runs the z and y passes and leaves the x pass pending;
  // capow_bin transforms and bins at once; anything else that reads the
  // k-space grids runs the pending pass first (finish_x)
  bool xbin_ok = false;
  int xbin_env = -1;        // CAPOW_XBIN: 0 off
  bool x_pending = false;
  double xbin_factor = 0.0; // 4 cor_max NX NY
  double* xbin_ssq = nullptr;
  double* xbin_q = nullptr;
  long long* xbin_part = nullptr;
  int xbin_rows = 0;
  int xbin_ssq_rows = 0;    // rows of xbin_ssq the last deferred transform wrote
  int nspec = 1;
  int codes[3] = {0, 0, 0};
  capow_assign_stats last{};
  // pinned host staging of the stats, error flags and bins of the last call
  double* h_stats = nullptr;
  unsigned int* h_errs = nullptr;
  double* h_bins = nullptr;
  bool cap_fixed = false;   // sorted buffer capacity fixed (batch); overflow is flagged
  // backward buffers
  double* bwd_G = nullptr;
  size_t bwd_G_cap = 0;
  double* bwd_g0 = nullptr;
  void* bwd_out = nullptr;
  size_t bwd_out_cap = 0;
  void* bwd_w = nullptr;
  size_t bwd_w_cap = 0;
  // input index of every sorted row, written by the sort while a backward
  // recomputes the assignment (the tile gather writes gradients back there)
  uint32_t* sorted_idx = nullptr;
  size_t sorted_idx_cap = 0;
  bool want_idx = false;
  unsigned int* bwd_err = nullptr;   // tile gather: a tap outside its region
  bool bwd_tiles = true;             // tile gather (CAPOW_GATHER_TILES=0: per particle)
  // capow_keep_forward: the forward's sort records the rows' input indices
  // and capow_backward reuses the forward's sort and transform while
  // fwd_kept holds: the grids are the transform (capow_fft) of the
  // deposit of the current particles (fwd_assigned: the last assignment
  // came from capow_assign; fwd_idx: it recorded the indices), unchanged
  // since
  bool keep_fwd = false;
  bool fwd_assigned = false, fwd_idx = false, fwd_kept = false;
  bool bwd_reused = false;           // the last backward reused the forward
  int fft_col_mul = 2;   // column batch multiplier of the y/x FFT passes (CAPOW_FFT_COL_MUL=1 halves it)
  // survey estimator: copies of the real-space fields, the monopole k-space
  // fields, bin partials and sums (allocated on first use)
  void* sv_src[4] = {nullptr, nullptr, nullptr, nullptr};
  void* sv_F0 = nullptr;
  void* sv_F0m = nullptr;
  double* sv_partials = nullptr;
  double* sv_acc = nullptr;
  double* h_sv = nullptr;
  int sv_nwg = 0;
  // stacked batch context: nstack catalogues of this parameter set share
  // one launch per stage; every per-catalogue buffer (grids, sort counters,
  // sums, flags, bins, the fused binning's rows) holds nstack copies
  int nstack = 1;
  MockDev* stack_tab = nullptr;     // device table of the catalogues of a run
  MockDev* h_stack_tab = nullptr;   // its pinned staging
  int64_t stack_rows = 0;           // sorted rows (the sum of the catalogue sizes)
};

static std::string g_create_err;

#define HIPCHK(ctx, call)                                                          \
  do {                                                                             \
    hipError_t _e = (call);                                                        \
    if (_e != hipSuccess) {                                                        \
      (ctx)->err = std::string(#call) + ": " + hipGetErrorString(_e);              \
      return CAPOW_ERR_HIP;                                                        \
    }                                                                              \
  } while (0)

static int fail(capow_ctx* c, int code, const std::string& m) {
  if (c) c->err = m;
  else g_create_err = m;
  return code;
}

// Field exponents per Momentum (src/read_data.c:91-130): A = w v^a, B = w v^b.
static void field_exps(int m, int* a, int* b) {
  static const int ta[9] = {0, 0, 1, 0, 1, 2, 0, 1, 0};
  static const int tb[9] = {-1, 1, -1, 2, 2, -1, 3, 3, 4};
  *a = ta[m];
  *b = tb[m];
}

static int auto_code(int e) { return e == 0 ? 0 : (e == 1 ? 2 : (e == 2 ? 5 : -1)); }

extern "C" int capow_spectra(const capow_params* p, int32_t* n_spectra, int32_t codes[3]) {
  if (!p || p->momentum < 0 || p->momentum > 8) return CAPOW_ERR_ARG;
  int a, b;
  field_exps(p->momentum, &a, &b);
  int n = 0;
  codes[n++] = p->momentum;
  if (b >= 0 && (p->extra_spectra & CAPOW_SPEC_AUTO_A)) codes[n++] = auto_code(a);
  if (b >= 0 && b <= 2 && (p->extra_spectra & CAPOW_SPEC_AUTO_B)) codes[n++] = auto_code(b);
  *n_spectra = n;
  return CAPOW_OK;
}

static int spec_kind(int code, int momentum) {
  // the kind of combination a spectrum code needs from the (A, B) fields
  if (code == momentum) {
    if (code == 0 || code == 2 || code == 5) return SPEC_AUTO_A;
    if (code == 1 || code == 4 || code == 6) return SPEC_CROSS_IM;
    return SPEC_CROSS_RE;
  }
  int a, b;
  field_exps(momentum, &a, &b);
  return code == auto_code(a) ? SPEC_AUTO_A : SPEC_AUTO_B;
}

static bool make_plan(int64_t M, FftPlan* pl) {
  pl->npass = 0;
  static const int rad[5] = {8, 4, 2, 3, 5};
  for (int q = 0; q < 5; q++) {
    while (M % rad[q] == 0 && M > 1) {
      if (pl->npass >= 24) return false;
      pl->radix[pl->npass++] = rad[q];
      M /= rad[q];
    }
  }
  return M == 1;
}

// Largest tile extent <= tmax dividing n (at least 4 unless n itself fits);
// `even` for the z extent, whose cells are written in pairs.
static int pick_tile(int n, int tmax, bool even = false) {
  if (n <= tmax && (!even || n % 2 == 0)) return n;
  for (int t = tmax; t >= 4; t--)
    if (n % t == 0 && (!even || t % 2 == 0)) return t;
  return -1;
}

// check_inputs (src/read_params.c:294-428) for the periodic tags, plus the
// shapes the kernels handle.
static int validate(const capow_params* p, std::string& m);

namespace capow {
int validate_params(const capow_params* p, std::string& m) { return validate(p, m); }
void set_global_error(const std::string& m) { g_create_err = m; }
}  // namespace capow

static int validate(const capow_params* p, std::string& m) {
  if (p->los < 0 || p->los > 3) { m = "`LOS' is neither 0 (real-space), 1, 2 nor 3."; return CAPOW_ERR_ARG; }
  if (p->interp_order < 1 || p->interp_order > 3) { m = "`InterpOrder' is neither 1, 2 nor 3."; return CAPOW_ERR_ARG; }
  if (p->do_interlacing != 0 && p->do_interlacing != 1) { m = "`DoInterlacing' is neither 0 nor 1."; return CAPOW_ERR_ARG; }
  if (p->output_2d != 0 && p->output_2d != 1) { m = "`Output2D' is neither 0 nor 1."; return CAPOW_ERR_ARG; }
  if (p->output_2d && p->nmu <= 0) { m = "`NMU' is less than 1."; return CAPOW_ERR_ARG; }
  if (p->output_log != 0 && p->output_log != 1) { m = "`OutputLog' is neither 0 nor 1."; return CAPOW_ERR_ARG; }
  if (p->nx <= 0 || p->ny <= 0 || p->nz <= 0) { m = "`NX', `NY' or `NZ' is less than 1."; return CAPOW_ERR_ARG; }
  if (p->nk <= 0) { m = "`NK' is less than 1."; return CAPOW_ERR_ARG; }
  if (p->xmax < p->xmin || p->ymax < p->ymin || p->zmax < p->zmin) { m = "`XMAX' is less than `XMIN'."; return CAPOW_ERR_ARG; }
  if (p->maxk < p->mink) { m = "`Maxk' is less than `Mink'."; return CAPOW_ERR_ARG; }
  if (p->momentum < 0 || p->momentum > 8) { m = "`Momentum' must be 0..8."; return CAPOW_ERR_ARG; }
  if (p->precision != CAPOW_FP32 && p->precision != CAPOW_FP64) { m = "precision must be CAPOW_FP32 or CAPOW_FP64."; return CAPOW_ERR_ARG; }
  if (p->edge_mode != CAPOW_EDGE_REFERENCE && p->edge_mode != CAPOW_EDGE_PERIODIC &&
      p->edge_mode != CAPOW_EDGE_SURVEY) { m = "edge_mode must be CAPOW_EDGE_REFERENCE, CAPOW_EDGE_PERIODIC or CAPOW_EDGE_SURVEY."; return CAPOW_ERR_ARG; }
  if (p->edge_mode == CAPOW_EDGE_SURVEY && p->momentum != 0 && p->momentum != 1 && p->momentum != 3) {
    m = "survey `Momentum' must be 0, 1 or 3 (2 has no ddg_mom grid in the reference).";
    return CAPOW_ERR_UNSUPPORTED;
  }
  if (p->nz % 2 != 0) { m = "NZ must be even (r2c row transform)."; return CAPOW_ERR_UNSUPPORTED; }
  if (p->nx > 65536 || p->ny > 65536 || p->nz > 65536) { m = "grid dimension above 65536."; return CAPOW_ERR_UNSUPPORTED; }
  FftPlan pl;
  if (!make_plan(p->nx, &pl) || !make_plan(p->ny, &pl) || !make_plan(p->nz / 2, &pl)) {
    m = "grid dimensions must factor into 2, 3 and 5";
    return CAPOW_ERR_UNSUPPORTED;
  }
  const int64_t dims[3] = {p->nx, p->ny, p->nz};
  const int tmax[3] = {16, 16, p->precision == CAPOW_FP64 ? 16 : 32};
  int64_t ntiles = 1;
  for (int a = 0; a < 3; a++) {
    const int t = pick_tile((int)dims[a], tmax[a], a == 2);
    if (t < 0) {
      m = "grid dimension has no tile divisor in [4, 32]";
      return CAPOW_ERR_UNSUPPORTED;
    }
    ntiles *= dims[a] / t;
  }
  // sort keys tile * 27 + border class stay within 32 bits
  if (ntiles > (int64_t)4096 * 4096) {
    m = "grid has more than 16,777,216 deposit tiles";
    return CAPOW_ERR_UNSUPPORTED;
  }
  return CAPOW_OK;
}

template <typename T>
static std::vector<T> twiddles(int64_t N) {
  std::vector<T> v(2 * N);
  for (int64_t j = 0; j < N; j++) {
    double a = 2.0 * M_PI * (double)j / (double)N;
    v[2 * j] = (T)cos(a);
    v[2 * j + 1] = (T)(-sin(a));
  }
  return v;
}

// One table entry of axis `axis` (0 x, 1 y, 2 z), as src/main.c:329-352
// and 382-389 form them: frequency with the float casts, the window
// deconvolution pow(1/sinc, 2 InterpOrder) and the interlace phase.
static void bin_table_entry(int axis, unsigned long long N, int i, double s, double d, int po,
                            double* f_out, double* cor, double* c_ph, double* s_ph) {
  double f;
  if (axis == 2) f = (float)i * s;
  else if ((unsigned long long)i <= N / 2) f = (float)i * s;
  else f = ((float)i - (float)N) * s;
  *f_out = f;
  const double a = M_PI * f * d;
  double sn = 1.0;
  if (f != 0.0) sn = sin(a) / a;
  *cor = pow(1.0 / sn, po);
  *c_ph = cos(a);
  *s_ph = sin(a);
}

namespace capow {

// Table layout: fx[NX] fy[NY] fz[nzh] corx[NX] cory[NY] corz[nzh]
// phx[2 NX] phy[2 NY] phz[2 nzh]; entry idx computed on its own (the
// interlacing factor 0.25 of src/main.c:395 folded into corz).
double bin_table_at(const capow_params& p, size_t idx) {
  const unsigned long long N[3] = {(unsigned long long)p.nx, (unsigned long long)p.ny,
                                   (unsigned long long)p.nz};
  const double d[3] = {(p.xmax - p.xmin) / (double)N[0], (p.ymax - p.ymin) / (double)N[1],
                       (p.zmax - p.zmin) / (double)N[2]};
  const size_t cnt[3] = {(size_t)N[0], (size_t)N[1], (size_t)(N[2] / 2 + 1)};
  const size_t sect = cnt[0] + cnt[1] + cnt[2];
  // 0: frequencies, 1: deconvolution, 2: phases (two doubles per entry)
  int kind, axis = 0;
  size_t r;
  if (idx < sect) { kind = 0; r = idx; }
  else if (idx < 2 * sect) { kind = 1; r = idx - sect; }
  else { kind = 2; r = idx - 2 * sect; }
  const size_t per = kind == 2 ? 2 : 1;
  while (axis < 2 && r >= cnt[axis] * per) r -= cnt[axis++] * per;
  const int i = (int)(r / per);
  const double s = 1.0 / ((float)N[axis] * d[axis]);
  double f, cor, c, sn;
  bin_table_entry(axis, N[axis], i, s, d[axis], 2 * p.interp_order, &f, &cor, &c, &sn);
  if (kind == 0) return f;
  if (kind == 1) return (axis == 2 && p.do_interlacing) ? cor * 0.25 : cor;
  return (r % 2) ? sn : c;
}

size_t bin_table_len(const capow_params& p) { return 4 * (size_t)(p.nx + p.ny + p.nz / 2 + 1); }

void fill_bin_tables(const capow_params& p, double* tab) {
  const size_t n = bin_table_len(p);
  for (size_t i = 0; i < n; i++) tab[i] = bin_table_at(p, i);
}

int make_bin_plan(const capow_params& p, int64_t rowc, bool mirror_allowed, int ncu, BinPlan& bp,
                  std::string& m) {
  memset(&bp, 0, sizeof(bp));
  const unsigned long long NX = p.nx, NY = p.ny, NZ = p.nz;
  const double dx = (p.xmax - p.xmin) / (double)NX, dy = (p.ymax - p.ymin) / (double)NY,
               dz = (p.zmax - p.zmin) / (double)NZ;
  const double sx = 1.0 / ((float)NX * dx), sy = 1.0 / ((float)NY * dy);
  const int nzh = (int)(NZ / 2 + 1);
  const size_t ofx = 0, ofy = NX, ofz = NX + NY, ocx = ofz + nzh, ocy = ocx + NX, ocz = ocy + NY,
               opx = ocz + nzh, opy = opx + 2 * NX, opz = opy + 2 * NY, total = opz + 2 * nzh;
  bp.ntab = total;
  int fa, fb;
  field_exps(p.momentum, &fa, &fb);
  if (p.edge_mode == CAPOW_EDGE_SURVEY) fb = (p.momentum == 3) ? 1 : -1;
  const bool hasB = fb >= 0;
  // mirror-group binning: f(N - i) = -f(i) and cor(N - i) = cor(i) exactly
  // on x and y, even NX and NY, and mu's sign irrelevant (line of sight
  // along z, or even multipoles only)
  {
    const int po = 2 * p.interp_order;
    bool sym = NX % 2 == 0 && NY % 2 == 0;
    double f1, c1, f2, c2, t0, t1;
    for (unsigned long long i = 1; sym && 2 * i < NX; i++) {
      bin_table_entry(0, NX, (int)i, sx, dx, po, &f1, &c1, &t0, &t1);
      bin_table_entry(0, NX, (int)(NX - i), sx, dx, po, &f2, &c2, &t0, &t1);
      sym = f2 == -f1 && c2 == c1;
    }
    for (unsigned long long j = 1; sym && 2 * j < NY; j++) {
      bin_table_entry(1, NY, (int)j, sy, dy, po, &f1, &c1, &t0, &t1);
      bin_table_entry(1, NY, (int)(NY - j), sy, dy, po, &f2, &c2, &t0, &t1);
      sym = f2 == -f1 && c2 == c1;
    }
    const bool mu_even = !p.odd_multipoles || (p.los != 1 && p.los != 2);
    bp.mirror = sym && mu_even && mirror_allowed;
    bp.sym = sym;
  }
  BinArgsDev& b = bp.b;
  b.nx = (int)NX; b.ny = (int)NY; b.nzh = nzh;
  b.fd_nzh.init((uint32_t)nzh);
  b.rowc = (int)rowc;
  b.fd_rowc.init((uint32_t)rowc);
  b.nk = p.nk; b.nmu = p.output_2d ? p.nmu : 1;
  b.out2d = p.output_2d; b.odd = p.odd_multipoles; b.logbins = p.output_log;
  b.los = p.los; b.interlace = p.do_interlacing; b.hasB = hasB;
  b.mink = p.mink;
  b.binwidth = (p.maxk - p.mink) / (double)p.nk;  // src/main.c:305
  b.inv_binwidth = 1.0 / b.binwidth;
  double nyx = M_PI / dx, nyy = M_PI / dy, nyz = M_PI / dz, mn = nyx;
  if (nyy < mn) mn = nyy;
  if (nyz < mn) mn = nyz;
  b.min_nyq = mn;
  b.ofx = (int64_t)ofx; b.ofy = (int64_t)ofy; b.ofz = (int64_t)ofz;
  b.ocx = (int64_t)ocx; b.ocy = (int64_t)ocy; b.ocz = (int64_t)ocz;
  b.opx = (int64_t)opx; b.opy = (int64_t)opy; b.opz = (int64_t)opz;
  int32_t ns, codes[3];
  capow_spectra(&p, &ns, codes);
  b.nspec = ns;
  for (int s = 0; s < 3; s++) bp.codes[s] = s < ns ? codes[s] : 0;
  for (int s = 0; s < ns; s++) b.kind[s] = spec_kind(codes[s], p.momentum);
  const int nl = p.odd_multipoles ? 5 : 3;
  const int n2d = p.output_2d ? p.nk * p.nmu : 0;
  bp.per_copy = ns * nl * p.nk + ns * n2d + p.nk + n2d;
  const size_t bytes = (size_t)bp.per_copy * 8;
  const size_t lds_max = 160 * 1024;
  if (bytes > lds_max) {
    m = "NK/NMU too large for the LDS bin arrays";
    return CAPOW_ERR_UNSUPPORTED;
  }
  // one wavefront per private bin copy, as many as fit in one CU's LDS
  // beside the z tables (up to 8); a persistent grid of one workgroup per CU.
  // The branch-free row form (linear bins, no 2-D bins) and the mirror-group
  // form (linear or log bins) keep only the sums in LDS: their mode counts
  // are computed once.  (k, mu) bins in the mirror form (line of sight
  // along z, where the four mirror modes share mu): a 1-D pass with per-k
  // bounds, then one int64 (k, mu) table per workgroup.
  const size_t ztab = (size_t)nzh * 4 * 8;
  bp.row_form = bytes + ztab <= lds_max;
  bp.mirror2d = bp.row_form && bp.mirror && b.out2d && p.los != 1 && p.los != 2 &&
                (size_t)ns * n2d * 8 + ztab <= lds_max &&
                (size_t)(ns * nl + 1) * p.nk * 8 + ztab <= lds_max;
  const bool fast = bp.row_form && !b.out2d && (!b.logbins || bp.mirror);
  const size_t lbytes = bp.mirror2d ? (size_t)(ns * nl + 1) * p.nk * 8
                                    : (fast ? (size_t)ns * nl * p.nk * 8 : bytes);
  const int copies = (int)((lds_max - (bp.row_form ? ztab : 0)) / lbytes);
  b.copies = copies > 8 ? 8 : copies;
  const bool fp64 = p.precision == CAPOW_FP64;
  const int64_t modes = (int64_t)NX * NY * rowc;
  const int64_t per_wg = (int64_t)64 * b.copies * (fp64 ? 4 : 8);   // modes per iteration
  const int64_t nbatch = ((int64_t)NX * NY + 63) / 64;              // row form: 64-row batches
  const int64_t want = bp.row_form ? (nbatch + b.copies - 1) / b.copies : (modes + per_wg - 1) / per_wg;
  bp.nwg = (int)(want < ncu ? want : ncu);
  return CAPOW_OK;
}

}  // namespace capow

static int setup(capow_ctx* c) {
  const capow_params& p = c->p;
  Geom& g = c->g;
  memset(&g, 0, sizeof(g));
  g.n[0] = (int)p.nx; g.n[1] = (int)p.ny; g.n[2] = (int)p.nz;
  g.nzh = (int)(p.nz / 2 + 1);
  {
    // rows start on 128-byte lines: 16 complex (FP32) or 8 (FP64)
    const int64_t align = p.precision == CAPOW_FP64 ? 8 : 16;
    g.rowc = ((int64_t)g.nzh + align - 1) / align * align;
    g.rowlen = 2 * g.rowc;
  }
  const double lo[3] = {p.xmin, p.ymin, p.zmin}, hi[3] = {p.xmax, p.ymax, p.zmax};
  for (int a = 0; a < 3; a++) {
    g.lo[a] = lo[a];
    g.hi[a] = hi[a];
    g.L[a] = hi[a] - lo[a];
    g.d[a] = (hi[a] - lo[a]) / (double)(unsigned long long)g.n[a];  // src/main.c:1126-1128
    g.invd[a] = 1.0 / g.d[a];
  }
  g.order = p.interp_order;
  g.interlace = p.do_interlacing;
  g.los = p.los;
  g.edge = p.edge_mode;
  g.periodic = (p.edge_mode == CAPOW_EDGE_PERIODIC) || (p.los == 0);
  field_exps(p.momentum, &g.fa, &g.fb);
  g.mean_sub_a = (g.fa == 0);
  g.rsd_axis = (p.los >= 1 && p.los <= 3) ? p.los - 1 : -1;
  if (p.edge_mode == CAPOW_EDGE_SURVEY) {
    // assign_survey_data (src/read_data.c:461-486): ddg = w (codes 0, 3) or
    // w v (code 1); ddg_mom = w v (code 3); randoms arrive with -alpha w.
    // Non-periodic taps (the reference-rule path with LOS != 0), no RSD
    // shift and no mean subtraction (src/main.c:226-228).
    g.los = 3;
    g.edge = CAPOW_EDGE_REFERENCE;
    g.periodic = false;
    g.rsd_axis = -1;
    g.fa = (p.momentum == 1) ? 1 : 0;
    g.fb = (p.momentum == 3) ? 1 : -1;
    g.mean_sub_a = 0;
  }
  g.rsd_f1 = 1.0 + p.redshift;
  g.rsd_den = 100.0 * sqrt(p.omega_m * pow((1.0 + p.redshift), 3.0) + 1.0 - p.omega_m);
  if (g.order == 3) { g.reach_lo = -1; g.reach_hi = g.interlace ? 2 : 1; }
  else if (g.order == 2) { g.reach_lo = 0; g.reach_hi = g.interlace ? 2 : 1; }
  else { g.reach_lo = 0; g.reach_hi = g.interlace ? 1 : 0; }
  // deposit tile: one phase of accumulators (2 fields) in <= 64 KB of LDS
  const int tmax[3] = {16, 16, p.precision == CAPOW_FP64 ? 16 : 32};
  for (int a = 0; a < 3; a++) {
    int t = pick_tile(g.n[a], tmax[a], a == 2);
    if (t < 0) return fail(c, CAPOW_ERR_UNSUPPORTED, "grid dimension has no tile divisor in [4, 32]");
    g.tdim[a] = t;
    g.tdiv[a].init((uint32_t)t);
    g.ntile[a] = g.n[a] / t;
  }
  // one bucket per (tile, border class)
  g.nbuckets = (int64_t)g.ntile[0] * g.ntile[1] * g.ntile[2] * 27;
  g.ntiles_mock = (int64_t)g.ntile[0] * g.ntile[1] * g.ntile[2];
  g.mock_stride = 0;
  const int M = c->nstack;
  if ((double)g.nbuckets * M > (double)UINT32_MAX - 1)
    return fail(c, CAPOW_ERR_UNSUPPORTED, "too many sort buckets for this stack");
  // LDS hash probes per key in the counting sort; CAPOW_SORT_PROBES=0 sends
  // every copy through the global counters (the table-full path)
  g.sort_probes = 32;
  if (const char* e = getenv("CAPOW_SORT_PROBES")) {
    const int v = atoi(e);
    g.sort_probes = v < 0 ? 0 : (v > 32 ? 32 : v);
  }
  c->fp64 = p.precision == CAPOW_FP64;
  if (c->fft_order < 0) c->fft_order = c->fp64 ? 0 : 5;
  c->esz = c->fp64 ? 8 : 4;

  // grids
  // host layout: the in-place r2c layout [NX][NY][2(NZ/2+1)]; device rows
  // are padded to rowlen reals
  c->grid_elems = (int64_t)p.nx * p.ny * 2 * g.nzh;
  const bool hasB = g.fb >= 0;
  // intermediate layouts of the out-of-place FFT: [x][kz][y] and [kz][ky][x]
  // with 128-byte row pitches; every buffer holds the largest layout
  {
    const int64_t al = c->fp64 ? 8 : 16;
    c->fft_py = (p.ny + al - 1) / al * al;
    c->fft_px = (p.nx + al - 1) / al * al;
  }
  int64_t cap_c = (int64_t)p.nx * p.ny * g.rowc;
  const int64_t l1 = (int64_t)p.nx * g.nzh * c->fft_py, l2 = (int64_t)g.nzh * p.ny * c->fft_px;
  if (l1 > cap_c) cap_c = l1;
  if (l2 > cap_c) cap_c = l2;
  const size_t gbytes = (size_t)cap_c * 2 * c->esz;
  c->grid_bytes = gbytes;
  int ngr = 0;
  // a stack: the catalogues' grids back to back in one buffer per slot
  if (M > 1) g.mock_stride = cap_c * 2;
  for (int s = 0; s < 4; s++) {
    bool need = (s == 0) || (s == 1 && g.interlace) || (s == 2 && hasB) || (s == 3 && hasB && g.interlace);
    if (need) {
      HIPCHK(c, hipMalloc(&c->grid[s], gbytes * M));
      ngr++;
    }
  }
  if (M > 1) {
    // a scratch buffer per grid slot, or no stack
    for (int s = 0; s < 4; s++)
      if (c->grid[s]) HIPCHK(c, hipMalloc(&c->scratch[s], gbytes * M));
  }
  bool want_scratch = M == 1;
  if (const char* e = getenv("CAPOW_FFT_TRANSPOSED")) want_scratch = atoi(e) != 0;
  // CAPOW_FFT_ONE_SPARE=1: the one-spare pipeline even when a full second
  // set of grids would fit (for tests at small sizes)
  bool force_one = false;
  if (const char* e = getenv("CAPOW_FFT_ONE_SPARE")) force_one = atoi(e) != 0;
  if (force_one && want_scratch && ngr > 1) {
    HIPCHK(c, hipMalloc(&c->scratch[0], gbytes));
    c->fft_one_spare = true;
    want_scratch = false;
  }
  if (want_scratch) {
    // only with room to spare: the second set of grids plus 16 GB
    size_t fre = 0, tot = 0;
    if (hipMemGetInfo(&fre, &tot) == hipSuccess && fre > (size_t)ngr * gbytes + ((size_t)16 << 30)) {
      for (int s = 0; s < 4; s++) {
        if (!c->grid[s]) continue;
        if (hipMalloc(&c->scratch[s], gbytes) != hipSuccess) {
          hipGetLastError();
          for (int q = 0; q < 4; q++)
            if (c->scratch[q]) { hipFree(c->scratch[q]); c->scratch[q] = nullptr; }
          break;
        }
      }
    } else if (ngr > 1 && hipMemGetInfo(&fre, &tot) == hipSuccess &&
               fre > gbytes + ((size_t)16 << 30)) {
      // room for one spare: the transposed passes run grid by grid
      if (hipMalloc(&c->scratch[0], gbytes) == hipSuccess) c->fft_one_spare = true;
      else hipGetLastError();
    }
  }
  // sort buffers
  const int64_t nbk = g.nbuckets * M;
  HIPCHK(c, hipMalloc((void**)&c->counts, (size_t)nbk * 4));
  HIPCHK(c, hipMalloc((void**)&c->offsets, (size_t)(nbk + 1) * 4));
  HIPCHK(c, hipMalloc((void**)&c->cursor, (size_t)nbk * 4));
  int64_t nblk = (nbk + 4095) / 4096;
  HIPCHK(c, hipMalloc((void**)&c->bsum, (size_t)(nblk + 1) * 4));
  HIPCHK(c, hipMalloc((void**)&c->boff, (size_t)(nblk + 1) * 4));
  HIPCHK(c, hipMalloc((void**)&c->partials, (size_t)c->count_blocks * STAT_N * 8 * M));
  HIPCHK(c, hipMalloc((void**)&c->errs, 4 * sizeof(unsigned int) * M));
  HIPCHK(c, hipMalloc((void**)&c->stats, STAT_TOTAL * 8 * M));
  HIPCHK(c, hipHostMalloc((void**)&c->h_stats, STAT_TOTAL * 8 * M, hipHostMallocDefault));
  HIPCHK(c, hipHostMalloc((void**)&c->h_errs, 4 * sizeof(unsigned int) * M, hipHostMallocDefault));
  if (M > 1) {
    HIPCHK(c, hipMalloc((void**)&c->stack_tab, sizeof(MockDev) * M));
    HIPCHK(c, hipHostMalloc((void**)&c->h_stack_tab, sizeof(MockDev) * M, hipHostMallocDefault));
  }

  // FFT plans + twiddles
  if (!make_plan(p.nx, &c->plan[0]) || !make_plan(p.ny, &c->plan[1]) ||
      !make_plan(p.nz / 2, &c->plan[2]))
    return fail(c, CAPOW_ERR_UNSUPPORTED, "grid dimensions must factor into 2, 3 and 5");
  if (p.nx * 8 > 8192 * 2 || p.ny * 8 > 8192 * 2 || (p.nz / 2) > 8192)
    return fail(c, CAPOW_ERR_UNSUPPORTED, "FFT length above the per-workgroup limit");
  const int64_t tn[3] = {p.nx, p.ny, p.nz};
  for (int a = 0; a < 3; a++) {
    if (c->fp64) {
      auto v = twiddles<double>(tn[a]);
      HIPCHK(c, hipMalloc(&c->tw[a], v.size() * 8));
      HIPCHK(c, hipMemcpy(c->tw[a], v.data(), v.size() * 8, hipMemcpyHostToDevice));
    } else {
      auto v = twiddles<float>(tn[a]);
      HIPCHK(c, hipMalloc(&c->tw[a], v.size() * 4));
      HIPCHK(c, hipMemcpy(c->tw[a], v.data(), v.size() * 4, hipMemcpyHostToDevice));
    }
  }

  // binning plan and tables (src/main.c:305-352, 382-389)
  {
    int dev = 0, ncu = 256;
    if (hipGetDevice(&dev) == hipSuccess)
      hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    if (ncu < 1) ncu = 256;
    c->ncu = ncu;
    BinPlan bp;
    std::string m;
    const int rc = make_bin_plan(p, g.rowc, c->bin_mirror_env != 0, ncu, bp, m);
    if (rc) return fail(c, rc, m);
    std::vector<double> tab(bp.ntab);
    fill_bin_tables(p, tab.data());
    HIPCHK(c, hipMalloc((void**)&c->bin_tabs, bp.ntab * 8));
    HIPCHK(c, hipMemcpy(c->bin_tabs, tab.data(), bp.ntab * 8, hipMemcpyHostToDevice));
    c->bin = bp.b;
    c->bin_mirror = bp.mirror;
    c->bin_sym = bp.sym && c->bin_mirror_env != 0;
    c->bin_row_form = bp.row_form;
    c->bin_mirror2d = bp.mirror2d;
    c->nspec = bp.b.nspec;
    for (int s = 0; s < 3; s++) c->codes[s] = bp.codes[s];
    c->bin_per_copy = bp.per_copy;
    c->bin_nwg = bp.nwg;
    const size_t bytes = (size_t)bp.per_copy * 8;
    const int n2d = p.output_2d ? p.nk * p.nmu : 0;
    HIPCHK(c, hipMalloc((void**)&c->bin_partials, (size_t)c->bin_nwg * bytes));
    HIPCHK(c, hipMalloc((void**)&c->bin_out, bytes * M));
    HIPCHK(c, hipMalloc((void**)&c->bin_counts, (size_t)(p.nk + n2d) * 8));
    if (c->bin_mirror2d) {
      HIPCHK(c, hipMalloc((void**)&c->bin_partials2, (size_t)c->bin_nwg * bp.b.nspec * n2d * 8));
      HIPCHK(c, hipMalloc((void**)&c->bin_qscale, (size_t)2 * p.nk * 8));
    }
    c->bin_counts_valid = false;
    HIPCHK(c, hipHostMalloc((void**)&c->h_bins, bytes * M, hipHostMallocDefault));
    // the fused x pass + binning: FP32 out-of-place transform (scratch
    // grids), persistent row passes, mirror-group 1-D bins
    const size_t xl = xbin_lds(g, bp.b);
    c->xbin_ok = !c->fp64 && c->scratch[0] && c->fft_persist && bp.mirror && bp.row_form &&
                 !p.output_2d && xl > 0 && c->xbin_env != 0 && p.edge_mode != CAPOW_EDGE_SURVEY;
    if (c->xbin_ok) {
      // cor_max: the largest deconvolution factor of each axis (x 0.25 with
      // interlacing, folded into the z table)
      double cm[3] = {0.0, 0.0, 0.0};
      const size_t nax[3] = {(size_t)p.nx, (size_t)p.ny, (size_t)(p.nz / 2 + 1)};
      const size_t sect = nax[0] + nax[1] + nax[2];
      size_t off = sect;   // the deconvolution tables follow the frequencies
      for (int ax = 0; ax < 3; ax++) {
        for (size_t i = 0; i < nax[ax]; i++) {
          const double v = bin_table_at(p, off + i);
          if (v > cm[ax]) cm[ax] = v;
        }
        off += nax[ax];
      }
      c->xbin_factor = 4.0 * cm[0] * cm[1] * cm[2] * (double)p.nx * (double)p.ny;
      const int per_cu = (int)((160 * 1024) / xl) > 0 ? (int)((160 * 1024) / xl) : 1;
      c->xbin_rows = c->ncu * (per_cu < 8 ? per_cu : 8);
      const int F = bp.b.nspec * (p.odd_multipoles ? 5 : 3);
      HIPCHK(c, hipMalloc((void**)&c->xbin_ssq, (size_t)XBIN_SSQ_MAX * 4 * 8 * M));
      HIPCHK(c, hipMalloc((void**)&c->xbin_q, 6 * 8 * M));
      HIPCHK(c, hipMalloc((void**)&c->xbin_part, (size_t)c->xbin_rows * F * p.nk * 8 * M));
    }
    if (M > 1 && !c->xbin_ok)
      return fail(c, CAPOW_ERR_UNSUPPORTED,
                  "stacked batches need the fused x-pass binning (FP32, 1-D bins, register FFT lengths)");
  }
  return CAPOW_OK;
}

static void release(capow_ctx* c) {
  for (int s = 0; s < 4; s++) if (c->grid[s]) hipFree(c->grid[s]);
  for (int s = 0; s < 4; s++) if (c->scratch[s]) hipFree(c->scratch[s]);
  void* ps[] = {c->own_parts, c->own_w, c->sorted, c->sorted_w, c->counts, c->offsets,
                c->cursor, c->bsum, c->boff, c->partials, c->errs, c->stats,
                c->tw[0], c->tw[1], c->tw[2], c->bin_tabs, c->bin_partials, c->bin_out,
                c->bin_counts, c->bin_partials2, c->bin_qscale, c->xbin_ssq, c->xbin_q,
                c->xbin_part};
  for (void* q : ps) if (q) hipFree(q);
  void* bs[] = {c->bwd_G, c->bwd_g0, c->bwd_out, c->bwd_w, c->sorted_idx, c->bwd_err};
  for (void* q : bs) if (q) hipFree(q);
  void* hs[] = {c->h_stats, c->h_errs, c->h_bins, c->h_sv, c->h_stack_tab};
  if (c->stack_tab) hipFree(c->stack_tab);
  for (void* q : hs) if (q) hipHostFree(q);
  void* sv[] = {c->sv_src[0], c->sv_src[1], c->sv_src[2], c->sv_src[3], c->sv_F0, c->sv_F0m,
                c->sv_partials, c->sv_acc};
  for (void* q : sv) if (q) hipFree(q);
}

// nstack > 1: a stacked batch context (internal to the batches)
static int create_ctx(const capow_params* p, int nstack, capow_ctx** out) {
  if (!p || !out) return fail(nullptr, CAPOW_ERR_ARG, "null argument");
  *out = nullptr;
  std::string m;
  int rc = validate(p, m);
  if (rc) return fail(nullptr, rc, m);
  capow_ctx* c = new capow_ctx();
  c->p = *p;
  c->nstack = nstack < 1 ? 1 : nstack;
  if (const char* e = getenv("CAPOW_FFT_COL_MUL")) c->fft_col_mul = atoi(e) == 1 ? 1 : 2;
  if (const char* e = getenv("CAPOW_FFT_ROWS_MUL")) c->fft_rows_mul = atoi(e) == 2 ? 2 : (atoi(e) == 1 ? 1 : 0);
  if (const char* e = getenv("CAPOW_FFT_PERSIST")) c->fft_persist = atoi(e) != 0;
  if (const char* e = getenv("CAPOW_DEPOSIT_HALO")) c->dep_halo = atoi(e) != 0;
  if (const char* e = getenv("CAPOW_DEPOSIT_BANK")) c->dep_bank = atoi(e) != 0;
  if (const char* e = getenv("CAPOW_GATHER_TILES")) c->bwd_tiles = atoi(e) != 0;
  if (const char* e = getenv("CAPOW_FFT_ORDER")) c->fft_order = atoi(e) & 7;
  if (const char* e = getenv("CAPOW_BIN_MIRROR")) c->bin_mirror_env = atoi(e) != 0;
  if (const char* e = getenv("CAPOW_XBIN")) c->xbin_env = atoi(e) != 0;
  rc = setup(c);
  if (rc) {
    g_create_err = c->err;
    release(c);
    delete c;
    return rc;
  }
  *out = c;
  return CAPOW_OK;
}

extern "C" int capow_create(const capow_params* p, capow_ctx** out) {
  return create_ctx(p, 1, out);
}

extern "C" int capow_destroy(capow_ctx* c) {
  if (!c) return CAPOW_OK;
  hipStreamSynchronize(c->stream);
  release(c);
  delete c;
  return CAPOW_OK;
}

extern "C" const char* capow_last_error(const capow_ctx* c) {
  return c ? c->err.c_str() : g_create_err.c_str();
}

extern "C" int capow_set_stream(capow_ctx* c, void* s) {
  if (!c) return CAPOW_ERR_ARG;
  c->stream = (hipStream_t)s;
  c->timer.stream = c->stream;
  return CAPOW_OK;
}

extern "C" int capow_synchronize(capow_ctx* c) {
  if (!c) return CAPOW_ERR_ARG;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->timer.collect();
  return CAPOW_OK;
}

extern "C" int capow_set_timing(capow_ctx* c, int on) {
  if (!c) return CAPOW_ERR_ARG;
  c->timer.on = on != 0;
  c->timer.stream = c->stream;
  return CAPOW_OK;
}

static int ensure(capow_ctx* c, void** buf, size_t* cap, size_t bytes) {
  if (*cap >= bytes) return CAPOW_OK;
  if (*buf) hipFree(*buf);
  *buf = nullptr;
  *cap = 0;
  HIPCHK(c, hipMalloc(buf, bytes));
  *cap = bytes;
  return CAPOW_OK;
}

extern "C" int capow_set_particles(capow_ctx* c, const void* xyzv, const void* w, int64_t n,
                                   int32_t dtype, int32_t where) {
  if (!c) return CAPOW_ERR_ARG;
  if (n < 0 || (n > 0 && !xyzv)) return fail(c, CAPOW_ERR_ARG, "bad particle array");
  if ((dtype == CAPOW_DTYPE_F64) != c->fp64)
    return fail(c, CAPOW_ERR_ARG, "particle dtype must match the context precision");
  // one sorted row per particle, addressed by 32-bit bucket offsets
  if (n > (int64_t)UINT32_MAX)
    return fail(c, CAPOW_ERR_UNSUPPORTED, "more than 2^32 - 1 particles (32-bit row offsets)");
  const size_t pb = (size_t)n * 4 * c->esz, wb = (size_t)n * c->esz;
  if (where == CAPOW_HOST) {
    int rc = ensure(c, &c->own_parts, &c->own_parts_cap, pb > 0 ? pb : 16);
    if (rc) return rc;
    if (n) HIPCHK(c, hipMemcpyAsync(c->own_parts, xyzv, pb, hipMemcpyHostToDevice, c->stream));
    c->parts = c->own_parts;
    if (w) {
      rc = ensure(c, &c->own_w, &c->own_w_cap, wb > 0 ? wb : 16);
      if (rc) return rc;
      if (n) HIPCHK(c, hipMemcpyAsync(c->own_w, w, wb, hipMemcpyHostToDevice, c->stream));
      c->wts = c->own_w;
    } else {
      c->wts = nullptr;
    }
  } else if (where == CAPOW_DEVICE) {
    c->parts = xyzv;
    c->wts = w;
  } else {
    return fail(c, CAPOW_ERR_ARG, "where must be CAPOW_HOST or CAPOW_DEVICE");
  }
  c->n = n;
  c->state = 1;
  c->fwd_assigned = c->fwd_kept = false;
  return CAPOW_OK;
}

static void fill_stats(capow_ctx* c, const double* s) {
  capow_assign_stats& o = c->last;
  o.nread = (double)c->n;
  o.nout = s[STAT_NOUT];
  o.ngrid = s[STAT_NGRID];
  o.nsq = s[STAT_W];
  o.vr = s[STAT_V1];
  o.vr2 = s[STAT_V2];
  o.vr3 = s[STAT_V3];
  o.vr4 = s[STAT_V4];
  o.alpha = s[STAT_ALPHA];
}

// Enqueues the assignment on the context stream: sort, deposit, and
// asynchronous copies of the sums and error flags into h_stats / h_errs.
// exact: size the sorted buffer from the copy count (one synchronisation);
// otherwise the buffer keeps its capacity and an overflow is flagged in
// h_errs[2] (the caller re-runs that catalogue).
static int assign_enqueue(capow_ctx* c, double* h_stats, unsigned int* h_errs, bool exact) {
  if (c->state < 1) return fail(c, CAPOW_ERR_STATE, "capow_assign before capow_set_particles");
  c->x_pending = false;   // the deposit rewrites every grid
  c->fwd_kept = false;
  Geom& g = c->g;
  g.has_w = c->wts != nullptr;
  // cleared by kernels (no memset nodes when the stream is captured)
  if (zero_u32(c->counts, g.nbuckets, c->stream) || zero_u32(c->errs, 4, c->stream))
    return fail(c, CAPOW_ERR_HIP, std::string("clearing the sort counters failed: ") +
                                      hipGetErrorString(hipGetLastError()));
  DepositArgs a;
  a.g = &g;
  a.stream = c->stream;
  a.timer = &c->timer;
  a.parts = c->parts;
  a.wts = c->wts;
  a.n = c->n;
  int64_t want = (c->n + BK_CHUNK - 1) / BK_CHUNK;
  a.count_blocks = (int)(want < 1 ? 1 : (want < c->count_blocks ? want : c->count_blocks));
  a.counts = c->counts;
  a.offsets = c->offsets;
  a.cursor = c->cursor;
  a.bsum = c->bsum;
  a.boff = c->boff;
  a.partials = c->partials;
  a.errs = c->errs;
  a.stats = c->stats;
  for (int s = 0; s < 4; s++) a.grid[s] = c->grid[s];
  a.halo = c->dep_halo;
  a.bank = c->dep_bank;
  int rc = run_count(a, c->fp64);
  if (rc) return fail(c, rc, std::string("deposit count failed: ") + hipGetErrorString(hipGetLastError()));
  if (exact) {
    // one sorted copy per accepted particle: n rows always suffice
    const size_t rows = c->n > 0 ? (size_t)c->n : 1;
    rc = ensure(c, &c->sorted, &c->sorted_cap, rows * 4 * c->esz);
    if (rc) return rc;
    if (c->wts) {
      rc = ensure(c, &c->sorted_w, &c->sorted_w_cap, rows * c->esz);
      if (rc) return rc;
    }
  } else if (!c->sorted || (c->wts && !c->sorted_w)) {
    return fail(c, CAPOW_ERR_STATE, "sorted buffer not reserved");
  }
  a.sorted = c->sorted;
  a.sorted_w = c->sorted_w;
  int64_t cap = (int64_t)(c->sorted_cap / (4 * c->esz));
  if (c->wts) {
    const int64_t capw = (int64_t)(c->sorted_w_cap / c->esz);
    cap = capw < cap ? capw : cap;
  }
  const bool idx = c->want_idx || (c->keep_fwd && !c->cap_fixed);
  if (idx) {
    rc = ensure(c, (void**)&c->sorted_idx, &c->sorted_idx_cap, (size_t)(cap > 0 ? cap : 1) * 4);
    if (rc) return rc;
    a.sorted_idx = c->sorted_idx;
  }
  c->fwd_idx = idx;
  a.sorted_cap = cap;
  rc = run_scatter_deposit(a, c->fp64);
  if (rc) return fail(c, rc, std::string("deposit launch failed: ") + hipGetErrorString(hipGetLastError()));
  HIPCHK(c, hipMemcpyAsync(h_stats, c->stats, STAT_TOTAL * 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(h_errs, c->errs, 4 * sizeof(unsigned int), hipMemcpyDeviceToHost, c->stream));
  c->state = 2;
  c->fwd_assigned = true;
  return CAPOW_OK;
}

// After the stream has completed: sums into the stats, error flags into a
// status.  CAPOW_ERR_NOMEM: the sorted buffer overflowed (fixed capacity).
static int assign_finish(capow_ctx* c, const double* h_stats, const unsigned int* h_errs,
                         capow_assign_stats* st) {
  fill_stats(c, h_stats);
  if (st) *st = c->last;
  if (h_errs[0]) {
    c->state = 1;
    return fail(c, CAPOW_ERR_OUT_OF_BOX,
                "particle out of grid bounds with LOS = 0 (" + std::to_string(h_errs[0]) + " particles)");
  }
  if (h_errs[2]) {
    c->state = 1;
    return fail(c, CAPOW_ERR_NOMEM, "sorted particle buffer too small for this catalogue");
  }
  if (h_errs[1]) {
    c->state = 1;
    return fail(c, CAPOW_ERR_INTERNAL, "deposit reach self-check failed");
  }
  return CAPOW_OK;
}

extern "C" int capow_assign(capow_ctx* c, capow_assign_stats* st) {
  if (!c) return CAPOW_ERR_ARG;
  int rc = assign_enqueue(c, c->h_stats, c->h_errs, true);
  if (rc) return rc;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->timer.collect();
  return assign_finish(c, c->h_stats, c->h_errs, st);
}

static void fft_args(capow_ctx* c, FftArgs& a) {
  a.g = &c->g;
  a.stream = c->stream;
  a.timer = &c->timer;
  a.ngrids = 0;
  for (int s = 0; s < 4; s++)
    if (c->grid[s]) a.grid[a.ngrids++] = c->grid[s];
  for (int q = 0; q < 3; q++) {
    a.plan[q] = c->plan[q];
    a.tw[q] = c->tw[q];
  }
  a.col_mul = c->fft_col_mul;
  int ns = 0;
  for (int s = 0; s < 4; s++)
    if (c->grid[s]) a.scratch[ns++] = c->scratch[s];
  if (!c->scratch[0]) a.scratch[0] = nullptr;
  a.one_spare = c->fft_one_spare;
  if (c->fft_one_spare) {
    a.scratch[0] = c->scratch[0];
    for (int q = 1; q < 4; q++) a.scratch[q] = nullptr;
  }
  a.px = c->fft_px;
  a.py = c->fft_py;
  a.rows_mul = c->fft_rows_mul;
  a.ncu = c->ncu;
  a.persist = c->fft_persist;
  a.order = c->fft_order;
}

// Takes back the buffers a transform rotated through the one spare: the
// i-th transformed grid belongs to slot slot_of[i].
static void keep_rotated(capow_ctx* c, const FftArgs& a, const int* slot_of) {
  for (int i = 0; i < a.ngrids; i++) c->grid[slot_of[i]] = a.grid[i];
  c->scratch[0] = a.scratch[0];
}

static int grid_slots(const capow_ctx* c, int* slot_of) {
  int n = 0;
  for (int s = 0; s < 4; s++)
    if (c->grid[s]) slot_of[n++] = s;
  return n;
}

extern "C" int capow_fft(capow_ctx* c) {
  if (!c) return CAPOW_ERR_ARG;
  if (c->p.edge_mode == CAPOW_EDGE_SURVEY)
    return fail(c, CAPOW_ERR_UNSUPPORTED, "survey context: use capow_survey_power");
  if (c->state < 2) return fail(c, CAPOW_ERR_STATE, "capow_fft before capow_assign");
  if (c->state >= 3) return fail(c, CAPOW_ERR_STATE, "grids already transformed");
  FftArgs a;
  fft_args(c, a);
  if (c->xbin_ok) {
    a.defer_x = true;
    a.ssq = c->xbin_ssq;
    a.xbin_factor = c->xbin_factor;
    a.xq = c->xbin_q;
  }
  int rc = run_fft(a, c->fp64);
  if (rc) return fail(c, rc, std::string("fft launch failed: ") + hipGetErrorString(hipGetLastError()));
  if (a.rotated) {
    int slot_of[4];
    grid_slots(c, slot_of);
    keep_rotated(c, a, slot_of);
  } else if (a.swapped) {
    for (int s = 0; s < 4; s++) {
      void* t = c->grid[s];
      c->grid[s] = c->scratch[s];
      c->scratch[s] = t;
    }
  }
  c->x_pending = a.x_deferred;
  c->xbin_ssq_rows = a.ssq_rows;
  c->state = 3;
  c->fwd_kept = c->fwd_assigned;
  return CAPOW_OK;
}

// Runs the x pass capow_fft left pending, so that the grids hold the
// k-space fields (for anything but the fused binning).
static int finish_x(capow_ctx* c) {
  if (!c->x_pending) return CAPOW_OK;
  FftArgs a;
  fft_args(c, a);
  const int rc = run_fft_x(a);
  if (rc) return fail(c, rc, std::string("fft x pass failed: ") + hipGetErrorString(hipGetLastError()));
  if (a.rotated) {
    int slot_of[4];
    grid_slots(c, slot_of);
    keep_rotated(c, a, slot_of);
  } else {
    for (int s = 0; s < 4; s++) {
      void* t = c->grid[s];
      c->grid[s] = c->scratch[s];
      c->scratch[s] = t;
    }
  }
  c->x_pending = false;
  return CAPOW_OK;
}

static void bin_launch(capow_ctx* c, BinLaunch& L) {
  L.stream = c->stream;
  L.timer = &c->timer;
  for (int s = 0; s < 4; s++) L.grid[s] = c->grid[s] ? c->grid[s] : c->grid[0];
  L.dev = c->bin;
  L.tabs = c->bin_tabs;
  L.row_form = c->bin_row_form;
  L.mirror = c->bin_mirror;
  L.mirror2d = c->bin_mirror2d;
  L.partials2 = c->bin_partials2;
  L.qscale = c->bin_qscale;
  L.nwg = c->bin_nwg;
  L.per_copy = c->bin_per_copy;
  L.partials = c->bin_partials;
  L.out = c->bin_out;
  L.counts = c->bin_counts;
  L.counts_valid = &c->bin_counts_valid;
}

// Enqueues the binning and an asynchronous copy of the bins into h_bins.
static int bin_enqueue(capow_ctx* c, double* h_bins) {
  if (c->state < 3) return fail(c, CAPOW_ERR_STATE, "capow_bin before capow_fft");
  BinLaunch L;
  int rc = CAPOW_ERR_UNSUPPORTED;
  if (c->x_pending) {
    FftArgs a;
    fft_args(c, a);
    a.ssq = c->xbin_ssq;
    a.ssq_rows = c->xbin_ssq_rows;
    a.xbin_factor = c->xbin_factor;
    a.xq = c->xbin_q;
    bin_launch(c, L);
    rc = run_fft_x_bin(a, L, c->xbin_part, c->xbin_rows);
    if (rc && rc != CAPOW_ERR_UNSUPPORTED)
      return fail(c, rc, std::string("fused x pass and binning failed: ") + hipGetErrorString(hipGetLastError()));
  }
  if (rc == CAPOW_ERR_UNSUPPORTED) {
    rc = finish_x(c);
    if (rc) return rc;
    bin_launch(c, L);
    rc = run_bin(L, c->fp64);
  }
  if (rc) return fail(c, rc, std::string("binning launch failed: ") + hipGetErrorString(hipGetLastError()));
  HIPCHK(c, hipMemcpyAsync(h_bins, c->bin_out, (size_t)c->bin_per_copy * 8, hipMemcpyDeviceToHost,
                           c->stream));
  return CAPOW_OK;
}

// Unpacks staged bins (LDS order per spectrum: l = 0, 2, 4 [, 1, 3]).
static void bin_finish(capow_ctx* c, const double* h, capow_bins* out) {
  const capow_params& p = c->p;
  const int nk = p.nk, nl = p.odd_multipoles ? 5 : 3, ns = c->nspec;
  const int n2d = p.output_2d ? p.nk * p.nmu : 0;
  out->n_spectra = ns;
  for (int s = 0; s < 3; s++) out->codes[s] = s < ns ? c->codes[s] : -1;
  static const int lmap[5] = {0, 2, 4, 1, 3};
  if (out->pk) {
    for (int s = 0; s < ns; s++) {
      for (int l = 0; l < 5; l++)
        for (int k = 0; k < nk; k++) out->pk[(s * 5 + l) * nk + k] = 0.0;
      for (int q = 0; q < nl; q++)
        for (int k = 0; k < nk; k++)
          out->pk[(s * 5 + lmap[q]) * nk + k] = h[(size_t)(s * nl + q) * nk + k];
    }
  }
  const double* acc2 = h + (size_t)ns * nl * nk;
  const double* cnt = acc2 + (size_t)ns * n2d;
  const double* cnt2 = cnt + nk;
  if (out->nmodes)
    for (int k = 0; k < nk; k++) out->nmodes[k] = (int64_t)cnt[k];
  if (p.output_2d) {
    if (out->pk2d)
      for (int q = 0; q < ns * n2d; q++) out->pk2d[q] = acc2[q];
    if (out->nmodes2d)
      for (int q = 0; q < n2d; q++) out->nmodes2d[q] = (int64_t)cnt2[q];
  }
}

extern "C" int capow_bin(capow_ctx* c, capow_bins* out) {
  if (!c || !out) return CAPOW_ERR_ARG;
  if (c->p.edge_mode == CAPOW_EDGE_SURVEY)
    return fail(c, CAPOW_ERR_UNSUPPORTED, "survey context: use capow_survey_power");
  int rc = bin_enqueue(c, c->h_bins);
  if (rc) return rc;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->timer.collect();
  bin_finish(c, c->h_bins, out);
  return CAPOW_OK;
}

extern "C" int capow_power(capow_ctx* c, capow_assign_stats* st, capow_bins* out) {
  if (!c || !out) return CAPOW_ERR_ARG;
  int rc = capow_assign(c, st);
  if (rc) return rc;
  rc = capow_fft(c);
  if (rc) return rc;
  return capow_bin(c, out);
}

// The (ii, jj, kk) index sets of the survey multipole loops, in the
// reference's order: src/main.c:540 (l = 1), 594-595 (l = 2, ii <= jj),
// 649-657 (l = 3), 709-726 (l = 4).
static int survey_combos(int l, int idx[][3]) {
  int n = 0;
  auto put = [&](int a, int b, int d) { idx[n][0] = a; idx[n][1] = b; idx[n][2] = d; n++; };
  if (l == 0) {
    put(0, 0, 0);
  } else if (l == 1) {
    for (int ii = 0; ii < 3; ii++) put(ii, 0, 0);
  } else if (l == 2) {
    for (int ii = 0; ii < 3; ii++)
      for (int jj = ii; jj < 3; jj++) put(ii, jj, 0);
  } else if (l == 3) {
    for (int ii = 0; ii < 3; ii++)
      for (int jj = ii; jj < (ii == 0 ? 2 : ii + 1); jj++)
        for (int kk = (jj != ii ? 2 : 0); kk < 3; kk++) put(ii, jj, kk);
  } else {
    for (int ii = 0; ii < 3; ii++)
      for (int jj = 0; jj < 3; jj++) {
        int lo = 0, hi = 3;
        if (jj != ii) {
          if (ii == 0) lo = jj;
          else if (ii == 1) lo = 2;
          else if (jj == 1) continue;
          else { lo = 1; hi = 2; }
        }
        for (int kk = lo; kk < hi; kk++) put(ii, jj, kk);
      }
  }
  return n;
}

extern "C" int capow_survey_power(capow_ctx* c, const double origin[3], double* pk,
                                  int64_t* nmodes) {
  if (!c || !origin || !pk || !nmodes) return CAPOW_ERR_ARG;
  const capow_params& p = c->p;
  if (c->state < 2) return fail(c, CAPOW_ERR_STATE, "capow_survey_power before the grids are filled");
  if (c->state >= 3) return fail(c, CAPOW_ERR_STATE, "grids already transformed");
  if (p.momentum == 2 || p.momentum == 5)
    return fail(c, CAPOW_ERR_UNSUPPORTED, "survey estimator: Momentum 2 and 5 have no ddg_mom grid");
  const int NK = p.nk;
  const int per_copy = 6 * NK;   // accumulated P0..P4 and mode counts
  // LDS per workgroup: 4 waves x up to 4 sums per k bin
  if ((size_t)4 * 4 * NK * 8 > (size_t)160 * 1024)
    return fail(c, CAPOW_ERR_UNSUPPORTED, "survey estimator: NK too large for the LDS bins");
  const bool mom = p.momentum != 0 && p.momentum != 1;
  const bool used[4] = {true, p.do_interlacing != 0, mom, mom && p.do_interlacing != 0};
  const size_t rbytes = (size_t)p.nx * p.ny * c->g.rowlen * c->esz;
  // one spare buffer per used grid, of the grids' capacity: it takes the
  // grid's place for the weighted copies while the real-space field stays
  // untouched in the original buffer, which becomes the grid again at the end
  for (int s = 0; s < 4; s++)
    if (used[s] && !c->sv_src[s]) HIPCHK(c, hipMalloc(&c->sv_src[s], c->grid_bytes));
  if (!c->sv_F0) HIPCHK(c, hipMalloc(&c->sv_F0, rbytes));
  if (mom && !c->sv_F0m) HIPCHK(c, hipMalloc(&c->sv_F0m, rbytes));
  if (!c->sv_acc) {
    const int64_t modes = (int64_t)p.nx * p.ny * c->g.rowc;
    // up to 5 workgroups of 4 waves per CU (the registers' limit)
    int64_t want = (modes + 255) / 256;
    if (want > 5 * (int64_t)c->ncu) want = 5 * (int64_t)c->ncu;
    c->sv_nwg = (int)(want < 1 ? 1 : want);
    HIPCHK(c, hipMalloc((void**)&c->sv_partials, (size_t)c->sv_nwg * 4 * NK * 8));
    HIPCHK(c, hipMalloc((void**)&c->sv_acc, (size_t)per_copy * 8));
    HIPCHK(c, hipHostMalloc((void**)&c->h_sv, (size_t)per_copy * 8, hipHostMallocDefault));
  }
  HIPCHK(c, hipMemsetAsync(c->sv_acc, 0, (size_t)per_copy * 8, c->stream));
  // swap the spare buffers in: sv_src holds the fields, grid the work copies
  auto swap_in_out = [&]() {
    for (int s = 0; s < 4; s++)
      if (used[s]) std::swap(c->grid[s], c->sv_src[s]);
  };
  swap_in_out();
  SurveyLaunch L;
  L.stream = c->stream;
  L.timer = &c->timer;
  L.g = &c->g;
  for (int s = 0; s < 4; s++) {
    L.src[s] = used[s] ? c->sv_src[s] : nullptr;
    L.work[s] = used[s] ? c->grid[s] : nullptr;
  }
  L.F0 = c->sv_F0;
  L.F0m = mom ? c->sv_F0m : nullptr;
  L.dev = c->bin;
  L.dev.out2d = 0;
  L.dev.nmu = 1;
  L.tabs = c->bin_tabs;
  for (int a = 0; a < 3; a++) L.origin[a] = origin[a];
  L.mom = mom ? 1 : 0;
  L.mirror = c->bin_sym;
  L.nwg = c->sv_nwg;
  L.partials = c->sv_partials;
  L.acc = c->sv_acc;
  int idx[15][3];
  for (int l = 0; l <= 4; l += p.odd_multipoles ? 1 : 2) {
    const int nc = survey_combos(l, idx);
    for (int q = 0; q < nc; q++) {
      int rc = run_survey_weight(L, c->fp64, l, idx[q]);
      if (rc) {
        swap_in_out();
        return fail(c, rc, std::string("survey weighting failed: ") + hipGetErrorString(hipGetLastError()));
      }
      // transform the weighted copies (the slots in use only)
      FftArgs a;
      fft_args(c, a);
      a.ngrids = 0;
      int slot_of[4];
      for (int s = 0; s < 4; s++)
        if (used[s]) {
          slot_of[a.ngrids] = s;
          a.grid[a.ngrids] = c->grid[s];
          a.scratch[a.ngrids] = c->fft_one_spare ? nullptr : c->scratch[s];
          a.ngrids++;
        }
      if (c->fft_one_spare) a.scratch[0] = c->scratch[0];
      if (!c->scratch[0]) a.scratch[0] = nullptr;
      rc = run_fft(a, c->fp64);
      if (rc) {
        swap_in_out();
        return fail(c, rc, std::string("fft launch failed: ") + hipGetErrorString(hipGetLastError()));
      }
      if (a.rotated) {
        keep_rotated(c, a, slot_of);
        for (int s = 0; s < 4; s++)
          if (used[s]) L.work[s] = c->grid[s];
      } else if (a.swapped) {
        for (int s = 0; s < 4; s++)
          if (used[s]) {
            void* t = c->grid[s];
            c->grid[s] = c->scratch[s];
            c->scratch[s] = t;
            L.work[s] = c->grid[s];
          }
      }
      rc = run_survey_bin(L, c->fp64, l, idx[q]);
      if (rc) {
        swap_in_out();
        return fail(c, rc, std::string("survey binning failed: ") + hipGetErrorString(hipGetLastError()));
      }
    }
  }
  // the grids are the untouched real-space fields again
  swap_in_out();
  HIPCHK(c, hipMemcpyAsync(c->h_sv, c->sv_acc, (size_t)per_copy * 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->timer.collect();
  for (int l = 0; l < 5; l++)
    for (int k = 0; k < NK; k++)
      pk[l * NK + k] = (l % 2 && !p.odd_multipoles) ? 0.0 : c->h_sv[l * NK + k];
  for (int k = 0; k < NK; k++) nmodes[k] = (int64_t)c->h_sv[5 * NK + k];
  return CAPOW_OK;
}

extern "C" int capow_grid_elems(const capow_ctx* c, int64_t* n) {
  if (!c || !n) return CAPOW_ERR_ARG;
  *n = c->grid_elems;
  return CAPOW_OK;
}

extern "C" int capow_get_grid(capow_ctx* c, int32_t slot, void* host, int64_t n) {
  if (!c || !host || slot < 0 || slot > 3) return CAPOW_ERR_ARG;
  if (!c->grid[slot]) return fail(c, CAPOW_ERR_ARG, "grid slot not allocated");
  if (n < c->grid_elems) return fail(c, CAPOW_ERR_ARG, "host buffer too small");
  int rc = finish_x(c);
  if (rc) return rc;
  const size_t w = (size_t)2 * c->g.nzh * c->esz, pitch = (size_t)c->g.rowlen * c->esz;
  HIPCHK(c, hipMemcpy2DAsync(host, w, c->grid[slot], pitch, w, (size_t)c->p.nx * c->p.ny,
                             hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return CAPOW_OK;
}

extern "C" int capow_set_grid(capow_ctx* c, int32_t slot, const void* host, int64_t n) {
  if (!c || !host || slot < 0 || slot > 3) return CAPOW_ERR_ARG;
  if (!c->grid[slot]) return fail(c, CAPOW_ERR_ARG, "grid slot not allocated");
  if (n < c->grid_elems) return fail(c, CAPOW_ERR_ARG, "host buffer too small");
  int rc = finish_x(c);   // the other slots stay consistent
  if (rc) return rc;
  c->fwd_assigned = c->fwd_kept = false;
  const size_t w = (size_t)2 * c->g.nzh * c->esz, pitch = (size_t)c->g.rowlen * c->esz;
  HIPCHK(c, hipMemcpy2DAsync(c->grid[slot], pitch, host, w, w, (size_t)c->p.nx * c->p.ny,
                             hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return CAPOW_OK;
}

extern "C" int capow_mark_assigned(capow_ctx* c, const capow_assign_stats* st) {
  if (!c) return CAPOW_ERR_ARG;
  c->x_pending = false;   // the caller's grids replace the pending transform
  if (st) c->last = *st;
  c->state = 2;
  c->fwd_assigned = c->fwd_kept = false;
  return CAPOW_OK;
}

extern "C" int capow_mark_transformed(capow_ctx* c) {
  if (!c) return CAPOW_ERR_ARG;
  c->x_pending = false;
  c->state = 3;
  c->fwd_assigned = c->fwd_kept = false;
  return CAPOW_OK;
}

extern "C" int capow_keep_forward(capow_ctx* c, int32_t on) {
  if (!c) return CAPOW_ERR_ARG;
  c->keep_fwd = on != 0;
  return CAPOW_OK;
}

extern "C" int capow_backward_reused(const capow_ctx* c, int32_t* reused) {
  if (!c || !reused) return CAPOW_ERR_ARG;
  *reused = c->bwd_reused ? 1 : 0;
  return CAPOW_OK;
}

extern "C" int capow_kernel_times(const capow_ctx* c, char* names, int32_t names_len, double* ms,
                                  int32_t* launches, int32_t max, int32_t* count) {
  if (!c || !count) return CAPOW_ERR_ARG;
  int n = (int)c->timer.names.size();
  if (n > max) n = max;
  int pos = 0;
  for (int i = 0; i < n; i++) {
    if (ms) ms[i] = c->timer.ms[i];
    if (launches) launches[i] = c->timer.launches[i];
    if (names) {
      const std::string& s = c->timer.names[i];
      if (pos + (int)s.size() + 1 <= names_len) {
        memcpy(names + pos, s.c_str(), s.size() + 1);
        pos += (int)s.size() + 1;
      }
    }
  }
  *count = n;
  return CAPOW_OK;
}

extern "C" int capow_reset_kernel_times(capow_ctx* c) {
  if (!c) return CAPOW_ERR_ARG;
  c->timer.reset();
  return CAPOW_OK;
}

// ---------------------------------------------------------------------------
// Backward: dL/d(raw bins) -> dL/d(x, y, z, v_los), dL/dw.
// ---------------------------------------------------------------------------
extern "C" int capow_backward(capow_ctx* c, const double* dL_draw, void* grad_xyzv, void* grad_w,
                              int32_t where) {
  if (!c) return CAPOW_ERR_ARG;
  if (c->p.edge_mode == CAPOW_EDGE_SURVEY)
    return fail(c, CAPOW_ERR_UNSUPPORTED, "capow_backward covers the periodic estimator only");
  if (!dL_draw || !grad_xyzv) return fail(c, CAPOW_ERR_ARG, "null gradient array");
  if (where != CAPOW_HOST && where != CAPOW_DEVICE)
    return fail(c, CAPOW_ERR_ARG, "where must be CAPOW_HOST or CAPOW_DEVICE");
  if (c->state < 1) return fail(c, CAPOW_ERR_STATE, "capow_backward before capow_set_particles");
  // the forward grids: kept from the forward (capow_keep_forward, nothing
  // changed since), or recomputed; the sort also records each row's input
  // index for the tile gather
  const bool tiles = c->bwd_tiles && c->n < ((int64_t)1 << 32) &&
                     gather_tiles_lds(c->g, c->fp64) <= 160 * 1024;
  const bool reuse = c->fwd_kept && c->state >= 3 && (!tiles || c->fwd_idx);
  int rc = CAPOW_OK;
  if (!reuse) {
    c->want_idx = tiles;
    rc = capow_assign(c, nullptr);
    c->want_idx = false;
    if (rc) return rc;
    rc = capow_fft(c);
  }
  c->bwd_reused = reuse;
  if (!rc) rc = finish_x(c);   // the binning backward needs the k-space grids
  if (rc) return rc;
  const size_t gbytes = (size_t)c->nspec * 5 * c->p.nk * 8;
  rc = ensure(c, (void**)&c->bwd_G, &c->bwd_G_cap, gbytes);
  if (rc) return rc;
  if (!c->bwd_g0) HIPCHK(c, hipMalloc((void**)&c->bwd_g0, 8 * sizeof(double)));
  HIPCHK(c, hipMemcpyAsync(c->bwd_G, dL_draw, gbytes, hipMemcpyHostToDevice, c->stream));
  BinLaunch L;
  bin_launch(c, L);
  rc = run_bin_backward(L, c->bwd_G, c->bwd_g0, c->fp64);
  if (rc) return fail(c, rc, std::string("binning backward failed: ") + hipGetErrorString(hipGetLastError()));
  FftArgs f;
  f.g = &c->g;
  f.stream = c->stream;
  f.timer = &c->timer;
  f.ngrids = 0;
  for (int s = 0; s < 4; s++)
    if (c->grid[s]) f.grid[f.ngrids++] = c->grid[s];
  for (int q = 0; q < 3; q++) {
    f.plan[q] = c->plan[q];
    f.tw[q] = c->tw[q];
  }
  f.col_mul = c->fft_col_mul;
  f.ncu = c->ncu;
  f.persist = c->fft_persist;
  rc = run_ifft(f, c->fp64);
  c->state = 1;   // the grids now hold gradients
  c->fwd_assigned = c->fwd_kept = false;
  if (rc == CAPOW_ERR_UNSUPPORTED)
    return fail(c, rc, "the backward needs register FFT sizes (16..1024 of 2^a 3^b with a split)");
  if (rc) return fail(c, rc, std::string("inverse fft failed: ") + hipGetErrorString(hipGetLastError()));
  GatherArgs ga;
  ga.g = &c->g;
  ga.stream = c->stream;
  ga.timer = &c->timer;
  ga.parts = c->parts;
  ga.wts = c->wts;
  ga.n = c->n;
  for (int s = 0; s < 4; s++) ga.grid[s] = c->grid[s] ? c->grid[s] : c->grid[0];
  ga.g0 = c->bwd_g0;
  if (tiles) {
    if (!c->bwd_err) HIPCHK(c, hipMalloc((void**)&c->bwd_err, sizeof(unsigned int)));
    if (zero_u32(c->bwd_err, 1, c->stream) != CAPOW_OK)
      return fail(c, CAPOW_ERR_HIP, "clearing the gather flag failed");
    ga.sorted = c->sorted;
    ga.sorted_w = c->sorted_w;
    ga.sorted_idx = c->sorted_idx;
    ga.offsets = c->offsets;
    ga.errs = c->bwd_err;
  }
  const size_t ob = (size_t)(c->n > 0 ? c->n : 1) * 4 * c->esz, wb = (size_t)(c->n > 0 ? c->n : 1) * c->esz;
  if (where == CAPOW_DEVICE) {
    ga.gout = grad_xyzv;
    ga.gw = grad_w;
  } else {
    rc = ensure(c, &c->bwd_out, &c->bwd_out_cap, ob);
    if (!rc && grad_w) rc = ensure(c, &c->bwd_w, &c->bwd_w_cap, wb);
    if (rc) return rc;
    ga.gout = c->bwd_out;
    ga.gw = grad_w ? c->bwd_w : nullptr;
  }
  c->g.has_w = c->wts != nullptr;
  rc = run_gather(ga, c->fp64);
  if (rc) return fail(c, rc, std::string("gather failed: ") + hipGetErrorString(hipGetLastError()));
  if (where == CAPOW_HOST && c->n > 0) {
    HIPCHK(c, hipMemcpyAsync(grad_xyzv, c->bwd_out, (size_t)c->n * 4 * c->esz, hipMemcpyDeviceToHost, c->stream));
    if (grad_w) HIPCHK(c, hipMemcpyAsync(grad_w, c->bwd_w, (size_t)c->n * c->esz, hipMemcpyDeviceToHost, c->stream));
  }
  unsigned int gerr = 0;
  if (tiles) HIPCHK(c, hipMemcpyAsync(&gerr, c->bwd_err, sizeof gerr, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (gerr) return fail(c, CAPOW_ERR_INTERNAL, "tile gather: a tap outside its tile's region");
  c->timer.collect();
  return CAPOW_OK;
}

// ---------------------------------------------------------------------------
// Batches of catalogues.
// ---------------------------------------------------------------------------
struct capow_batch {
  std::vector<capow_params> cfg;
  int nconf = 0, nstreams = 0;
  int nactive = 0;               // streams the runs use (<= nstreams; the autotuner picks)
  int64_t max_n = 0;
  std::vector<hipStream_t> streams;
  std::vector<capow_ctx*> ctx;   // [stream * nconf + config]
  size_t stride = 0;             // doubles of staging per mock
  double* stage = nullptr;
  size_t stage_mocks = 0;
  std::vector<hipEvent_t> ev;
  std::string err;
  // graph mode: the whole run (every stream's stages of every catalogue)
  // captured into one hipGraph, replayed while the catalogue list (device
  // pointers, counts, parameter sets), the streams in use and the staging
  // buffer are unchanged
  bool graph = false;
  hipGraphExec_t gexec = nullptr;
  std::vector<capow_mock> gkey;
  int gactive = 0;
  double* gstage = nullptr;
  std::vector<hipEvent_t> gev;   // fork / join of the captured streams
  // stacked mode (capow_batch_set_stack): one stack context per parameter
  // set, on stream config % nactive; a run orders the catalogues by
  // (parameter set, weights or not) and sends each such group through its
  // set's context in stacks of up to `stack` catalogues, one launch per
  // stage per stack.  The staging then holds n stats rows, n flag rows and
  // the stacks' bins (stack at position j0: from j0 * stride_bins on).
  int stack = 0;
  std::vector<capow_ctx*> sctx;  // [config]
  MockDev* tab = nullptr;        // device table, one entry per catalogue position
  MockDev* htab = nullptr;       // its pinned staging
  size_t tab_mocks = 0;
  int pc_max = 0;                // largest bin copy (doubles) over the sets
  int gstack = 0;
};

// A run's stacks: the catalogue at each position (grouped by parameter set
// and weights) and every stack's first position and size.
struct StackPlan {
  std::vector<int> order;        // position -> catalogue index
  std::vector<int> first, size;  // per stack
};

static void stack_plan(const capow_batch* b, const capow_mock* mocks, int n, StackPlan& sp) {
  sp.order.clear();
  sp.first.clear();
  sp.size.clear();
  for (int q = 0; q < b->nconf; q++)
    for (int hw = 0; hw < 2; hw++) {
      int in_stack = 0;
      for (int i = 0; i < n; i++) {
        if (mocks[i].config != q || (mocks[i].w != nullptr) != (hw != 0)) continue;
        if (in_stack == 0) {
          sp.first.push_back((int)sp.order.size());
          sp.size.push_back(0);
        }
        sp.order.push_back(i);
        sp.size.back()++;
        if (++in_stack == b->stack) in_stack = 0;
      }
    }
}

static int bfail(capow_batch* b, int code, const std::string& m) {
  if (b) b->err = m;
  else g_create_err = m;
  return code;
}

extern "C" int capow_batch_destroy(capow_batch* b) {
  if (!b) return CAPOW_OK;
  for (hipStream_t s : b->streams) hipStreamSynchronize(s);
  if (b->gexec) hipGraphExecDestroy(b->gexec);
  for (hipEvent_t e : b->gev) hipEventDestroy(e);
  for (capow_ctx* c : b->ctx) capow_destroy(c);
  for (capow_ctx* c : b->sctx) capow_destroy(c);
  if (b->tab) hipFree(b->tab);
  if (b->htab) hipHostFree(b->htab);
  for (hipEvent_t e : b->ev) hipEventDestroy(e);
  for (hipStream_t s : b->streams) hipStreamDestroy(s);
  if (b->stage) hipHostFree(b->stage);
  delete b;
  return CAPOW_OK;
}

static void batch_drop_graph(capow_batch* b);

extern "C" int capow_batch_set_graph(capow_batch* b, int32_t on) {
  if (!b) return bfail(b, CAPOW_ERR_ARG, "null batch");
  for (hipStream_t s : b->streams) hipStreamSynchronize(s);
  batch_drop_graph(b);
  b->graph = on != 0;
  return CAPOW_OK;
}

// Stacked mode: a stack context per parameter set holding `m` catalogues
// (fewer when m times the largest catalogue would pass 2^32 - 4097 sorted
// rows), its sort buffers sized for m of the largest catalogues and its
// tunables copied from the set's stream-0 context.  m = 0 returns to the
// per-catalogue streams.  Every set must support stacks, or the mode stays
// as it was.
extern "C" int capow_batch_set_stack(capow_batch* b, int32_t m) {
  if (!b || m < 0) return bfail(b, CAPOW_ERR_ARG, "bad stack size");
  for (hipStream_t s : b->streams) hipStreamSynchronize(s);
  batch_drop_graph(b);
  if (m == 0) {
    b->stack = 0;
    for (capow_ctx* d : b->sctx) capow_destroy(d);
    b->sctx.clear();
    return CAPOW_OK;
  }
  const int64_t per = b->max_n > 0 ? b->max_n : 1;
  int64_t mm = std::min<int64_t>(m, ((int64_t)UINT32_MAX - 4096) / per);
  if (mm < 1) mm = 1;
  std::vector<capow_ctx*> made;
  for (int q = 0; q < b->nconf; q++) {
    capow_ctx* c = nullptr;
    int rc = create_ctx(&b->cfg[q], (int)mm, &c);
    if (!rc) {
      capow_ctx* c0 = b->ctx[q];
      c->count_blocks = c0->count_blocks;
      c->g.sort_probes = c0->g.sort_probes;
      c->dep_halo = c0->dep_halo;
      c->dep_bank = c0->dep_bank;
      c->fft_rows_mul = c0->fft_rows_mul;
      c->fft_order = c0->fft_order;
      c->cap_fixed = true;
      const size_t rows = (size_t)mm * (size_t)per + 4096;
      rc = ensure(c, &c->sorted, &c->sorted_cap, rows * 4 * c->esz);
      if (!rc) rc = ensure(c, &c->sorted_w, &c->sorted_w_cap, rows * c->esz);
    }
    if (rc) {
      const std::string m2 = c ? c->err : g_create_err;
      if (c) capow_destroy(c);
      for (capow_ctx* d : made) capow_destroy(d);
      return bfail(b, rc, "stack for config " + std::to_string(q) + ": " + m2);
    }
    made.push_back(c);
  }
  for (capow_ctx* d : b->sctx) capow_destroy(d);
  b->sctx = made;
  b->stack = (int)mm;
  return CAPOW_OK;
}

extern "C" const char* capow_batch_last_error(const capow_batch* b) {
  return b ? b->err.c_str() : g_create_err.c_str();
}

extern "C" int capow_batch_create(const capow_params* configs, int32_t nconf,
                                  int64_t max_particles, int32_t nstreams, capow_batch** out) {
  if (!configs || !out || nconf < 1 || nstreams < 1 || max_particles < 0)
    return bfail(nullptr, CAPOW_ERR_ARG, "bad batch arguments");
  if (max_particles > (int64_t)UINT32_MAX - 4096)
    return bfail(nullptr, CAPOW_ERR_UNSUPPORTED, "more than 2^32 - 4097 particles per catalogue");
  for (int q = 0; q < nconf; q++)
    if (configs[q].edge_mode == CAPOW_EDGE_SURVEY)
      return bfail(nullptr, CAPOW_ERR_UNSUPPORTED, "batches cover the periodic estimator only");
  *out = nullptr;
  capow_batch* b = new capow_batch();
  b->cfg.assign(configs, configs + nconf);
  b->nconf = nconf;
  b->nstreams = nstreams;
  b->nactive = nstreams;
  b->max_n = max_particles;
  size_t per_copy_max = 0;
  for (int s = 0; s < nstreams; s++) {
    hipStream_t st;
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) {
      capow_batch_destroy(b);
      return bfail(nullptr, CAPOW_ERR_HIP, "hipStreamCreate failed");
    }
    b->streams.push_back(st);
    for (int q = 0; q < nconf; q++) {
      capow_ctx* c = nullptr;
      int rc = capow_create(&configs[q], &c);
      if (rc) {
        std::string m = g_create_err;
        capow_batch_destroy(b);
        return bfail(nullptr, rc, "batch config " + std::to_string(q) + ": " + m);
      }
      b->ctx.push_back(c);
      capow_set_stream(c, st);
      c->cap_fixed = true;
      // one sorted copy per particle
      const size_t rows = (size_t)max_particles + 4096;
      int r1 = ensure(c, &c->sorted, &c->sorted_cap, rows * 4 * c->esz);
      int r2 = r1 ? r1 : ensure(c, &c->sorted_w, &c->sorted_w_cap, rows * c->esz);
      if (r1 || r2) {
        std::string m = c->err;
        capow_batch_destroy(b);
        return bfail(nullptr, CAPOW_ERR_NOMEM, "batch sort buffers: " + m);
      }
      if ((size_t)c->bin_per_copy > per_copy_max) per_copy_max = (size_t)c->bin_per_copy;
    }
  }
  b->stride = STAT_TOTAL + 2 + per_copy_max;
  b->pc_max = (int)per_copy_max;
  *out = b;
  return CAPOW_OK;
}

// Enqueues every catalogue's stages (particles, assignment, transform,
// binning) round-robin over the active streams; timed: events around each
// catalogue's stages (not while capturing).  Returns the first failure.
static int batch_enqueue(capow_batch* b, const capow_mock* mocks, int n, bool timed,
                         std::vector<int>& launched) {
  int first = CAPOW_OK;
  for (int i = 0; i < n; i++) {
    const int s = i % b->nactive;
    capow_ctx* c = b->ctx[(size_t)s * b->nconf + mocks[i].config];
    double* st = b->stage + (size_t)i * b->stride;
    unsigned int* er = reinterpret_cast<unsigned int*>(st + STAT_TOTAL);
    double* bins = st + STAT_TOTAL + 2;
    if (timed) hipEventRecord(b->ev[2 * i], b->streams[s]);
    int rc = capow_set_particles(c, mocks[i].xyzv, mocks[i].w, mocks[i].n,
                                 c->fp64 ? CAPOW_DTYPE_F64 : CAPOW_DTYPE_F32, CAPOW_DEVICE);
    if (!rc) rc = assign_enqueue(c, st, er, false);
    if (!rc) rc = capow_fft(c);
    if (!rc) rc = bin_enqueue(c, bins);
    if (timed) hipEventRecord(b->ev[2 * i + 1], b->streams[s]);
    if (rc) {
      if (!first) {
        first = rc;
        b->err = "mock " + std::to_string(i) + ": " + c->err;
      }
      continue;
    }
    launched[i] = 1;
  }
  return first;
}

// Enqueues one stack (positions j0 .. j0 + k - 1 of the plan, catalogues of
// one parameter set with or without weights) on its set's stack context:
// the table upload, the sort and deposit of every catalogue, the z and y
// passes and the fused x pass + binning, each one launch for the stack, then
// the copies of the sums, flags and bins into the run's staging (n
// positions in all).
static int stack_enqueue(capow_batch* b, capow_ctx* c, const capow_mock* mocks,
                         const StackPlan& sp, int st, int n) {
  const int j0 = sp.first[st], k = sp.size[st];
  Geom& g = c->g;
  hipStream_t s = c->stream;
  int64_t rows = 0, want = 1;
  for (int j = j0; j < j0 + k; j++) {
    const capow_mock& m = mocks[sp.order[j]];
    b->htab[j] = MockDev{m.xyzv, m.w, m.n};
    rows += m.n;
    want = std::max(want, (m.n + BK_CHUNK - 1) / BK_CHUNK);
  }
  g.has_w = mocks[sp.order[j0]].w != nullptr;
  const int64_t cap = (int64_t)(c->sorted_cap / (4 * c->esz));
  if (rows > cap || (g.has_w && (int64_t)(c->sorted_w_cap / c->esz) < rows))
    return fail(c, CAPOW_ERR_NOMEM, "stack sort buffer too small");
  HIPCHK(c, hipMemcpyAsync(b->tab + j0, b->htab + j0, (size_t)k * sizeof(MockDev),
                           hipMemcpyHostToDevice, s));
  if (zero_u32(c->counts, g.nbuckets * k, s) || zero_u32(c->errs, 4 * (int64_t)k, s))
    return fail(c, CAPOW_ERR_HIP, "clearing the stack's sort counters failed");
  DepositArgs a;
  a.g = &g;
  a.stream = s;
  a.timer = &c->timer;
  a.parts = nullptr;
  a.wts = nullptr;
  a.n = 0;
  a.nmock = k;
  a.mocks = b->tab + j0;
  // each catalogue's k_bucket blocks: as for a context of its own
  a.count_blocks = (int)std::min<int64_t>(want, c->count_blocks);
  a.counts = c->counts;
  a.offsets = c->offsets;
  a.cursor = c->cursor;
  a.bsum = c->bsum;
  a.boff = c->boff;
  a.partials = c->partials;
  a.errs = c->errs;
  a.stats = c->stats;
  for (int q = 0; q < 4; q++) a.grid[q] = c->grid[q];
  a.halo = c->dep_halo;
  a.bank = c->dep_bank;
  int rc = run_count(a, c->fp64);
  if (rc) return fail(c, rc, std::string("stack count failed: ") + hipGetErrorString(hipGetLastError()));
  a.sorted = c->sorted;
  a.sorted_w = c->sorted_w;
  a.sorted_cap = cap;
  rc = run_scatter_deposit(a, c->fp64);
  if (rc) return fail(c, rc, std::string("stack deposit failed: ") + hipGetErrorString(hipGetLastError()));
  FftArgs f;
  fft_args(c, f);
  f.nmock = k;
  f.mock_stride = g.mock_stride;
  f.defer_x = true;
  f.ssq = c->xbin_ssq;
  f.xbin_factor = c->xbin_factor;
  f.xq = c->xbin_q;
  rc = run_fft(f, c->fp64);
  if (!rc && !f.x_deferred) rc = CAPOW_ERR_UNSUPPORTED;
  if (rc) return fail(c, rc, std::string("stack transform failed: ") + hipGetErrorString(hipGetLastError()));
  BinLaunch L;
  bin_launch(c, L);
  L.out_stride = c->bin_per_copy;
  rc = run_fft_x_bin(f, L, c->xbin_part, c->xbin_rows);
  if (rc) return fail(c, rc, std::string("stack binning failed: ") + hipGetErrorString(hipGetLastError()));
  double* stage = b->stage;
  HIPCHK(c, hipMemcpyAsync(stage + (size_t)j0 * STAT_TOTAL, c->stats, (size_t)k * STAT_TOTAL * 8,
                           hipMemcpyDeviceToHost, s));
  HIPCHK(c, hipMemcpyAsync(reinterpret_cast<unsigned int*>(stage + (size_t)n * STAT_TOTAL) + (size_t)j0 * 4,
                           c->errs, (size_t)k * 4 * sizeof(unsigned int), hipMemcpyDeviceToHost, s));
  HIPCHK(c, hipMemcpyAsync(stage + (size_t)n * (STAT_TOTAL + 2) + (size_t)j0 * b->pc_max, c->bin_out,
                           (size_t)k * c->bin_per_copy * 8, hipMemcpyDeviceToHost, s));
  return CAPOW_OK;
}

// Every stack of the plan, stack contexts on stream config % nactive;
// timed: events around each stack (not while capturing).
static int stacks_enqueue(capow_batch* b, const capow_mock* mocks, int n, const StackPlan& sp,
                          bool timed, std::vector<int>& launched) {
  int first = CAPOW_OK;
  for (size_t st = 0; st < sp.first.size(); st++) {
    const int j0 = sp.first[st];
    const int q = mocks[sp.order[j0]].config;
    capow_ctx* c = b->sctx[q];
    capow_set_stream(c, b->streams[q % b->nactive]);
    if (timed) hipEventRecord(b->ev[2 * j0], c->stream);
    const int rc = stack_enqueue(b, c, mocks, sp, (int)st, n);
    if (timed) hipEventRecord(b->ev[2 * j0 + 1], c->stream);
    if (rc) {
      if (!first) {
        first = rc;
        b->err = "stack at mock " + std::to_string(sp.order[j0]) + ": " + c->err;
      }
      continue;
    }
    for (int j = j0; j < j0 + sp.size[st]; j++) launched[sp.order[j]] = 1;
  }
  return first;
}

static int any_enqueue(capow_batch* b, const capow_mock* mocks, int n, bool timed,
                       std::vector<int>& launched) {
  if (b->stack <= 0) return batch_enqueue(b, mocks, n, timed, launched);
  StackPlan sp;
  stack_plan(b, mocks, n, sp);
  return stacks_enqueue(b, mocks, n, sp, timed, launched);
}

static void batch_drop_graph(capow_batch* b) {
  if (b->gexec) hipGraphExecDestroy(b->gexec);
  b->gexec = nullptr;
  b->gkey.clear();
}

// Captures the enqueue of `mocks` on every active stream into one graph
// (stream 0 forks to the others through an event and joins them at the
// end) and instantiates it; nothing runs.  The enqueue allocates nothing
// and synchronises nothing once the list has run eagerly, and every clear
// inside it is a kernel, so the graph holds kernels and copies only.
static int batch_capture(capow_batch* b, const capow_mock* mocks, int n) {
  batch_drop_graph(b);
  while (b->gev.size() < (size_t)b->nstreams) {
    hipEvent_t e;
    if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess)
      return bfail(b, CAPOW_ERR_HIP, "hipEventCreate failed");
    b->gev.push_back(e);
  }
  hipStream_t s0 = b->streams[0];
  if (hipStreamBeginCapture(s0, hipStreamCaptureModeThreadLocal) != hipSuccess)
    return bfail(b, CAPOW_ERR_HIP, std::string("hipStreamBeginCapture: ") + hipGetErrorString(hipGetLastError()));
  hipError_t e = hipEventRecord(b->gev[0], s0);
  for (int s = 1; s < b->nactive && e == hipSuccess; s++) e = hipStreamWaitEvent(b->streams[s], b->gev[0], 0);
  std::vector<int> launched(n, 0);
  int rc = e == hipSuccess ? any_enqueue(b, mocks, n, false, launched) : CAPOW_ERR_HIP;
  for (int s = 1; s < b->nactive && e == hipSuccess; s++) {
    e = hipEventRecord(b->gev[s], b->streams[s]);
    if (e == hipSuccess) e = hipStreamWaitEvent(s0, b->gev[s], 0);
  }
  hipGraph_t graph = nullptr;
  const hipError_t e1 = hipStreamEndCapture(s0, &graph);
  if (e == hipSuccess) e = e1;
  if (e == hipSuccess && !rc && graph) e = hipGraphInstantiate(&b->gexec, graph, nullptr, nullptr, 0);
  if (graph) hipGraphDestroy(graph);
  if (rc) {
    b->gexec = nullptr;
    return rc;
  }
  if (e != hipSuccess || !b->gexec) {
    b->gexec = nullptr;
    return bfail(b, CAPOW_ERR_HIP, std::string("batch graph capture failed: ") + hipGetErrorString(e));
  }
  b->gkey.assign(mocks, mocks + n);
  b->gactive = b->nactive;
  b->gstage = b->stage;
  b->gstack = b->stack;
  return CAPOW_OK;
}

extern "C" int capow_batch_run(capow_batch* b, const capow_mock* mocks, int32_t n,
                               capow_bins* out, capow_assign_stats* stats, float* ms) {
  if (!b || (n > 0 && (!mocks || !out))) return bfail(b, CAPOW_ERR_ARG, "bad batch run arguments");
  for (int i = 0; i < n; i++) {
    if (mocks[i].config < 0 || mocks[i].config >= b->nconf)
      return bfail(b, CAPOW_ERR_ARG, "mock " + std::to_string(i) + ": config out of range");
    if (mocks[i].n < 0 || mocks[i].n > b->max_n)
      return bfail(b, CAPOW_ERR_ARG, "mock " + std::to_string(i) + ": particle count above the batch maximum");
  }
  if (b->stage_mocks < (size_t)n) {
    if (b->stage) hipHostFree(b->stage);
    b->stage = nullptr;
    b->stage_mocks = 0;
    if (hipHostMalloc((void**)&b->stage, (size_t)n * b->stride * 8, hipHostMallocDefault) != hipSuccess)
      return bfail(b, CAPOW_ERR_NOMEM, "batch staging allocation failed");
    b->stage_mocks = (size_t)n;
  }
  while (b->ev.size() < (size_t)2 * n) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return bfail(b, CAPOW_ERR_HIP, "hipEventCreate failed");
    b->ev.push_back(e);
  }
  if (b->stack > 0 && b->tab_mocks < (size_t)n) {
    if (b->tab) hipFree(b->tab);
    if (b->htab) hipHostFree(b->htab);
    b->tab = nullptr;
    b->htab = nullptr;
    b->tab_mocks = 0;
    if (hipMalloc((void**)&b->tab, (size_t)n * sizeof(MockDev)) != hipSuccess ||
        hipHostMalloc((void**)&b->htab, (size_t)n * sizeof(MockDev), hipHostMallocDefault) != hipSuccess)
      return bfail(b, CAPOW_ERR_NOMEM, "batch catalogue table allocation failed");
    b->tab_mocks = (size_t)n;
  }
  int first = CAPOW_OK;
  bool fell_back = false;
  std::vector<int> launched(n, 0);
  const bool replay = b->graph && b->gexec && b->gkey.size() == (size_t)n &&
                      b->gactive == b->nactive && b->gstage == b->stage && b->gstack == b->stack &&
                      memcmp(b->gkey.data(), mocks, (size_t)n * sizeof(capow_mock)) == 0;
  if (replay) {
    if (hipGraphLaunch(b->gexec, b->streams[0]) != hipSuccess)
      return bfail(b, CAPOW_ERR_HIP, std::string("hipGraphLaunch: ") + hipGetErrorString(hipGetLastError()));
    for (int i = 0; i < n; i++) launched[i] = 1;
  } else {
    first = any_enqueue(b, mocks, n, true, launched);
  }
  for (hipStream_t s : b->streams)
    if (hipStreamSynchronize(s) != hipSuccess)
      return bfail(b, CAPOW_ERR_HIP, std::string("batch stream: ") + hipGetErrorString(hipGetLastError()));
  if (b->stack > 0) {
    // stacked: per position, the sums, flags and bins its stack staged
    StackPlan sp;
    stack_plan(b, mocks, n, sp);
    for (size_t st = 0; st < sp.first.size(); st++) {
      const int j0 = sp.first[st];
      float t = -1.f;   // the stack's GPU time (not while replaying)
      if (ms && !replay) hipEventElapsedTime(&t, b->ev[2 * j0], b->ev[2 * j0 + 1]);
      for (int j = j0; j < j0 + sp.size[st]; j++) {
        const int i = sp.order[j];
        if (ms) ms[i] = t;
        if (!launched[i]) continue;
        capow_ctx* c = b->sctx[mocks[i].config];
        const double* hs = b->stage + (size_t)j * STAT_TOTAL;
        const unsigned int* he =
            reinterpret_cast<const unsigned int*>(b->stage + (size_t)n * STAT_TOTAL) + (size_t)j * 4;
        c->n = mocks[i].n;
        int rc = assign_finish(c, hs, he, stats ? &stats[i] : nullptr);
        if (!rc)
          bin_finish(c, b->stage + (size_t)n * (STAT_TOTAL + 2) + (size_t)j0 * b->pc_max +
                            (size_t)(j - j0) * c->bin_per_copy,
                     &out[i]);
        if (rc && !first) {
          first = rc;
          b->err = "mock " + std::to_string(i) + ": " + c->err;
        }
      }
    }
    if (b->graph && !replay && !first) return batch_capture(b, mocks, n);
    return first;
  }
  for (int i = 0; i < n; i++) {
    if (!launched[i]) continue;
    const int s = i % b->nactive;
    capow_ctx* c = b->ctx[(size_t)s * b->nconf + mocks[i].config];
    const double* st = b->stage + (size_t)i * b->stride;
    const unsigned int* er = reinterpret_cast<const unsigned int*>(st + STAT_TOTAL);
    c->n = mocks[i].n;   // the context ran later mocks since
    int rc = assign_finish(c, st, er, stats ? &stats[i] : nullptr);
    if (rc == CAPOW_ERR_NOMEM) {
      // the catalogue needed more sorted copies than reserved: run it alone
      // with an exactly sized buffer
      // (its exactly sized buffer replaces the one a captured graph uses)
      batch_drop_graph(b);
      fell_back = true;
      int r = capow_set_particles(c, mocks[i].xyzv, mocks[i].w, mocks[i].n,
                                  c->fp64 ? CAPOW_DTYPE_F64 : CAPOW_DTYPE_F32, CAPOW_DEVICE);
      if (!r) r = capow_assign(c, stats ? &stats[i] : nullptr);
      if (!r) r = capow_fft(c);
      if (!r) r = capow_bin(c, &out[i]);
      rc = r;
    } else if (!rc) {
      bin_finish(c, st + STAT_TOTAL + 2, &out[i]);
    }
    if (ms) {
      float t = -1.f;   // a replayed graph does not time the catalogues one by one
      if (!replay) hipEventElapsedTime(&t, b->ev[2 * i], b->ev[2 * i + 1]);
      ms[i] = t;
    }
    if (rc && !first) {
      first = rc;
      b->err = "mock " + std::to_string(i) + ": " + c->err;
    }
  }
  // graph mode: after an eager run of a new list (which left every buffer
  // sized and the mode counts computed), capture the same list for replay
  if (b->graph && !replay && !first && !fell_back) return batch_capture(b, mocks, n);
  return first;
}

// ---- autotuner ------------------------------------------------------------
// The reference plans its FFTs with FFTW_ESTIMATE (src/main.c:1200-1217) and
// has no other tunable stage.  Here each tunable of a context is timed on
// the particles it holds, one tunable at a time (coordinate descent), and
// the fastest candidate kept:
//   sort_blocks  workgroups of the two sort passes per CU (4, 8, 12)
//   sort_lds     LDS run aggregation in the sort (on) or the global
//                counters for every run (off)
//   dep_halo     halo rows in the FP32 deposit tiles
//   dep_bank     bank-balanced copy order and masked rows in those tiles
//   fft_runs     write-run width of the out-of-place FFT passes (auto:
//                128 bytes, or 64)
//   fft_order    batch order of the z/y/x passes (bit per pass: outer
//                index fastest)
//   xbin         the x pass fused with the binning (FP32)
// Every choice but xbin gives bitwise the same grids and bins (integer
// deposit sums, the same per-row transforms); xbin's sums differ from the
// separate binning's at the FP32 rounding level.

struct TuneKnob {
  const char* name;
  int* value;
  std::vector<int> cand;
};

static int tune_time(capow_ctx* c, int reps, float* med) {
  hipEvent_t a, b;
  if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) return CAPOW_ERR_HIP;
  std::vector<float> t;
  capow_bins out;
  memset(&out, 0, sizeof out);   // null arrays: the sums stay on the device
  int rc = CAPOW_OK;
  for (int r = 0; r <= reps && !rc; r++) {   // r = 0: warm-up
    hipEventRecord(a, c->stream);
    rc = capow_assign(c, nullptr);
    if (!rc) rc = capow_fft(c);
    if (!rc) rc = capow_bin(c, &out);
    hipEventRecord(b, c->stream);
    hipEventSynchronize(b);
    float ms = 0.f;
    hipEventElapsedTime(&ms, a, b);
    if (r > 0) t.push_back(ms);
  }
  hipEventDestroy(a);
  hipEventDestroy(b);
  if (rc) return rc;
  std::sort(t.begin(), t.end());
  *med = t[t.size() / 2];
  return CAPOW_OK;
}

static void tune_note(std::string& rep, const char* fmt, ...) {
  char buf[256];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  rep += buf;
}

extern "C" int capow_autotune(capow_ctx* c, int32_t reps, char* report, int32_t report_len) {
  if (!c) return CAPOW_ERR_ARG;
  if (c->p.edge_mode == CAPOW_EDGE_SURVEY)
    return fail(c, CAPOW_ERR_UNSUPPORTED, "survey context: nothing to tune");
  if (c->state < 1) return fail(c, CAPOW_ERR_STATE, "capow_autotune before capow_set_particles");
  if (reps < 1) reps = 3;
  const bool xbin_avail = c->xbin_ok || c->xbin_part != nullptr;
  int sort_blocks = c->count_blocks / (c->ncu > 0 ? c->ncu : 256);
  int sort_lds = c->g.sort_probes > 0;
  int halo = c->dep_halo ? 1 : 0, runs = c->fft_rows_mul, order = c->fft_order;
  int bank = c->dep_bank ? 1 : 0;
  int xbin = c->xbin_ok ? 1 : 0;
  std::vector<TuneKnob> knobs = {
      {"sort_blocks", &sort_blocks, {12, 8, 4}},
      {"sort_lds", &sort_lds, {1, 0}},
      {"dep_halo", &halo, {1, 0}},
      {"dep_bank", &bank, {1, 0}},
      {"fft_runs", &runs, {0, 1}},
      {"fft_order", &order, {c->fp64 ? 0 : 5, 0, 7, 2}},
  };
  if (xbin_avail) knobs.push_back({"xbin", &xbin, {1, 0}});
  auto apply = [&]() {
    c->count_blocks = sort_blocks * (c->ncu > 0 ? c->ncu : 256);
    c->g.sort_probes = sort_lds ? 32 : 0;
    c->dep_halo = halo != 0;
    c->dep_bank = bank != 0;
    c->fft_rows_mul = runs;
    c->fft_order = order;
    c->xbin_ok = xbin_avail && xbin != 0;
  };
  std::string rep;
  tune_note(rep, "autotune %lldx%lldx%lld %s, %lld particles, median of %d\n", (long long)c->p.nx,
            (long long)c->p.ny, (long long)c->p.nz, c->fp64 ? "FP64" : "FP32", (long long)c->n, reps);
  for (auto& k : knobs) {
    int best = *k.value;
    float best_ms = 1e30f;
    for (int v : k.cand) {
      *k.value = v;
      apply();
      float ms = 0.f;
      const int rc = tune_time(c, reps, &ms);
      if (rc) return rc;
      tune_note(rep, "  %s=%d: %.3f ms\n", k.name, v, ms);
      if (ms < best_ms) {
        best_ms = ms;
        best = v;
      }
    }
    *k.value = best;
    apply();
  }
  tune_note(rep,
            "chosen: sort_blocks=%d sort_lds=%d dep_halo=%d dep_bank=%d fft_runs=%d fft_order=%d "
            "xbin=%d\n",
            sort_blocks, sort_lds, halo, bank, runs, order, c->xbin_ok ? 1 : 0);
  if (report && report_len > 0) {
    strncpy(report, rep.c_str(), (size_t)report_len - 1);
    report[report_len - 1] = 0;
  }
  return CAPOW_OK;
}

extern "C" int capow_batch_autotune(capow_batch* b, const capow_mock* mocks, int32_t n,
                                    int32_t reps, char* report, int32_t report_len) {
  if (!b || !mocks || n < 1) return bfail(b, CAPOW_ERR_ARG, "bad batch autotune arguments");
  if (reps < 1) reps = 3;
  for (hipStream_t st : b->streams) hipStreamSynchronize(st);
  batch_drop_graph(b);   // the launches it holds follow the old choices
  std::string rep;
  std::vector<char> buf(4096);
  for (int q = 0; q < b->nconf; q++) {
    int m = -1;
    for (int i = 0; i < n && m < 0; i++)
      if (mocks[i].config == q) m = i;
    if (m < 0) continue;
    capow_ctx* c0 = b->ctx[q];
    int rc = capow_set_particles(c0, mocks[m].xyzv, mocks[m].w, mocks[m].n,
                                 c0->fp64 ? CAPOW_DTYPE_F64 : CAPOW_DTYPE_F32, CAPOW_DEVICE);
    if (!rc) rc = capow_autotune(c0, reps, buf.data(), (int32_t)buf.size());
    if (rc) return bfail(b, rc, "autotune config " + std::to_string(q) + ": " + c0->err);
    rep += buf.data();
    for (int s = 1; s < b->nstreams; s++) {
      capow_ctx* c = b->ctx[(size_t)s * b->nconf + q];
      c->count_blocks = c0->count_blocks;
      c->g.sort_probes = c0->g.sort_probes;
      c->dep_halo = c0->dep_halo;
      c->dep_bank = c0->dep_bank;
      c->fft_rows_mul = c0->fft_rows_mul;
      c->fft_order = c0->fft_order;
      c->xbin_ok = c0->xbin_ok;
    }
  }
  // streams in use: the whole batch timed with each candidate count
  std::vector<capow_bins> out(n);
  for (auto& o : out) memset(&o, 0, sizeof o);
  int best = b->nactive;
  double best_s = 1e30;
  for (int ns = 2; ns <= b->nstreams; ns += 2) {
    b->nactive = ns;
    std::vector<double> t;
    for (int r = 0; r <= reps; r++) {
      for (hipStream_t st : b->streams) hipStreamSynchronize(st);
      const auto t0 = std::chrono::steady_clock::now();
      const int rc = capow_batch_run(b, mocks, n, out.data(), nullptr, nullptr);
      const auto t1 = std::chrono::steady_clock::now();
      if (rc) return rc;
      if (r > 0) t.push_back(std::chrono::duration<double>(t1 - t0).count());
    }
    std::sort(t.begin(), t.end());
    tune_note(rep, "  streams=%d: %.3f ms for %d catalogues\n", ns, t[t.size() / 2] * 1e3, n);
    if (t[t.size() / 2] < best_s) {
      best_s = t[t.size() / 2];
      best = ns;
    }
  }
  b->nactive = best;
  tune_note(rep, "chosen: streams=%d\n", best);
  // then stacks (one launch per stage for up to m catalogues of a set)
  // against the per-catalogue streams, where every set supports them
  const int stack0 = b->stack;
  int best_stack = 0;
  if (stack0 > 0) capow_batch_set_stack(b, 0);
  for (int m : {16, 64}) {
    if (capow_batch_set_stack(b, m) != CAPOW_OK) {
      tune_note(rep, "  stack=%d: unsupported (%s)\n", m, b->err.c_str());
      break;
    }
    std::vector<double> t;
    for (int r = 0; r <= reps; r++) {
      for (hipStream_t st : b->streams) hipStreamSynchronize(st);
      const auto t0 = std::chrono::steady_clock::now();
      const int rc = capow_batch_run(b, mocks, n, out.data(), nullptr, nullptr);
      const auto t1 = std::chrono::steady_clock::now();
      if (rc) return rc;
      if (r > 0) t.push_back(std::chrono::duration<double>(t1 - t0).count());
    }
    std::sort(t.begin(), t.end());
    tune_note(rep, "  stack=%d: %.3f ms for %d catalogues\n", m, t[t.size() / 2] * 1e3, n);
    if (t[t.size() / 2] < best_s) {
      best_s = t[t.size() / 2];
      best_stack = m;
    }
  }
  capow_batch_set_stack(b, best_stack);
  tune_note(rep, "chosen: stack=%d\n", best_stack);
  if (best_stack > 0) {
    // a stack's best settings can differ from one catalogue's (many
    // catalogues per launch), so each set's stack context is tuned again,
    // starting from the library defaults and timing the whole batch: the
    // sets' stacks run on their own streams and share the GPU, and a
    // setting timed on one set alone does not carry over to the shared run
    auto time_all = [&](double* med) -> int {
      std::vector<double> t;
      for (int r = 0; r <= reps; r++) {
        for (hipStream_t st : b->streams) hipStreamSynchronize(st);
        const auto t0 = std::chrono::steady_clock::now();
        const int rc = capow_batch_run(b, mocks, n, out.data(), nullptr, nullptr);
        const auto t1 = std::chrono::steady_clock::now();
        if (rc) return rc;
        if (r > 0) t.push_back(std::chrono::duration<double>(t1 - t0).count());
      }
      std::sort(t.begin(), t.end());
      *med = t[t.size() / 2];
      return CAPOW_OK;
    };
    // every stack context at the defaults first: a set is timed against
    // the others' defaults until their own turn
    for (capow_ctx* c : b->sctx) {
      c->count_blocks = 12 * (c->ncu > 0 ? c->ncu : 256);
      c->g.sort_probes = 32;
      c->dep_halo = true;
      c->dep_bank = true;
      c->fft_rows_mul = 0;
      c->fft_order = c->fp64 ? 0 : 5;
    }
    for (int q = 0; q < b->nconf; q++) {
      bool used = false;
      for (int i = 0; i < n && !used; i++) used = mocks[i].config == q;
      if (!used) continue;
      capow_ctx* c = b->sctx[q];
      const int ncu = c->ncu > 0 ? c->ncu : 256;
      const int order0 = c->fp64 ? 0 : 5;
      int sort_blocks = 12, sort_lds = 1, halo = 1, bank = 1, runs = 0, order = order0;
      std::vector<TuneKnob> knobs = {{"sort_blocks", &sort_blocks, {12, 8, 4}},
                                     {"sort_lds", &sort_lds, {1, 0}},
                                     {"dep_halo", &halo, {1, 0}},
                                     {"dep_bank", &bank, {1, 0}},
                                     {"fft_runs", &runs, {0, 1}},
                                     {"fft_order", &order, {order0, 0, 7, 2}}};
      auto apply = [&]() {
        c->count_blocks = sort_blocks * ncu;
        c->g.sort_probes = sort_lds ? 32 : 0;
        c->dep_halo = halo != 0;
        c->dep_bank = bank != 0;
        c->fft_rows_mul = runs;
        c->fft_order = order;
      };
      apply();
      for (TuneKnob& k : knobs) {
        int best_v = *k.value;
        double best_t = 1e30;
        for (size_t ci = 0; ci < k.cand.size(); ci++) {
          const int v = k.cand[ci];
          if (ci > 0 && v == k.cand[0]) continue;   // the current value, timed first
          *k.value = v;
          apply();
          double med = 0.0;
          const int rc = time_all(&med);
          if (rc) return rc;
          tune_note(rep, "  stack config %d %s=%d: %.3f ms\n", q, k.name, v, med * 1e3);
          if (med < best_t) {
            best_t = med;
            best_v = v;
          }
        }
        *k.value = best_v;
        apply();
      }
      tune_note(rep,
                "chosen: stack config %d sort_blocks=%d sort_lds=%d dep_halo=%d dep_bank=%d "
                "fft_runs=%d fft_order=%d\n",
                q, sort_blocks, sort_lds, halo, bank, runs, order);
    }
  }
  if (report && report_len > 0) {
    strncpy(report, rep.c_str(), (size_t)report_len - 1);
    report[report_len - 1] = 0;
  }
  return CAPOW_OK;
}
