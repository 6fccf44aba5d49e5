// kernels.h -- launch arguments shared by the API layer and the kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "common.h"

namespace capow {

// Global sums written by k_stats_final (doubles).
enum {
  STAT_NGRID = 0,   // sum w over main-accepted particles
  STAT_W = 1,       // sum w over all particles
  STAT_V1 = 2, STAT_V2 = 3, STAT_V3 = 4, STAT_V4 = 5,
  STAT_NOUT = 6,
  STAT_N = 8,       // stride of block partials
  STAT_ALPHA = 8,
  // 9..14 unused
  STAT_TOTAL = 16
};

// Kernel timer: HIP events around launches on the context stream.
struct Timer {
  bool on = false;
  hipStream_t stream = nullptr;
  struct Rec { std::string name; hipEvent_t a, b; };
  std::vector<Rec> pending;
  std::vector<hipEvent_t> pool;
  std::vector<std::string> names;
  std::vector<double> ms;
  std::vector<int> launches;
  hipEvent_t ev() {
    hipEvent_t e;
    if (!pool.empty()) { e = pool.back(); pool.pop_back(); }
    else hipEventCreate(&e);
    return e;
  }
  void push(const char* n, hipEvent_t a, hipEvent_t b) { pending.push_back({n, a, b}); }
  // after the stream is synchronised
  void collect() {
    for (auto& r : pending) {
      float t = 0.f;
      hipEventElapsedTime(&t, r.a, r.b);
      size_t k = 0;
      for (; k < names.size(); k++) if (names[k] == r.name) break;
      if (k == names.size()) { names.push_back(r.name); ms.push_back(0.0); launches.push_back(0); }
      ms[k] += t;
      launches[k] += 1;
      pool.push_back(r.a);
      pool.push_back(r.b);
    }
    pending.clear();
  }
  void reset() { names.clear(); ms.clear(); launches.clear(); }
  ~Timer() {
    for (auto& r : pending) { hipEventDestroy(r.a); hipEventDestroy(r.b); }
    for (auto e : pool) hipEventDestroy(e);
  }
};

#define TIMED(timer, name, launch)                                  \
  do {                                                              \
    Timer* _t = (timer);                                            \
    if (_t && _t->on) {                                             \
      hipEvent_t _a = _t->ev(), _b = _t->ev();                      \
      hipEventRecord(_a, _t->stream);                               \
      launch;                                                       \
      hipEventRecord(_b, _t->stream);                               \
      _t->push(name, _a, _b);                                       \
    } else {                                                        \
      launch;                                                       \
    }                                                               \
  } while (0)

// particles per bucket-kernel block chunk (one LDS aggregation round)
#define BK_CHUNK 2048

// One catalogue of a stacked batch (device table, one entry per catalogue):
// its particles (n x 4), weights (n or null) and count.
struct MockDev {
  const void* parts;
  const void* w;
  int64_t n;
};

struct DepositArgs {
  const Geom* g;
  hipStream_t stream;
  Timer* timer;
  const void* parts;      // n x 4 (x y z v_los)
  const void* wts;        // n or null
  int64_t n;
  int count_blocks;
  uint32_t* counts;       // nbuckets = ntiles * 27 (zeroed)
  uint32_t* offsets;      // nbuckets + 1
  uint32_t* cursor;       // nbuckets: scatter cursors
  uint32_t* bsum;         // scan block sums
  uint32_t* boff;         // scan block offsets (+1)
  void* sorted;           // sorted_cap x 4 (one row per particle copy)
  void* sorted_w;         // sorted_cap or null
  uint32_t* sorted_idx = nullptr;   // sorted_cap input indices of the rows, or null (not written)
  int64_t sorted_cap;     // rows the sorted buffers hold
  double* partials;       // count_blocks * STAT_N
  unsigned int* errs;     // 2 (zeroed): [0] LOS=0 out-of-box, [1] reach violations
  double* stats;          // STAT_TOTAL
  void* grid[4];          // A, A_int, B, B_int (null when unused)
  bool halo = true;       // halo-row deposit where the tile shape allows
  bool bank = true;       // halo-row deposit: bank-balanced copy order, masked rows
  // stacked batch: nmock catalogues of one geometry (g->mock_stride set),
  // their particles in the device table `mocks` (parts, wts, n unused);
  // counts, offsets and cursors hold nmock * nbuckets entries, partials
  // nmock * count_blocks rows, errs 4 and stats STAT_TOTAL per catalogue
  int nmock = 1;
  const MockDev* mocks = nullptr;
};
// count + stats + scan; then scatter + tile deposit
int run_count(DepositArgs& a, bool fp64);
int run_scatter_deposit(DepositArgs& a, bool fp64);

struct FftPlan {
  int npass;
  int radix[24];
};

struct FftArgs {
  const Geom* g;
  hipStream_t stream;
  Timer* timer;
  void* grid[4];          // compacted list of grids to transform
  int ngrids;
  FftPlan plan[3];        // x: NX, y: NY, z: NZ/2
  const void* tw[3];      // x: NX entries, y: NY, z: NZ (exp(-2 pi i j/N))
  bool fast = true;       // register two-level kernels where the size allows
  int col_mul = 1;        // column batch multiplier of the y/x passes (1 or 2)
  // out-of-place pipeline (null: in place): scratch grids of the same
  // capacity, row pitches of the intermediate layouts; swapped = the result
  // is in the scratch grids
  void* scratch[4] = {nullptr, nullptr, nullptr, nullptr};
  int64_t px = 0, py = 0;
  bool swapped = false;
  int rows_mul = 0;       // rows per workgroup multiplier of the out-of-place passes (0: per pass)
  int ncu = 256;          // compute units (persistent grids)
  int persist = 1;        // 0: one workgroup per batch instead of a persistent grid
  int order = 0;          // bit per pass (z, y, x): batches with the outer index fastest
  // FP32 out-of-place pipeline with the x pass deferred to the binning
  // (run_fft_x_bin): defer_x asks for the z and y passes only; x_deferred
  // reports that they ran (the y pass output stays in grid[]).  The y pass
  // then sums |u|^2 of its input per grid and workgroup into ssq (XBIN_SSQ_MAX
  // rows of 4 doubles) and k_xbin_scale turns them into the binning quantum
  // xq = {q, 1/q}: xbin_factor x NX NY sum_u |u|^2 bounds the sum of the
  // magnitudes of every binned contribution (Parseval).
  bool defer_x = false;
  bool x_deferred = false;
  double* ssq = nullptr;
  double xbin_factor = 0.0;
  double* xq = nullptr;
  int ssq_rows = 0;       // out: workgroups of the y pass (rows of ssq)
  int ssq_slot0 = 0;      // ssq slot of grid[0] (grid-by-grid passes)
  // one spare buffer (scratch[0] only, grids too large for a scratch each):
  // the passes run grid by grid, z: grid -> spare, y: spare -> grid, and an
  // x pass grid -> spare swaps the two; on return grid[] and scratch[0]
  // hold the rotated pointers (rotated = true), for the caller to keep
  bool one_spare = false;
  bool rotated = false;
  // stacked batch: nmock catalogues, each buffer's catalogue m at m *
  // mock_stride reals (ssq: nmock blocks of ssq_rows rows; xq: 6 doubles
  // per catalogue); transposed register pipeline only
  int nmock = 1;
  int64_t mock_stride = 0;
};
#define XBIN_SSQ_MAX 8192
int run_fft(FftArgs& a, bool fp64);
// zero fill by a kernel (deposit.hip)
int zero_u32(uint32_t* p, int64_t n, hipStream_t st);
// The pending x pass after run_fft with x_deferred (grid -> scratch; the
// caller swaps the buffers as after a whole run_fft).
int run_fft_x(FftArgs& a);

// adjoint of run_fft on gradient grids (register FFT sizes only)
int run_ifft(FftArgs& a, bool fp64);

enum { SPEC_AUTO_A = 0, SPEC_AUTO_B = 1, SPEC_CROSS_IM = 2, SPEC_CROSS_RE = 3 };

struct BinArgsDev {
  int nx, ny, nzh;
  FastDiv fd_nzh;
  int rowc;               // complex per row in device memory (>= nzh)
  FastDiv fd_rowc;
  int nk, nmu, out2d, odd, logbins, los, interlace, hasB;
  int nspec;
  int kind[3];
  int copies;
  double mink, binwidth, inv_binwidth, min_nyq;
  // offsets (doubles) of the per-axis tables in the table buffer:
  // frequencies f, deconvolution cor, interlace phases ph (re, im pairs)
  int64_t ofx, ofy, ofz, ocx, ocy, ocz, opx, opy, opz;
};

struct BinLaunch {
  hipStream_t stream;
  Timer* timer;
  const void* grid[4];    // A, A_int, B, B_int (unused -> any valid pointer)
  BinArgsDev dev;
  const double* tabs;     // per-axis tables (offsets in dev)
  bool row_form;          // row-walking kernel (z tables in LDS)
  bool mirror;            // mirror-group form (symmetric tables, even NX, NY)
  bool mirror2d;          // (k, mu) bins in the mirror form (line of sight z)
  long long* partials2;   // nwg * nspec * NK * NMU int64 (k, mu) tables
  double* qscale;         // 2 NK: per-k quanta and their inverses
  int nwg;
  int per_copy;
  double* partials;       // nwg * per_copy
  double* out;            // per_copy (stacked batch: per catalogue, out_stride apart)
  int64_t out_stride = 0;
  double* counts;         // nk: mode counts of the branch-free form
  bool* counts_valid;     // counts computed for this geometry
};
// The binning plan of a parameter set for grids of `rowc` complex values
// per row: the kernels' arguments, which form runs, the workgroup count of a
// device with `ncu` compute units and the buffer sizes (capow_create and
// capow_bin_noalloc share it); the per-axis tables it points into.
struct BinPlan {
  BinArgsDev b;
  bool mirror, row_form, mirror2d;
  bool sym;               // f(N - i) = -f(i), cor(N - i) = cor(i) on x and y, even NX, NY
  int codes[3];
  int per_copy;           // doubles per bin copy
  int nwg;                // workgroups of the accumulation kernels
  size_t ntab;            // doubles of the per-axis tables
};
int make_bin_plan(const capow_params& p, int64_t rowc, bool mirror_allowed, int ncu, BinPlan& bp,
                  std::string& err);
void fill_bin_tables(const capow_params& p, double* tab);
double bin_table_at(const capow_params& p, size_t idx);   // one entry of fill_bin_tables
size_t bin_table_len(const capow_params& p);
// capow_create's parameter check, and the message capow_last_error(NULL) returns
int validate_params(const capow_params* p, std::string& err);
void set_global_error(const std::string& err);
int run_bin(BinLaunch& L, bool fp64);
// The x pass fused with the mirror-group binning (FP32, 1-D bins): the y
// pass output in f.grid[] is transformed along x and binned without being
// written; int64 quanta of f.xq per workgroup into xpart, summed into L.out
// (sums, then the mode counts from L.counts, which must be valid).
// xbin_lds: the LDS the fused kernel needs (0: this geometry has no fused
// kernel).
size_t xbin_lds(const Geom& g, const BinArgsDev& b);
int run_fft_x_bin(FftArgs& f, BinLaunch& L, long long* xpart, int xpart_rows);
// The pieces of run_bin around the accumulation kernel: the once-per-context
// mode counts of the branch-free forms, and the final sum of the partials.
int run_bin_counts(BinLaunch& L);
int run_bin_final(BinLaunch& L, bool fast);

// Survey (Yamamoto) estimator passes (survey.hip).
struct SurveyLaunch {
  hipStream_t stream;
  Timer* timer;
  const Geom* g;
  const void* src[4];     // real-space fields A, A_int, B, B_int (null: unused)
  void* work[4];          // weighted copies, transformed in place by run_fft
  void* F0;               // monopole k-space fields kept for the cross terms
  void* F0m;
  BinArgsDev dev;         // geometry and k bins (2-D bins off)
  const double* tabs;
  double origin[3];
  int mom;                // momentum cross spectrum (B grids)
  bool mirror = false;    // mirror groups (symmetric x and y tables: BinPlan::sym)
  int nwg;
  double* partials;       // nwg * (up to 4) * NK: the sums one pass touches
  double* acc;            // 6 * NK: P0..P4, mode counts
};
// work[s] = src[s] * (product of r^ components idx[0..n-1]; idx[0] twice
// for n = 4) on the real cells
int run_survey_weight(SurveyLaunch& L, bool fp64, int n, const int idx[3]);
// one multipole pass over the transformed work grids, added into L.acc
int run_survey_bin(SurveyLaunch& L, bool fp64, int multipole, const int idx[3]);

// Backward of the binning: dL/draw (G: [n_spectra][5][NK]) -> gradients of
// the k-space grids, in place; g0: the k = 0 gradients (8 doubles).
int run_bin_backward(const BinLaunch& L, const double* G, double* g0, bool fp64);

struct GatherArgs {
  const Geom* g;
  hipStream_t stream;
  Timer* timer;
  const void* parts;      // n x 4
  const void* wts;        // n or null
  int64_t n;
  // the forward's tile sort (rows in tile order with their input indices):
  // when set, the tile gather runs; otherwise the per-particle gather
  const void* sorted = nullptr;
  const void* sorted_w = nullptr;
  const uint32_t* sorted_idx = nullptr;
  const uint32_t* offsets = nullptr;
  unsigned int* errs = nullptr;   // [1]: a tap outside its tile's region
  const void* grid[4];    // real-space gradient grids (A, A_int, B, B_int)
  const double* g0;       // k = 0 gradients from the binning backward
  void* gout;             // n x 4: dL/dx, dL/dy, dL/dz, dL/dv_los
  void* gw;               // n: dL/dw (or null)
};
int run_gather(const GatherArgs& a, bool fp64);
size_t gather_tiles_lds(const Geom& g, bool fp64);

}  // namespace capow
