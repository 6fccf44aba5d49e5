// noalloc.hip -- capow_bin_noalloc: the binning of capow_bin on
// caller-owned device grids, for callers that cannot allocate device memory.
//
// Every buffer the binning kernels use -- per-axis tables, per-workgroup
// partial bins, the reduced sums, mode counts, the (k, mu) tables and their
// quanta, and the unpacked result -- is carved out of one file-scope device
// array, `scratch`, of CAPOW_NOALLOC_BYTES.  Each GPU the library runs on
// holds its own copy until the process ends.  Host code never uses the
// array's name as an address: the first call on a device looks its address
// up with hipGetSymbolAddress (with that device current) and keeps it per
// device.  No hipMalloc, hipFree or allocating library routine is called.
//
// Every call on one GPU shares that GPU's array, which keeps the last
// call's contents: calls that overlap on one GPU (two host threads, or two
// streams) race.  What must start at zero (the mode counts) is cleared by a
// kernel in each call.
#include <mutex>
#include <string>

#include "capow.h"
#include "common.h"
#include "kernels.h"

namespace capow {

#define CAPOW_NOALLOC_BYTES (64ull << 20)

static __device__ __align__(256) unsigned char scratch[CAPOW_NOALLOC_BYTES];

static std::mutex g_scratch_mu;
static void* g_scratch_ptr[64];   // per device: the address of `scratch`

// The address of this device's `scratch`, looked up once per device with
// the device current (the caller's current device).
static int scratch_address(void** out, std::string& m) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) {
    m = "capow_bin_noalloc: no current HIP device (or more than 64 devices)";
    return CAPOW_ERR_HIP;
  }
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
      m = std::string("capow_bin_noalloc: hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
      return CAPOW_ERR_HIP;
    }
    g_scratch_ptr[dev] = ptr;
  }
  *out = g_scratch_ptr[dev];
  return CAPOW_OK;
}

// capow_bins layout from the kernels' bin order (per spectrum l = 0, 2, 4
// [, 1, 3], then the (k, mu) sums, the counts and the (k, mu) counts).
__global__ void k_bin_unpack(const double* __restrict__ h, int ns, int nl, int nk, int n2d,
                             double* __restrict__ pk, long long* __restrict__ nmodes,
                             double* __restrict__ pk2d, long long* __restrict__ nmodes2d) {
  const int npk = ns * 5 * nk;
  const double* acc2 = h + (size_t)ns * nl * nk;
  const double* cnt = acc2 + (size_t)ns * n2d;
  const double* cnt2 = cnt + nk;
  const int total = npk + nk + ns * n2d + n2d;
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < total; t += gridDim.x * blockDim.x) {
    if (t < npk) {
      const int s = t / (5 * nk), l = (t / nk) % 5, k = t % nk;
      // position of multipole l in the kernels' order 0, 2, 4, 1, 3
      const int q = l == 0 ? 0 : (l == 2 ? 1 : (l == 4 ? 2 : (l == 1 ? 3 : 4)));
      pk[t] = q < nl ? h[(size_t)(s * nl + q) * nk + k] : 0.0;
    } else if (t < npk + nk) {
      nmodes[t - npk] = (long long)cnt[t - npk];
    } else if (t < npk + nk + ns * n2d) {
      pk2d[t - npk - nk] = acc2[t - npk - nk];
    } else {
      const int q = t - npk - nk - ns * n2d;
      nmodes2d[q] = (long long)cnt2[q];
    }
  }
}

static size_t up256(size_t b) { return (b + 255) / 256 * 256; }

static int noalloc_fail(int code, const std::string& m) {
  set_global_error(m);
  return code;
}

}  // namespace capow

using namespace capow;

extern "C" int capow_bin_noalloc(const capow_params* p, const void* const grids[4], int64_t row_pitch,
                                 void* hip_stream, capow_bins* out) {
  if (!p || !grids || !out) return noalloc_fail(CAPOW_ERR_ARG, "capow_bin_noalloc: null argument");
  std::string m;
  int rc = validate_params(p, m);
  if (rc) return noalloc_fail(rc, m);
  if (p->edge_mode == CAPOW_EDGE_SURVEY)
    return noalloc_fail(CAPOW_ERR_UNSUPPORTED, "capow_bin_noalloc: survey parameters (use capow_survey_power)");
  const bool fp64 = p->precision == CAPOW_FP64;
  const int64_t nzh = p->nz / 2 + 1;
  // rows of whole 128-byte lines, as capow_create lays them out: the 16-byte
  // loads of the binning kernels never leave a row
  const int64_t align = fp64 ? 8 : 16;
  if (row_pitch < nzh || row_pitch % align != 0)
    return noalloc_fail(CAPOW_ERR_ARG, "capow_bin_noalloc: row_pitch must be >= NZ/2+1 complex values "
                                       "and a whole number of 128-byte lines");
  // exponent b of field B = w v^b per Momentum (src/read_data.c:91-130; -1:
  // no ddg_mom grid)
  static const int kFieldB[9] = {-1, 1, -1, 2, 2, -1, 3, 3, 4};
  const int fb = kFieldB[p->momentum];
  const bool need[4] = {true, p->do_interlacing != 0, fb >= 0, fb >= 0 && p->do_interlacing != 0};
  for (int s = 0; s < 4; s++) {
    if (!need[s]) continue;
    if (!grids[s] || ((uintptr_t)grids[s] & 255) != 0)
      return noalloc_fail(CAPOW_ERR_ARG, "capow_bin_noalloc: grid " + std::to_string(s) +
                                             " missing or not 256-byte aligned");
  }
  int dev = 0, ncu = 256;
  if (hipGetDevice(&dev) == hipSuccess) hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
  if (ncu < 1) ncu = 256;
  BinPlan bp;
  rc = make_bin_plan(*p, row_pitch, true, ncu, bp, m);
  if (rc) return noalloc_fail(rc, m);
  const int ns = bp.b.nspec, nk = p->nk, nl = p->odd_multipoles ? 5 : 3;
  const int n2d = p->output_2d ? p->nk * p->nmu : 0;
  // carve the regions (256-byte aligned)
  size_t off = 0;
  const size_t o_tab = off; off += up256(bp.ntab * 8);
  const size_t o_part = off; off += up256((size_t)bp.nwg * bp.per_copy * 8);
  const size_t o_out = off; off += up256((size_t)bp.per_copy * 8);
  const size_t o_cnt = off; off += up256((size_t)(nk + n2d) * 8);
  const size_t o_p2 = off; off += bp.mirror2d ? up256((size_t)bp.nwg * ns * n2d * 8) : 0;
  const size_t o_qs = off; off += bp.mirror2d ? up256((size_t)2 * nk * 8) : 0;
  const size_t o_pk = off; off += up256((size_t)ns * 5 * nk * 8);
  const size_t o_nm = off; off += up256((size_t)nk * 8);
  const size_t o_pk2 = off; off += up256((size_t)ns * n2d * 8);
  const size_t o_nm2 = off; off += up256((size_t)n2d * 8);
  if (off > CAPOW_NOALLOC_BYTES)
    return noalloc_fail(CAPOW_ERR_UNSUPPORTED,
                        "capow_bin_noalloc: these bins need " + std::to_string(off) +
                            " bytes of scratch, more than its " +
                            std::to_string((unsigned long long)CAPOW_NOALLOC_BYTES) + " bytes");
  unsigned char* base = nullptr;
  rc = scratch_address((void**)&base, m);
  if (rc) return noalloc_fail(rc, m);
  hipStream_t st = (hipStream_t)hip_stream;
  // tables: computed on the host in stack-sized chunks (the same entries
  // capow_create uploads), each copy finished before its chunk is reused
  {
    double chunk[2048];
    const size_t n = bp.ntab;
    for (size_t i0 = 0; i0 < n; i0 += 2048) {
      const size_t len = n - i0 < 2048 ? n - i0 : 2048;
      for (size_t i = 0; i < len; i++) chunk[i] = bin_table_at(*p, i0 + i);
      if (hipMemcpyAsync(base + o_tab + i0 * 8, chunk, len * 8, hipMemcpyHostToDevice, st) != hipSuccess ||
          hipStreamSynchronize(st) != hipSuccess)
        return noalloc_fail(CAPOW_ERR_HIP, "capow_bin_noalloc: table upload failed");
    }
  }
  Timer timer;
  bool counts_valid = false;   // mode counts are recomputed every call
  BinLaunch L;
  L.stream = st;
  L.timer = &timer;
  for (int s = 0; s < 4; s++) L.grid[s] = need[s] ? grids[s] : grids[0];
  L.dev = bp.b;
  L.tabs = (const double*)(base + o_tab);
  L.row_form = bp.row_form;
  L.mirror = bp.mirror;
  L.mirror2d = bp.mirror2d;
  L.partials2 = bp.mirror2d ? (long long*)(base + o_p2) : nullptr;
  L.qscale = bp.mirror2d ? (double*)(base + o_qs) : nullptr;
  L.nwg = bp.nwg;
  L.per_copy = bp.per_copy;
  L.partials = (double*)(base + o_part);
  L.out = (double*)(base + o_out);
  L.counts = (double*)(base + o_cnt);
  L.counts_valid = &counts_valid;
  rc = run_bin(L, fp64);
  if (rc) return noalloc_fail(rc, std::string("capow_bin_noalloc: binning launch failed: ") +
                                      hipGetErrorString(hipGetLastError()));
  double* d_pk = (double*)(base + o_pk);
  long long* d_nm = (long long*)(base + o_nm);
  double* d_pk2 = (double*)(base + o_pk2);
  long long* d_nm2 = (long long*)(base + o_nm2);
  k_bin_unpack<<<64, 256, 0, st>>>(L.out, ns, nl, nk, n2d, d_pk, d_nm, d_pk2, d_nm2);
  if (hipGetLastError() != hipSuccess)
    return noalloc_fail(CAPOW_ERR_HIP, "capow_bin_noalloc: unpack launch failed");
  out->n_spectra = ns;
  for (int s = 0; s < 3; s++) out->codes[s] = s < ns ? bp.codes[s] : -1;
  hipError_t e = hipSuccess;
  if (out->pk && e == hipSuccess)
    e = hipMemcpyAsync(out->pk, d_pk, (size_t)ns * 5 * nk * 8, hipMemcpyDeviceToHost, st);
  if (out->nmodes && e == hipSuccess)
    e = hipMemcpyAsync(out->nmodes, d_nm, (size_t)nk * 8, hipMemcpyDeviceToHost, st);
  if (n2d && out->pk2d && e == hipSuccess)
    e = hipMemcpyAsync(out->pk2d, d_pk2, (size_t)ns * n2d * 8, hipMemcpyDeviceToHost, st);
  if (n2d && out->nmodes2d && e == hipSuccess)
    e = hipMemcpyAsync(out->nmodes2d, d_nm2, (size_t)n2d * 8, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess)
    return noalloc_fail(CAPOW_ERR_HIP, std::string("capow_bin_noalloc: ") + hipGetErrorString(e));
  return CAPOW_OK;
}
