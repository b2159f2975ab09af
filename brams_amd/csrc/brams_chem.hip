// C-ABI of the CCATT RODAS3 chemistry step on MI355X (include/brams_chem.h).
//
// Host side of the drop-in for chem_rodas3_dyndt
// (src/brams/ccatt/chem_spack_rodas3_dyndt.f90:41-131): validates the
// arguments, keeps persistent device buffers per grid, moves host arrays
// to HBM and back for the Fortran entry point, and launches the kernels of
// rodas3_kernel.h:
//   1. block 1 alone (its value-based sparsity mask changes every attempt and
//      fixes the mask for every later block, rodas3:470-506);
//   2. the remaining blocks, 128-lane workgroups for the leading blocks with
//      more than 64 cells and wave64 workgroups for the rest.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include <algorithm>

#include "../../include/brams_chem.h"
#include "../../include/brams_surrogate.h"
#include "rodas3_kernel.h"
#include "rodas3_launch.h"
#include "rodas3_slab_launch.h"
#include "sched_kernels.h"
#include "jphoto_kernels.h"
#include "sur_step_kernels.h"

namespace {

thread_local std::string g_err = "";

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define HIP_TRY(expr)                                                              \
  do {                                                                             \
    hipError_t e_ = (expr);                                                        \
    if (e_ != hipSuccess)                                                          \
      return fail(-100 - (int)e_, "%s failed: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

// device buffer that only grows
struct DevBuf {
  void* p = nullptr;
  size_t n = 0;
  hipError_t ensure(size_t bytes) {
    if (bytes <= n) return hipSuccess;
    if (p) hipFree(p);
    p = nullptr;
    n = 0;
    hipError_t e = hipMalloc(&p, bytes);
    if (e == hipSuccess) n = bytes;
    return e;
  }
  void release() {
    if (p) hipFree(p);
    p = nullptr;
    n = 0;
  }
  template <class T> T* as() const { return static_cast<T*>(p); }
};

// Kernel forms of the 64-cell partition blocks: the slab layouts
// 0..kSlabVariants-1 (rodas3_slab.h) and the per-lane kernel
// (rodas3_kernel.h) at 1, 2 and 4 waves per SIMD.  All give bit-identical
// results; the autotuner keeps the fastest per grid.
constexpr int kFormPerLane = brams::kSlabVariants;      // first per-lane form
constexpr int kPerLaneWps[3] = {1, 2, 4};
constexpr int kForms = brams::kSlabVariants + 3;
// the per-lane forms the library holds (a lean build: 1 and 4 waves per SIMD)
constexpr bool form_built(int f) {
  return f >= kFormPerLane && f < kForms && (!brams::kLeanBuild || f != kFormPerLane + 1);
}
// List (fallback) launches: the per-lane kernel at 1 or 2 waves per SIMD,
// by the launch's block count.  Measured on the 1.6M-cell rollout (ms per
// launch, 1 / 2 waves per SIMD): 19 400 blocks 246 / 190; the 2000..6000
// blocks of the warm steps 35.5 / 39.6 on average over 24 steps.  Timing
// each call's choice online was too noisy to pick reliably (consecutive
// calls differ by 10%), so the crossover is fixed.
constexpr int kListWps[2] = {1, 2};
constexpr int kListManyBlocks = 8192;
// The slab forms are switched off: on gfx950 their kernels did not
// terminate, even for a one-block grid with the attempts capped at 300,
// while a host build of the same block integration matches the oracle bit
// for bit.  Until that is understood every call takes the per-lane form.
constexpr bool kSlabEnabled = false;

// Device state of one grid, created on the first call for that grid and
// kept until brams_chem_finalize (the reference allocates its block index
// per grid on the first chemistry call, chemistry.f90:310-320).  Grids are
// told apart by their extents and interior bounds, so nested grids called in
// turn keep their own partition, block history and fallback memory.
struct GridState {
  int m1 = 0, m2 = 0, m3 = 0, ia = 0, iz = 0, ja = 0, jz = 0;
  long long ncell = 0;
  std::vector<int> sizes;        // AllocIndex block sizes
  DevBuf block_start, mask, attempts, substeps, failed, forced, thresh, order_thresh;
  bool blk_hist = false;         // attempts[] holds a previous call on this grid
  long long hist_cells = -1;     // fallback memory of the surrogate step (-1: to be started)
  DevBuf cell_dt;
  std::vector<int> h_att, h_sub, h_fail, h_thresh;
  std::vector<float> h_forced;
  int form = -1;                 // kernel form of the 64-cell blocks (-1: not tuned yet)
  float tune_ms[kForms] = {};
  int ros2_form = -1;            // the same for the ROS2 step (forms 3 and 5)
  float ros2_tune_ms[kForms] = {};
  bool same(int a1, int a2, int a3, int b1, int b2, int b3, int b4) const {
    return m1 == a1 && m2 == a2 && m3 == a3 && ia == b1 && iz == b2 && ja == b3 && jz == b4;
  }
};

struct State {
  bool init = false;
  int device = 0;
  std::vector<GridState*> grids;
  GridState* cur = nullptr;      // grid of the last call
  // host-entry staging
  DevBuf sc_p, sc_t, sc_t_dyn, met, jphoto, last_dt, consts;
  // surrogate step
  DevBuf sur_X, sur_out, sur_flag, sur_list, sur_chunk, sur_reason, fb_dt, fb_att, mask_all;
  // work ordering (sched_kernels.h): sort scratch and block order
  DevBuf sort_keys, sort_ids, sort_count, sort_off, order;
  // slab kernels: per-slot scratch regions and the work-queue counter
  DevBuf slab_lu, slab_jac, slab_rk, slab_vec, qcount;
  long long slab_slots = 0;
  int n_cu = 0;
  int force_form = -1;           // brams_chem_set_form (-1: autotune per grid)
  // autotuning scratch (copies of the state a step writes)
  DevBuf tune_sc, tune_dt;
  // QSSA: count of cells that reached 1e20
  DevBuf qssa_fatal;
  // several grids in one pass: per-grid kernel parameters and work lists
  DevBuf multi_ps, multi_work;
  hipEvent_t ev0 = nullptr, ev1 = nullptr, evb = nullptr, evm = nullptr;
  hipEvent_t sev[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  long long main_cells = 0;
  int last_form = -1;
  brams_chem_stats stats{};
};
State g;

// the host entries' per-species constants in g.consts: weight f32,
// transport flag i32, atol f64, rtol f64, each region 256-byte aligned
constexpr size_t kConstRound(size_t x) { return (x + 255) / 256 * 256; }
constexpr size_t kConstOffT = kConstRound(sizeof(float) * MECH_NS);
constexpr size_t kConstOffA = kConstOffT + kConstRound(sizeof(int) * MECH_NS);
constexpr size_t kConstOffR = kConstOffA + kConstRound(sizeof(double) * MECH_NS);
constexpr size_t kConstBytes = kConstOffR + kConstRound(sizeof(double) * MECH_NS);

// Bookkeeping of brams_chem_rodas3_dyndt_device_noalloc (block starts, mask
// words, per-block counters, failure count) lives in one file-scope device
// array instead of allocated buffers.  Every device holds its own copy for
// the life of the process; on the host, `scratch` names a host-side stand-in,
// so only the device address taken per device by scratch_address() is ever
// passed to kernels or copies.  Every region is fully written by each call
// before it is read (nothing has to start at zero), and calls that overlap on
// one device race on it.
constexpr int kNoAllocMaxBlocks = 32768;   // 2 097 152 cells at 64 cells per block
constexpr size_t na_round(size_t x) { return (x + 255) / 256 * 256; }
constexpr size_t kNaStart = 0;                                                           // int[max + 1]
constexpr size_t kNaMask = na_round(kNaStart + sizeof(int) * (kNoAllocMaxBlocks + 1));    // unsigned[words]
constexpr size_t kNaAttempts = na_round(kNaMask + sizeof(unsigned) * MECH_MASK_WORDS);     // int[max]
constexpr size_t kNaSubsteps = na_round(kNaAttempts + sizeof(int) * kNoAllocMaxBlocks);    // int[max]
constexpr size_t kNaFailed = na_round(kNaSubsteps + sizeof(int) * kNoAllocMaxBlocks);      // int[max]
constexpr size_t kNaForced = na_round(kNaFailed + sizeof(int) * kNoAllocMaxBlocks);        // float[max]
constexpr size_t kNaThresh = na_round(kNaForced + sizeof(float) * kNoAllocMaxBlocks);      // int[max]
constexpr size_t kNaOrderThresh = na_round(kNaThresh + sizeof(int) * kNoAllocMaxBlocks);   // int[1]
constexpr size_t kNaCount = na_round(kNaOrderThresh + sizeof(int));                        // int[1]
constexpr size_t kNoAllocBytes = na_round(kNaCount + sizeof(int));

static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

constexpr int kMaxDevices = 64;
unsigned char* g_scratch[kMaxDevices];   // device address of scratch per device (null: not taken)

// the device address of scratch on the current device, taken on the first
// call there (hipGetSymbolAddress with that device current)
int scratch_address(unsigned char** out) {
  int dev = -1;
  HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices)
    return fail(-2, "device %d is beyond the %d devices the no-allocation entry serves", dev, kMaxDevices);
  if (!g_scratch[dev]) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      return fail(-100 - (int)e, "hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
    g_scratch[dev] = static_cast<unsigned char*>(ptr);
  }
  *out = g_scratch[dev];
  return 0;
}

// AllocIndex, chem_spack_utils.f90:93-125
int partition(int block_size, long long ncell, std::vector<int>& sizes) {
  if (ncell <= 0 || block_size <= 0) return -1;
  const int nob = (block_size >= ncell) ? 1 : (int)(ncell / block_size);
  sizes.assign(nob, (block_size >= ncell) ? (int)ncell : block_size);
  long long rest = ncell - (long long)sizes[0] * nob;
  while (rest > 0)
    for (int b = 0; b < nob && rest > 0; ++b) {
      ++sizes[b];
      --rest;
    }
  return nob;
}

// the state of the grid with these extents, created (partition, device
// bookkeeping) on its first call
int ensure_grid(int m1, int m2, int m3, int ia, int iz, int ja, int jz) {
  for (GridState* gs : g.grids)
    if (gs->same(m1, m2, m3, ia, iz, ja, jz)) {
      g.cur = gs;
      return 0;
    }
  const long long ncell = (long long)(m1 - 2) * (iz - ia + 1) * (jz - ja + 1);
  std::vector<int> sizes;
  int nob = partition(64, ncell, sizes);
  if (nob < 0) return fail(-2, "empty grid");
  for (int sz : sizes)
    if (sz > 128) return fail(-3, "block of %d cells exceeds 128", sz);
  GridState* gs = new GridState;
  gs->m1 = m1; gs->m2 = m2; gs->m3 = m3; gs->ia = ia; gs->iz = iz; gs->ja = ja; gs->jz = jz;
  gs->ncell = ncell;
  gs->sizes = sizes;
  std::vector<int> start(nob + 1, 0);
  for (int b = 0; b < nob; ++b) start[b + 1] = start[b] + sizes[b];
  hipError_t e = gs->block_start.ensure(sizeof(int) * (nob + 1));
  if (e == hipSuccess) e = hipMemcpy(gs->block_start.p, start.data(), sizeof(int) * (nob + 1), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = gs->mask.ensure(sizeof(unsigned) * MECH_MASK_WORDS);
  if (e == hipSuccess) e = gs->attempts.ensure(sizeof(int) * nob);
  if (e == hipSuccess) e = gs->substeps.ensure(sizeof(int) * nob);
  if (e == hipSuccess) e = gs->failed.ensure(sizeof(int) * nob);
  if (e == hipSuccess) e = gs->forced.ensure(sizeof(float) * nob);
  if (e == hipSuccess) e = gs->thresh.ensure(sizeof(int) * nob);
  if (e == hipSuccess) e = gs->order_thresh.ensure(sizeof(int));
  if (e != hipSuccess) {
    DevBuf* bufs[] = {&gs->block_start, &gs->mask, &gs->attempts, &gs->substeps, &gs->failed,
                      &gs->forced, &gs->thresh, &gs->order_thresh};
    for (DevBuf* b : bufs) b->release();
    delete gs;
    return fail(-100 - (int)e, "grid state allocation failed: %s", hipGetErrorString(e));
  }
  g.grids.push_back(gs);
  g.cur = gs;
  return 0;
}

// stable descending order of items[0..n) by keys (0..kSortBins-1) into out
int sort_desc(const int* keys, const int* items, int n, int* out, hipStream_t st) {
  const int nchunk = (n + brams::kSortChunk - 1) / brams::kSortChunk;
  const int nslot = brams::kSortBins * nchunk;
  HIP_TRY(g.sort_count.ensure(sizeof(int) * (nslot + 1)));
  HIP_TRY(g.sort_off.ensure(sizeof(int) * (nslot + 1)));
  const unsigned nb = (unsigned)((nchunk + 63) / 64);
  hipLaunchKernelGGL(brams::sort_count, dim3(nb), dim3(64), 0, st, keys, n, nchunk, g.sort_count.as<int>());
  hipLaunchKernelGGL(brams::compact_scan, dim3(1), dim3(256), 0, st, g.sort_count.as<int>(), nslot,
                     g.sort_off.as<int>());
  hipLaunchKernelGGL(brams::sort_scatter, dim3(nb), dim3(64), 0, st, keys, items, n, nchunk,
                     g.sort_off.as<int>(), out);
  HIP_TRY(hipGetLastError());
  return 0;
}

// blocks b0..b0+n-1, longest first by the attempts of the previous call
int order_blocks(int b0, int n, hipStream_t st) {
  HIP_TRY(g.sort_keys.ensure(sizeof(int) * n));
  HIP_TRY(g.sort_ids.ensure(sizeof(int) * n));
  HIP_TRY(g.order.ensure(sizeof(int) * n));
  hipLaunchKernelGGL(brams::block_keys, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st,
                     g.cur->attempts.as<int>(), b0, n, g.sort_keys.as<int>(), g.sort_ids.as<int>());
  HIP_TRY(hipGetLastError());
  return sort_desc(g.sort_keys.as<int>(), g.sort_ids.as<int>(), n, g.order.as<int>(), st);
}

// slots of a slab variant: every CU times the workgroups resident per CU
int slab_slots(int variant, long long* nslots) {
  const int per_cu = brams::slab_resident_per_cu(variant);
  if (per_cu <= 0) return fail(-5, "slab kernel variant %d cannot be resident", variant);
  *nslots = (long long)g.n_cu * per_cu;
  return 0;
}

// per-slot scratch for n slots
int ensure_slabs(long long n) {
  if (n <= g.slab_slots) return 0;
  HIP_TRY(g.slab_lu.ensure(sizeof(double) * 64 * brams::kSlabLuRows * n));
  HIP_TRY(g.slab_jac.ensure(sizeof(double) * 64 * brams::kSlabJacRows * n));
  HIP_TRY(g.slab_rk.ensure(sizeof(double) * 64 * brams::kSlabRkRows * n));
  HIP_TRY(g.slab_vec.ensure(sizeof(double) * 64 * brams::kSlabVecRows * n));
  g.slab_slots = n;
  return 0;
}

brams::QueueParams queue(const int* order, int b0, int nitems) {
  brams::QueueParams q{};
  q.counter = g.qcount.as<int>();
  q.order = order;
  q.b0 = b0;
  q.nitems = nitems;
  q.lu = g.slab_lu.as<double>();
  q.jac = g.slab_jac.as<double>();
  q.rk = g.slab_rk.as<double>();
  q.vec = g.slab_vec.as<double>();
  return q;
}

// the 64-cell blocks [first64, nob) in the given kernel form; order: the
// blocks' launch order (null: ascending)
int launch_main(int form, const brams::StepParams& p, int first64, int nob, const int* order,
                hipStream_t st) {
  const int n = nob - first64;
  if (n <= 0) return 0;
  if (form >= kFormPerLane) {
    brams::StepParams pm = p;
    pm.block_order = order;
    HIP_TRY(brams::launch_rodas3(64, false, (unsigned)n, pm, first64, st, kPerLaneWps[form - kFormPerLane]));
    return 0;
  }
  long long nslots = 0;
  int rc = slab_slots(form, &nslots);
  if (rc) return rc;
  nslots = std::min<long long>(nslots, n);
  if ((rc = ensure_slabs(nslots))) return rc;
  HIP_TRY(brams::launch_rodas3_slab(form, (unsigned)nslots, p, queue(order, first64, n), st));
  return 0;
}

// block 1 alone: its value-based mask fixes the mask of every later block
int launch_block1(const brams::StepParams& p, hipStream_t st) {
  if (!kSlabEnabled || g.cur->sizes[0] > 64) {
    HIP_TRY(brams::launch_rodas3(g.cur->sizes[0] > 64 ? 128 : 64, true, 1, p, 0, st));
    return 0;
  }
  if (g.cur->sizes[0] > 64) {
    HIP_TRY(brams::launch_rodas3(128, true, 1, p, 0, st));
    return 0;
  }
  int rc = ensure_slabs(1);
  if (rc) return rc;
  HIP_TRY(brams::launch_rodas3_slab(brams::kSlabBlock1, 1, p, queue(nullptr, 0, 1), st));
  return 0;
}

// Autotuner: on the first call of a grid, each kernel form (the per-lane
// kernel at 1, 2 and 4 waves per SIMD, and the slab layouts when they are
// switched on) integrates the call's 64-cell blocks on a scratch copy of the
// state (block 1 has run on it first, so the mask is the call's), timed
// with events; the fastest is kept for the grid.  The forms give
// bit-identical results, so the choice changes speed only.  The caller's
// arrays are not touched.
int autotune(brams::StepParams p, int first64, int nob, hipStream_t st) {
  GridState& gs = *g.cur;
  const size_t nsc = sizeof(float) * MECH_NS * (size_t)p.n3d;
  const bool to_t = (p.split == brams::kSplitParallel) && (p.ndyn == 1);
  HIP_TRY(g.tune_sc.ensure(2 * nsc * (to_t ? 2 : 1)));
  HIP_TRY(g.tune_dt.ensure(2 * sizeof(double) * nob));
  float* sc0 = g.tune_sc.as<float>();
  float* sc1 = sc0 + (size_t)MECH_NS * p.n3d;
  float* st0 = to_t ? sc1 + (size_t)MECH_NS * p.n3d : nullptr;
  float* st1 = to_t ? st0 + (size_t)MECH_NS * p.n3d : nullptr;
  double* dt0 = g.tune_dt.as<double>();
  double* dt1 = dt0 + nob;
  HIP_TRY(hipMemcpyAsync(sc0, p.sc_p, nsc, hipMemcpyDeviceToDevice, st));
  if (to_t) HIP_TRY(hipMemcpyAsync(st0, p.sc_t, nsc, hipMemcpyDeviceToDevice, st));
  HIP_TRY(hipMemcpyAsync(dt0, p.last_dt, sizeof(double) * nob, hipMemcpyDeviceToDevice, st));
  brams::StepParams q = p;
  q.sc_p = sc0;
  q.sc_t = st0;
  q.last_dt = dt0;
  int rc = launch_block1(q, st);
  if (rc) return rc;
  q.sc_p = sc1;
  q.sc_t = st1;
  q.last_dt = dt1;
  const int f0 = kSlabEnabled ? 0 : kFormPerLane;
  int best = f0;
  for (int f = f0; f < kForms; ++f) {
    if (f >= kFormPerLane && !form_built(f)) continue;
    HIP_TRY(hipMemcpyAsync(sc1, sc0, nsc, hipMemcpyDeviceToDevice, st));
    if (to_t) HIP_TRY(hipMemcpyAsync(st1, st0, nsc, hipMemcpyDeviceToDevice, st));
    HIP_TRY(hipMemcpyAsync(dt1, dt0, sizeof(double) * nob, hipMemcpyDeviceToDevice, st));
    HIP_TRY(hipEventRecord(g.ev0, st));
    if ((rc = launch_main(f, q, first64, nob, nullptr, st))) return rc;
    HIP_TRY(hipEventRecord(g.ev1, st));
    HIP_TRY(hipEventSynchronize(g.ev1));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, g.ev0, g.ev1));
    gs.tune_ms[f] = ms;
    if (ms < gs.tune_ms[best]) best = f;
  }
  gs.form = best;
  return 0;
}

// Launch the blocks of the current grid; p has every pointer filled except
// the per-call bookkeeping arrays, which this sets.
int launch_step(brams::StepParams p, hipStream_t st) {
  GridState& gs = *g.cur;
  const int nob = (int)gs.sizes.size();
  p.block_start = gs.block_start.as<int>();
  p.mask = gs.mask.as<unsigned>();
  p.blk_attempts = gs.attempts.as<int>();
  p.blk_substeps = gs.substeps.as<int>();
  p.blk_failed = gs.failed.as<int>();
  p.blk_forced = gs.forced.as<float>();
  p.blk_thresh = gs.thresh.as<int>();
  p.order_thresh = gs.order_thresh.as<int>();
  const std::vector<int>& sz = gs.sizes;
  int nbig = 0;
  while (nbig < nob && sz[nbig] > 64) ++nbig;
  const int first64 = nbig > 1 ? nbig : 1;
  int form = g.force_form >= 0 ? g.force_form : gs.form;
  if (form < 0 && nob > first64) {
    int rc = autotune(p, first64, nob, st);
    if (rc) return rc;
    form = gs.form;
  }
  // the 64-cell blocks run longest first once a call on this grid has
  // counted their attempts (the order changes no block's result)
  const int* order = nullptr;
  if (nob > first64 && gs.blk_hist) {
    int rc = order_blocks(first64, nob - first64, st);
    if (rc) return rc;
    order = g.order.as<int>();
  }
  HIP_TRY(hipEventRecord(g.ev0, st));
  int rc = launch_block1(p, st);
  if (rc) return rc;
  HIP_TRY(hipEventRecord(g.evb, st));
  if (nbig > 1) HIP_TRY(brams::launch_rodas3(128, false, (unsigned)(nbig - 1), p, 1, st));
  HIP_TRY(hipEventRecord(g.evm, st));
  g.main_cells = 0;
  if (nob > first64) {
    if ((rc = launch_main(form, p, first64, nob, order, st))) return rc;
    for (int b = first64; b < nob; ++b) g.main_cells += sz[b];
  }
  HIP_TRY(hipEventRecord(g.ev1, st));
  gs.blk_hist = true;
  g.last_form = nob > first64 ? form : -1;
  return 0;
}

// The mask a ROS2 step reports: every structural entry, bits 0..MECH_NSTRUCT-1
// (ROS2 factors the full pattern, chem_spack_ros_dyndt.f90 has no value mask).
struct StructMask {
  unsigned w[MECH_MASK_WORDS];
};
constexpr StructMask struct_mask() {
  StructMask m{};
  for (int w = 0; w < MECH_MASK_WORDS; ++w) {
    const int bits = MECH_NSTRUCT - 32 * w;
    m.w[w] = bits >= 32 ? 0xffffffffu : bits > 0 ? (1u << bits) - 1u : 0u;
  }
  return m;
}
constexpr StructMask kStructMask = struct_mask();

// The ROS2 solver (CHEMISTRY=3, chem_spack_ros_dyndt.f90) on the current
// grid: every block alike on the full structural Jacobian, so no block-1 form
// and no mask; the 65..127-cell blocks on 128 lanes, the 64-cell blocks
// longest first from the second call on the grid.
// The ROS2 step's 64-cell blocks at one and four waves per SIMD (forms 3
// and 5, both built in every library), timed on a scratch copy of the state
// on the first ROS2 call of a grid, as autotune() does for RODAS3; the
// blocks are independent (no mask), so no other block runs first.
int ros2_autotune(const brams::StepParams& p, int first64, int nob, hipStream_t st) {
  GridState& gs = *g.cur;
  const size_t nsc = sizeof(float) * MECH_NS * (size_t)p.n3d;
  const bool to_t = (p.split == brams::kSplitParallel) && (p.ndyn == 1);
  HIP_TRY(g.tune_sc.ensure(nsc * (to_t ? 2 : 1)));
  HIP_TRY(g.tune_dt.ensure(sizeof(double) * nob));
  brams::StepParams q = p;
  q.sc_p = g.tune_sc.as<float>();
  q.sc_t = to_t ? q.sc_p + (size_t)MECH_NS * p.n3d : nullptr;
  q.last_dt = g.tune_dt.as<double>();
  const int forms[2] = {kFormPerLane, kFormPerLane + 2};
  int best = forms[0];
  for (int f : forms) {
    HIP_TRY(hipMemcpyAsync(q.sc_p, p.sc_p, nsc, hipMemcpyDeviceToDevice, st));
    if (to_t) HIP_TRY(hipMemcpyAsync(q.sc_t, p.sc_t, nsc, hipMemcpyDeviceToDevice, st));
    HIP_TRY(hipMemcpyAsync(q.last_dt, p.last_dt, sizeof(double) * nob, hipMemcpyDeviceToDevice, st));
    HIP_TRY(hipEventRecord(g.ev0, st));
    HIP_TRY(brams::launch_ros2(64, (unsigned)(nob - first64), q, first64, st, kPerLaneWps[f - kFormPerLane]));
    HIP_TRY(hipEventRecord(g.ev1, st));
    HIP_TRY(hipEventSynchronize(g.ev1));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, g.ev0, g.ev1));
    gs.ros2_tune_ms[f] = ms;
    if (ms < gs.ros2_tune_ms[best]) best = f;
  }
  gs.ros2_form = best;
  return 0;
}

int launch_ros2_step(brams::StepParams p, hipStream_t st) {
  GridState& gs = *g.cur;
  const int nob = (int)gs.sizes.size();
  p.block_start = gs.block_start.as<int>();
  p.mask = gs.mask.as<unsigned>();
  p.blk_attempts = gs.attempts.as<int>();
  p.blk_substeps = gs.substeps.as<int>();
  p.blk_failed = gs.failed.as<int>();
  p.blk_forced = gs.forced.as<float>();
  p.blk_thresh = gs.thresh.as<int>();
  p.order_thresh = nullptr;
  HIP_TRY(hipMemcpyAsync(gs.mask.p, kStructMask.w, sizeof kStructMask.w, hipMemcpyHostToDevice, st));
  HIP_TRY(hipMemsetAsync(gs.order_thresh.p, 0, sizeof(int), st));
  int nbig = 0;
  while (nbig < nob && gs.sizes[nbig] > 64) ++nbig;
  // brams_chem_set_form forces the ROS2 form too (3 or 5: ROS2 has no
  // two-wave kernel)
  if (g.force_form == kFormPerLane + 1)
    return fail(-6, "ROS2 has no kernel form %d (two waves per SIMD)", g.force_form);
  if (g.force_form < 0 && gs.ros2_form < 0 && nob > nbig) {
    int rc = ros2_autotune(p, nbig, nob, st);
    if (rc) return rc;
  }
  const int form = g.force_form >= 0 ? g.force_form : gs.ros2_form;
  const int* order = nullptr;
  if (nob > nbig && gs.blk_hist) {
    int rc = order_blocks(nbig, nob - nbig, st);
    if (rc) return rc;
    order = g.order.as<int>();
  }
  HIP_TRY(hipEventRecord(g.ev0, st));
  HIP_TRY(brams::launch_ros2(128, (unsigned)nbig, p, 0, st));
  HIP_TRY(hipEventRecord(g.evb, st));
  HIP_TRY(hipEventRecord(g.evm, st));
  g.main_cells = 0;
  if (nob > nbig) {
    brams::StepParams pm = p;
    pm.block_order = order;
    HIP_TRY(brams::launch_ros2(64, (unsigned)(nob - nbig), pm, nbig, st, kPerLaneWps[form - kFormPerLane]));
    for (int b = nbig; b < nob; ++b) g.main_cells += gs.sizes[b];
  }
  HIP_TRY(hipEventRecord(g.ev1, st));
  gs.blk_hist = true;
  g.last_form = nob > nbig ? form : -1;
  return 0;
}

int collect_stats(hipStream_t st) {
  GridState& gs = *g.cur;
  const int nob = (int)gs.sizes.size();
  HIP_TRY(hipStreamSynchronize(st));
  gs.h_att.resize(nob);
  gs.h_sub.resize(nob);
  gs.h_fail.resize(nob);
  gs.h_forced.resize(nob);
  gs.h_thresh.resize(nob);
  unsigned mw[MECH_MASK_WORDS];
  int order_viol = 0;
  HIP_TRY(hipMemcpy(gs.h_att.data(), gs.attempts.p, sizeof(int) * nob, hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(gs.h_sub.data(), gs.substeps.p, sizeof(int) * nob, hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(gs.h_fail.data(), gs.failed.p, sizeof(int) * nob, hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(gs.h_forced.data(), gs.forced.p, sizeof(float) * nob, hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(gs.h_thresh.data(), gs.thresh.p, sizeof(int) * nob, hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(&order_viol, gs.order_thresh.p, sizeof(int), hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(mw, gs.mask.p, sizeof mw, hipMemcpyDeviceToHost));
  brams_chem_stats s{};
  s.cells = gs.ncell;
  s.blocks = nob;
  for (int b = 0; b < nob; ++b) {
    s.attempts += gs.h_att[b];
    s.substeps += gs.h_sub[b];
    s.failed_blocks += gs.h_fail[b];
    s.cell_attempts += (long long)gs.h_att[b] * gs.sizes[b];
    s.cell_substeps += (long long)gs.h_sub[b] * gs.sizes[b];
    if (gs.h_forced[b] > 1.0f) ++s.forced_blocks;
    s.max_forced_err = std::max(s.max_forced_err, gs.h_forced[b]);
    s.threshold_pivots += gs.h_thresh[b];
  }
  s.order_threshold_pivots = order_viol;
  int nz = 0;
  for (int w = 0; w < MECH_MASK_WORDS; ++w) nz += __builtin_popcount(mw[w]);
  s.mask_nnz = nz;
  float ms = 0.f;
  HIP_TRY(hipEventElapsedTime(&ms, g.ev0, g.ev1));
  s.kernel_ms = ms;
  HIP_TRY(hipEventElapsedTime(&ms, g.ev0, g.evb));
  s.block1_ms = ms;
  HIP_TRY(hipEventElapsedTime(&ms, g.evm, g.ev1));
  s.main_ms = ms;
  s.main_cells = g.main_cells;
  s.kernel_form = g.last_form;
  g.stats = s;
  return (int)s.failed_blocks;
}

// grid, photolysis method and time step: the checks every solver shares
int check_grid(int m1, int m2, int m3, int ia, int iz, int ja, int jz, int photoj, float dtlt) {
  if (m1 < 3 || m2 < 1 || m3 < 1) return fail(-2, "bad grid extents %d %d %d", m1, m2, m3);
  if (ia < 1 || iz > m2 || ia > iz || ja < 1 || jz > m3 || ja > jz)
    return fail(-2, "bad interior bounds ia=%d iz=%d ja=%d jz=%d", ia, iz, ja, jz);
  if (photoj == BRAMS_PHOTOJ_LUT)
    return fail(-4, "PhotojMethod LUT is not supported (the TUV mechanisms use FAST-TUV)");
  if (photoj < 0 || photoj > 2) return fail(-2, "unknown PhotojMethod %d", photoj);
  if (!(dtlt > 0.f)) return fail(-2, "dtlt must be positive");
  return 0;
}

// The surrogate's feature and target layout (include/brams_surrogate.h) is
// RELACS_TUV's: a build for another mechanism has no surrogate step.
int check_surrogate_mechanism() {
  if (brams::kSurNin != BRAMS_SUR_NIN || brams::kSurNout != BRAMS_SUR_NOUT || !brams::kSurrogateBuilt)
    return fail(-4, "the surrogate step exists for RELACS_TUV only (%d features, %d species; this build has "
                "%d, %d)", BRAMS_SUR_NIN, BRAMS_SUR_NOUT, brams::kSurNin, brams::kSurNout);
  return 0;
}

int check_common(int m1, int m2, int m3, int ia, int iz, int ja, int jz, int ndyn, int split,
                 int photoj, float dtlt) {
  int rc = check_grid(m1, m2, m3, ia, iz, ja, jz, photoj, dtlt);
  if (rc) return rc;
  if (ndyn < 1) return fail(-2, "N_DYN_CHEM must be >= 1");
  if (split < 0 || split > 2) return fail(-2, "unknown split_method %d", split);
  return 0;
}

}  // namespace

extern "C" {

// shared by the surrogate translation unit: record an error, return code
int brams_chem_set_error(int code, const char* msg) { return fail(code, "%s", msg); }

int brams_chem_init(int device) {
  if (g.init && g.device == device) return 0;
  int n = 0;
  HIP_TRY(hipGetDeviceCount(&n));
  if (device < 0 || device >= n) return fail(-1, "device %d not present (%d devices)", device, n);
  HIP_TRY(hipSetDevice(device));
  hipDeviceProp_t prop;
  HIP_TRY(hipGetDeviceProperties(&prop, device));
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    return fail(-1, "device %d is %s; this library is built for gfx950", device, prop.gcnArchName);
  g.device = device;
  if (!g.ev0) HIP_TRY(hipEventCreate(&g.ev0));
  if (!g.ev1) HIP_TRY(hipEventCreate(&g.ev1));
  if (!g.evb) HIP_TRY(hipEventCreate(&g.evb));
  if (!g.evm) HIP_TRY(hipEventCreate(&g.evm));
  for (auto& e : g.sev)
    if (!e) HIP_TRY(hipEventCreate(&e));
  g.n_cu = prop.multiProcessorCount;
  HIP_TRY(g.qcount.ensure(sizeof(int) * 64));
  unsigned char* na = nullptr;
  if (int rc = scratch_address(&na)) return rc;
  g.init = true;
  return 0;
}

int brams_chem_alloc_index(int32_t block_size, int32_t m1, int32_t ia, int32_t iz, int32_t ja,
                           int32_t jz, int32_t* block_end, int32_t max_nob) {
  std::vector<int> sz;
  long long ncell = (long long)(m1 - 2) * (iz - ia + 1) * (jz - ja + 1);
  int nob = partition(block_size, ncell, sz);
  if (nob < 0) return fail(-2, "empty grid");
  if (block_end) {
    if (nob > max_nob) return fail(-2, "max_nob %d < nob %d", max_nob, nob);
    for (int b = 0; b < nob; ++b) block_end[b] = sz[b];
  }
  return nob;
}

const char* brams_chem_mechanism(int32_t* nspecies, int32_t* nr, int32_t* nr_photo) {
  if (nspecies) *nspecies = MECH_NS;
  if (nr) *nr = MECH_NR;
  if (nr_photo) *nr_photo = MECH_NPHOT;
  return MECH_NAME;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// dt history record (host code only).  Layout, little-endian:
//   0  char[8]  "BRAMSDTH"          8  uint32 version (1)   12 uint32 header bytes (72)
//   16 char[16] mechanism name, NUL-padded
//   32 int32[8] m1 m2 m3 ia iz ja jz block_size
//   64 int32    nob                 68 int32 0
//   72 double[nob] last_accepted_dt
//   72+8*nob uint64 FNV-1a (64-bit) of every byte before it
namespace {

constexpr char kDthMagic[8] = {'B', 'R', 'A', 'M', 'S', 'D', 'T', 'H'};
constexpr uint32_t kDthVersion = 1;
constexpr int64_t kDthHeader = 72;

uint64_t fnv1a64(const unsigned char* p, int64_t n) {
  uint64_t h = 0xcbf29ce484222325ull;
  for (int64_t i = 0; i < n; ++i) {
    h ^= p[i];
    h *= 0x100000001b3ull;
  }
  return h;
}

// the partition's block count of a grid, or <0 (with the error set)
int dth_nob(const brams_dt_history_grid* gr) {
  if (!gr) return fail(-2, "null grid");
  if (gr->m1 < 3 || gr->ia < 1 || gr->iz < gr->ia || gr->iz > gr->m2 || gr->ja < 1 || gr->jz < gr->ja ||
      gr->jz > gr->m3 || gr->block_size < 1)
    return fail(-2, "bad grid (m1 %d m2 %d m3 %d ia..iz %d..%d ja..jz %d..%d block_size %d)", gr->m1, gr->m2,
                gr->m3, gr->ia, gr->iz, gr->ja, gr->jz, gr->block_size);
  std::vector<int> sz;
  return partition(gr->block_size, (long long)(gr->m1 - 2) * (gr->iz - gr->ia + 1) * (gr->jz - gr->ja + 1), sz);
}

bool dt_ok(double v) { return std::isfinite(v) && v > 0.0; }

}  // namespace

extern "C" {

int64_t brams_chem_dt_history_bytes(int32_t nob) {
  if (nob < 1) return fail(-2, "nob %d < 1", nob);
  return kDthHeader + 8 * (int64_t)nob + 8;
}

int64_t brams_chem_dt_history_pack(const brams_dt_history_grid* grid, const double* last_accepted_dt,
                                   int32_t nob, void* buf, int64_t cap) {
  const int want = dth_nob(grid);
  if (want < 0) return want;
  if (nob != want) return fail(-8, "nob %d is not the grid's block count %d", nob, want);
  if (!last_accepted_dt || !buf) return fail(-2, "null pointer");
  const int64_t bytes = brams_chem_dt_history_bytes(nob);
  if (cap < bytes) return fail(-2, "record needs %lld bytes, buffer holds %lld", (long long)bytes, (long long)cap);
  for (int b = 0; b < nob; ++b)
    if (!dt_ok(last_accepted_dt[b])) return fail(-7, "block %d: dt %g is not finite and positive", b + 1,
                                                 last_accepted_dt[b]);
  unsigned char* o = static_cast<unsigned char*>(buf);
  std::memset(o, 0, kDthHeader);
  std::memcpy(o, kDthMagic, 8);
  const uint32_t ver = kDthVersion, hdr = (uint32_t)kDthHeader;
  std::memcpy(o + 8, &ver, 4);
  std::memcpy(o + 12, &hdr, 4);
  std::strncpy(reinterpret_cast<char*>(o + 16), MECH_NAME, 15);
  const int32_t gv[8] = {grid->m1, grid->m2, grid->m3, grid->ia, grid->iz, grid->ja, grid->jz, grid->block_size};
  std::memcpy(o + 32, gv, sizeof gv);
  std::memcpy(o + 64, &nob, 4);
  std::memcpy(o + kDthHeader, last_accepted_dt, 8 * (size_t)nob);
  const uint64_t h = fnv1a64(o, bytes - 8);
  std::memcpy(o + bytes - 8, &h, 8);
  return bytes;
}

int brams_chem_dt_history_unpack(const void* buf, int64_t nbytes, const brams_dt_history_grid* grid,
                                 double* last_accepted_dt, int32_t max_nob) {
  const int want = dth_nob(grid);
  if (want < 0) return want;
  if (!buf || !last_accepted_dt) return fail(-2, "null pointer");
  const unsigned char* r = static_cast<const unsigned char*>(buf);
  if (nbytes < kDthHeader + 16 || std::memcmp(r, kDthMagic, 8) != 0)
    return fail(-7, "not a dt history record");
  uint32_t ver, hdr;
  int32_t nob;
  std::memcpy(&ver, r + 8, 4);
  std::memcpy(&hdr, r + 12, 4);
  std::memcpy(&nob, r + 64, 4);
  if (ver != kDthVersion || hdr != (uint32_t)kDthHeader)
    return fail(-7, "dt history record version %u (header %u bytes); this library reads version %u", ver, hdr,
                kDthVersion);
  if (nob < 1 || nbytes != kDthHeader + 8 * (int64_t)nob + 8)
    return fail(-7, "dt history record of %lld bytes does not hold %d blocks", (long long)nbytes, nob);
  uint64_t h;
  std::memcpy(&h, r + nbytes - 8, 8);
  if (h != fnv1a64(r, nbytes - 8)) return fail(-7, "dt history record checksum mismatch (corrupt record)");
  char mech[17] = {};
  std::memcpy(mech, r + 16, 16);
  if (std::strcmp(mech, MECH_NAME) != 0)
    return fail(-8, "dt history record of mechanism %s restored into a %s library", mech, MECH_NAME);
  int32_t gv[8];
  std::memcpy(gv, r + 32, sizeof gv);
  const int32_t ex[8] = {grid->m1, grid->m2, grid->m3, grid->ia, grid->iz, grid->ja, grid->jz, grid->block_size};
  if (std::memcmp(gv, ex, sizeof gv) != 0)
    return fail(-8, "dt history record of grid %dx%dx%d (%d..%d, %d..%d, blocks of %d) restored into grid "
                "%dx%dx%d (%d..%d, %d..%d, blocks of %d)", gv[0], gv[1], gv[2], gv[3], gv[4], gv[5], gv[6], gv[7],
                ex[0], ex[1], ex[2], ex[3], ex[4], ex[5], ex[6], ex[7]);
  if (nob != want) return fail(-8, "record holds %d blocks, the grid has %d", nob, want);
  if (max_nob < nob) return fail(-2, "max_nob %d < nob %d", max_nob, nob);
  for (int b = 0; b < nob; ++b) {
    double v;
    std::memcpy(&v, r + kDthHeader + 8 * (int64_t)b, 8);
    if (!dt_ok(v)) return fail(-7, "record block %d: dt %g is not finite and positive", b + 1, v);
  }
  std::memcpy(last_accepted_dt, r + kDthHeader, 8 * (size_t)nob);
  return nob;
}

}  // extern "C"

namespace {

// argument checks shared by the device entries (no device work)
int check_device_args(const brams_chem_dev_args* a) {
  if (!a) return fail(-2, "null args");
  int rc = check_common(a->m1, a->m2, a->m3, a->ia, a->iz, a->ja, a->jz, a->n_dyn_chem,
                        a->split_method, a->photoj_method, a->dtlt);
  if (rc) return rc;
  if (!a->pp || !a->pi0 || !a->theta || !a->rv || !a->jphoto || !a->weight || !a->transported ||
      !a->sc_p || !a->last_accepted_dt || !a->atol || !a->rtol)
    return fail(-2, "null device pointer");
  if (a->split_method == BRAMS_SPLIT_PARALLEL && a->n_dyn_chem == 1 && !a->sc_t)
    return fail(-2, "PARALLEL with N_DYN_CHEM=1 needs sc_t");
  if (a->split_method == BRAMS_SPLIT_PARALLEL && a->n_dyn_chem > 1 && !a->sc_t_dyn)
    return fail(-2, "PARALLEL with N_DYN_CHEM>1 needs sc_t_dyn");
  return 0;
}

// the kernel parameters that come from the caller's arguments
void fill_params(const brams_chem_dev_args* a, brams::StepParams& p) {
  p = brams::StepParams{};
  p.m1 = a->m1; p.m2 = a->m2; p.m3 = a->m3;
  p.ia = a->ia; p.iz = a->iz; p.ja = a->ja; p.jz = a->jz;
  p.nk = a->m1 - 2; p.ni = a->iz - a->ia + 1;
  p.dtlt = a->dtlt; p.ndyn = a->n_dyn_chem; p.split = a->split_method;
  p.cp = a->cp; p.cpor = a->cpor; p.p00 = a->p00;
  p.pp = a->pp; p.pi0 = a->pi0; p.theta = a->theta; p.rv = a->rv;
  p.jphoto = a->jphoto;
  p.weight = a->weight; p.transported = a->transported;
  p.sc_p = a->sc_p; p.sc_t = a->sc_t; p.sc_t_dyn = a->sc_t_dyn;
  p.n3d = (long long)a->m1 * a->m2 * a->m3;
  p.last_dt = a->last_accepted_dt;
  p.atol = a->atol; p.rtol = a->rtol;
}

// validate device arguments and fill the kernel parameters (no launch)
int prepare_device(const brams_chem_dev_args* a, brams::StepParams& p, long long& ncell) {
  if (!a) return fail(-2, "null args");
  if (!g.init) {
    int rc = brams_chem_init(0);
    if (rc) return rc;
  }
  int rc = check_device_args(a);
  if (rc) return rc;
  if ((rc = ensure_grid(a->m1, a->m2, a->m3, a->ia, a->iz, a->ja, a->jz))) return rc;
  ncell = g.cur->ncell;
  const int nob = (int)g.cur->sizes.size();
  if (a->nob != nob) return fail(-2, "nob %d does not match the AllocIndex partition (%d)", a->nob, nob);
  fill_params(a, p);
  return 0;
}

}  // namespace

extern "C" {

int brams_chem_rodas3_dyndt_device(const brams_chem_dev_args* a) {
  brams::StepParams p;
  long long ncell = 0;
  int rc = prepare_device(a, p, ncell);
  if (rc) return rc;
  hipStream_t st = (hipStream_t)a->stream;
  if ((rc = launch_step(p, st))) return rc;
  return collect_stats(st);
}

// The same step with no allocation: the bookkeeping sits in `scratch`, the
// blocks run in partition order through the same kernels (block 1, then the
// 65..127-cell blocks on 128 lanes, then the 64-cell blocks).
int brams_chem_rodas3_dyndt_device_noalloc(const brams_chem_dev_args* a) {
  int rc = check_device_args(a);
  if (rc) return rc;
  const long long ncell = (long long)(a->m1 - 2) * (a->iz - a->ia + 1) * (a->jz - a->ja + 1);
  const long long nobl = ncell <= 64 ? 1 : ncell / 64;
  if (nobl > kNoAllocMaxBlocks)
    return fail(-7, "%lld blocks exceed the no-allocation limit of %d blocks (%lld cells)", nobl,
                kNoAllocMaxBlocks, 64LL * kNoAllocMaxBlocks);
  const int nob = (int)nobl;
  if (a->nob != nob) return fail(-2, "nob %d does not match the AllocIndex partition (%d)", a->nob, nob);
  // AllocIndex sizes: base + (b < r) cells
  const long long rest = ncell <= 64 ? 0 : ncell - 64LL * nob;
  const int base = ncell <= 64 ? (int)ncell : 64 + (int)(rest / nob);
  const int r = (int)(rest % nob);
  const int size0 = base + (r > 0 ? 1 : 0);
  if (size0 > 128) return fail(-3, "block of %d cells exceeds 128", size0);
  const int nbig = base > 64 ? nob : (size0 > 64 ? r : 0);   // blocks of more than 64 cells
  unsigned char* s = nullptr;
  if ((rc = scratch_address(&s))) return rc;
  brams::StepParams p;
  fill_params(a, p);
  int* starts = reinterpret_cast<int*>(s + kNaStart);
  p.block_start = starts;
  p.mask = reinterpret_cast<unsigned*>(s + kNaMask);
  p.blk_attempts = reinterpret_cast<int*>(s + kNaAttempts);
  p.blk_substeps = reinterpret_cast<int*>(s + kNaSubsteps);
  p.blk_failed = reinterpret_cast<int*>(s + kNaFailed);
  p.blk_forced = reinterpret_cast<float*>(s + kNaForced);
  p.blk_thresh = reinterpret_cast<int*>(s + kNaThresh);
  p.order_thresh = reinterpret_cast<int*>(s + kNaOrderThresh);
  int* count = reinterpret_cast<int*>(s + kNaCount);
  hipStream_t st = (hipStream_t)a->stream;
  hipLaunchKernelGGL(brams::partition_starts, dim3((unsigned)((nob + 256) / 256)), dim3(256), 0, st, nob,
                     base, r, starts);
  HIP_TRY(hipGetLastError());
  HIP_TRY(brams::launch_rodas3(size0 > 64 ? 128 : 64, true, 1, p, 0, st));
  if (nbig > 1) HIP_TRY(brams::launch_rodas3(128, false, (unsigned)(nbig - 1), p, 1, st));
  const int first64 = nbig > 1 ? nbig : 1;
  if (nob > first64) HIP_TRY(brams::launch_rodas3(64, false, (unsigned)(nob - first64), p, first64, st));
  hipLaunchKernelGGL(brams::count_nonzero, dim3(1), dim3(256), 0, st, p.blk_failed, nob, count);
  HIP_TRY(hipGetLastError());
  int nfail = 0;
  HIP_TRY(hipMemcpyAsync(&nfail, count, sizeof(int), hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  return nfail;
}

// chem_accum on device (chemistry.f90:528-604), for callers that keep the
// tracers in HBM: no synchronisation, stream-ordered.
int brams_chem_accum_device(int32_t m1, int32_t m2, int32_t m3, int32_t ntask, int32_t n_dyn_chem,
                            int32_t n_transported, const int32_t* transp_chem_index, float* sc_t_dyn,
                            const float* sc_t, void* stream) {
  if (m1 < 1 || m2 < 1 || m3 < 1) return fail(-2, "bad grid extents %d %d %d", m1, m2, m3);
  if (n_dyn_chem < 1) return fail(-2, "N_DYN_CHEM must be >= 1");
  if (n_transported < 0 || n_transported > MECH_NS) return fail(-2, "n_transported %d out of range", n_transported);
  if (n_transported > 0 && !transp_chem_index) return fail(-2, "null transp_chem_index");
  if (!sc_t_dyn || (ntask == 1 && !sc_t)) return fail(-2, "null device pointer");
  brams::AccumSpecies sp{};
  sp.n = n_transported;
  for (int t = 0; t < n_transported; ++t) {
    const int n = transp_chem_index[t];
    if (n < 1 || n > MECH_NS) return fail(-2, "transp_chem_index(%d) = %d out of range", t + 1, n);
    sp.idx[t] = n - 1;
  }
  if (n_transported == 0) return 0;
  const long long n3d = (long long)m1 * m2 * m3;
  const float ni = 1.0f / (float)n_dyn_chem;      // REAL(1./N_DYN_CHEM)
  const unsigned gx = (unsigned)std::min<long long>((n3d + 255) / 256, 1024);
  hipLaunchKernelGGL(brams::chem_accum, dim3(gx, (unsigned)n_transported), dim3(256), 0, (hipStream_t)stream,
                     n3d, ni, ntask == 1 ? 1 : 0, sp, sc_t_dyn, ntask == 1 ? sc_t : sc_t_dyn);
  HIP_TRY(hipGetLastError());
  return 0;
}

// Several grids in one pass (the nested grids of a model step).  Block 1 of
// every grid runs first, in one launch (each publishes its own grid's mask),
// then the other 65..127-cell blocks of every grid, then the 64-cell blocks
// of every grid, each in one launch.  From the second call on, the 64-cell
// blocks launch longest first across the grids by their previous attempts.
int brams_chem_rodas3_dyndt_device_grids(const brams_chem_dev_args* a, int32_t ngrids) {
  constexpr int kMaxGrids = 16;
  if (!a || ngrids < 1 || ngrids > kMaxGrids) return fail(-2, "ngrids must be 1..%d", kMaxGrids);
  if (!g.init) {
    int rc = brams_chem_init(0);
    if (rc) return rc;
  }
  std::vector<GridState*> gss(ngrids);
  std::vector<brams::StepParams> ps(ngrids);
  std::vector<int> w1, w128, w64;
  for (int i = 0; i < ngrids; ++i) {
    int rc = check_device_args(&a[i]);
    if (rc) return rc;
    if (a[i].stream != a[0].stream) return fail(-2, "grid %d is on another stream than grid 1", i + 1);
    if ((rc = ensure_grid(a[i].m1, a[i].m2, a[i].m3, a[i].ia, a[i].iz, a[i].ja, a[i].jz))) return rc;
    GridState* gs = g.cur;
    for (int j = 0; j < i; ++j)
      if (gss[j] == gs) return fail(-2, "grids %d and %d have the same extents", j + 1, i + 1);
    const int nob = (int)gs->sizes.size();
    if (a[i].nob != nob)
      return fail(-2, "grid %d: nob %d does not match the AllocIndex partition (%d)", i + 1, a[i].nob, nob);
    if (nob > brams::kGridBlockMask) return fail(-2, "grid %d has too many blocks", i + 1);
    gss[i] = gs;
    brams::StepParams& p = ps[i];
    fill_params(&a[i], p);
    p.block_start = gs->block_start.as<int>();
    p.mask = gs->mask.as<unsigned>();
    p.blk_attempts = gs->attempts.as<int>();
    p.blk_substeps = gs->substeps.as<int>();
    p.blk_failed = gs->failed.as<int>();
    p.blk_forced = gs->forced.as<float>();
    p.blk_thresh = gs->thresh.as<int>();
    p.order_thresh = gs->order_thresh.as<int>();
    const int tag = i << brams::kGridShift;
    int nbig = 0;
    while (nbig < nob && gs->sizes[nbig] > 64) ++nbig;
    w1.push_back(tag);
    for (int b = 1; b < nbig; ++b) w128.push_back(tag | b);
    for (int b = std::max(nbig, 1); b < nob; ++b) w64.push_back(tag | b);
  }
  bool hist = true;
  for (GridState* gs : gss) hist = hist && gs->blk_hist && (int)gs->h_att.size() == (int)gs->sizes.size();
  if (hist)
    std::stable_sort(w64.begin(), w64.end(), [&](int x, int y) {
      return gss[x >> brams::kGridShift]->h_att[x & brams::kGridBlockMask] >
             gss[y >> brams::kGridShift]->h_att[y & brams::kGridBlockMask];
    });
  hipStream_t st = (hipStream_t)a[0].stream;
  const size_t nw = w1.size() + w128.size() + w64.size();
  HIP_TRY(g.multi_ps.ensure(sizeof(brams::StepParams) * ngrids));
  HIP_TRY(g.multi_work.ensure(sizeof(int) * nw));
  std::vector<int> work;
  work.reserve(nw);
  work.insert(work.end(), w1.begin(), w1.end());
  work.insert(work.end(), w128.begin(), w128.end());
  work.insert(work.end(), w64.begin(), w64.end());
  const brams::StepParams* dps = g.multi_ps.as<brams::StepParams>();
  const int* dw = g.multi_work.as<int>();
  HIP_TRY(hipMemcpyAsync(g.multi_ps.p, ps.data(), sizeof(brams::StepParams) * ngrids, hipMemcpyHostToDevice, st));
  HIP_TRY(hipMemcpyAsync(g.multi_work.p, work.data(), sizeof(int) * nw, hipMemcpyHostToDevice, st));
  HIP_TRY(hipEventRecord(g.ev0, st));
  if (brams::kLeanBuild) {
    // a lean library has no rodas3_grids kernels: the same three phases, one
    // launch per grid of that grid's kernels (blocks are independent, so the
    // results are the same; the 64-cell blocks run in block order)
    std::vector<int> nbig(ngrids, 0);
    for (int i = 0; i < ngrids; ++i) {
      const std::vector<int>& sz = gss[i]->sizes;
      while (nbig[i] < (int)sz.size() && sz[nbig[i]] > 64) ++nbig[i];
      HIP_TRY(brams::launch_rodas3(sz[0] > 64 ? 128 : 64, true, 1, ps[i], 0, st));
    }
    HIP_TRY(hipEventRecord(g.evb, st));
    for (int i = 0; i < ngrids; ++i)
      if (nbig[i] > 1) HIP_TRY(brams::launch_rodas3(128, false, (unsigned)(nbig[i] - 1), ps[i], 1, st));
    HIP_TRY(hipEventRecord(g.evm, st));
    for (int i = 0; i < ngrids; ++i) {
      const int first64 = std::max(nbig[i], 1), nob = (int)gss[i]->sizes.size();
      if (nob > first64) HIP_TRY(brams::launch_rodas3(64, false, (unsigned)(nob - first64), ps[i], first64, st, 1));
    }
  } else {
    HIP_TRY(brams::launch_rodas3_grids(128, true, 1, (unsigned)w1.size(), dps, dw, st));
    HIP_TRY(hipEventRecord(g.evb, st));
    HIP_TRY(brams::launch_rodas3_grids(128, false, 1, (unsigned)w128.size(), dps, dw + w1.size(), st));
    HIP_TRY(hipEventRecord(g.evm, st));
    const int wps = w64.size() >= (size_t)brams::kRodas3ManyBlocks ? brams::kRodas3ManyWaves : 1;
    HIP_TRY(brams::launch_rodas3_grids(64, false, wps, (unsigned)w64.size(), dps,
                                       dw + w1.size() + w128.size(), st));
  }
  HIP_TRY(hipEventRecord(g.ev1, st));
  // per-grid counters, then their sums (mask size: the first grid's)
  brams_chem_stats tot{};
  long long main_cells = 0;
  for (int w : w64) main_cells += gss[w >> brams::kGridShift]->sizes[w & brams::kGridBlockMask];
  g.main_cells = main_cells;
  g.last_form = -1;
  for (int i = 0; i < ngrids; ++i) {
    g.cur = gss[i];
    gss[i]->blk_hist = true;
    int rc = collect_stats(st);
    if (rc < 0) return rc;
    const brams_chem_stats& s = g.stats;
    tot.cells += s.cells;
    tot.blocks += s.blocks;
    tot.substeps += s.substeps;
    tot.attempts += s.attempts;
    tot.failed_blocks += s.failed_blocks;
    tot.cell_substeps += s.cell_substeps;
    tot.cell_attempts += s.cell_attempts;
    tot.forced_blocks += s.forced_blocks;
    tot.max_forced_err = std::max(tot.max_forced_err, s.max_forced_err);
    tot.threshold_pivots += s.threshold_pivots;
    tot.order_threshold_pivots += s.order_threshold_pivots;
    if (i == 0) tot.mask_nnz = s.mask_nnz;
    tot.kernel_ms = s.kernel_ms;
    tot.block1_ms = s.block1_ms;
    tot.main_ms = s.main_ms;
    tot.main_cells = s.main_cells;
  }
  tot.kernel_form = -1;
  g.stats = tot;
  return (int)tot.failed_blocks;
}

// ------------------------------------------------------------------ surrogate
int brams_sur_features_raw(const brams_chem_dev_args* a, float* feat, double floor_c) {
  brams::StepParams p;
  long long ncell = 0;
  int rc = check_surrogate_mechanism();
  if (rc || (rc = prepare_device(a, p, ncell))) return rc;
  if (rc) return rc;
  hipStream_t st = (hipStream_t)a->stream;
  hipLaunchKernelGGL(brams::sur_raw_features, dim3((unsigned)((ncell + 255) / 256)), dim3(256), 0, st,
                     p, ncell, floor_c, feat);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(st));
  return 0;
}

int brams_sur_targets_raw(const brams_chem_dev_args* a, const float* sc_p_before, float* targ,
                          double floor_c) {
  brams::StepParams p;
  long long ncell = 0;
  if (int rc0 = check_surrogate_mechanism()) return rc0;
  int rc = prepare_device(a, p, ncell);
  if (rc) return rc;
  hipStream_t st = (hipStream_t)a->stream;
  hipLaunchKernelGGL(brams::sur_raw_targets, dim3((unsigned)((ncell + 255) / 256)), dim3(256), 0, st,
                     p, ncell, floor_c, sc_p_before, targ);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(st));
  return 0;
}

int brams_sur_moments(const float* x, int64_t n, int ncol, double* sum, double* sumsq) {
  if (!x || n <= 0 || ncol <= 0) return fail(-2, "bad moments arguments");
  HIP_TRY(hipMemset(sum, 0, sizeof(double) * ncol));
  HIP_TRY(hipMemset(sumsq, 0, sizeof(double) * ncol));
  const int blocks = (int)std::min<int64_t>(2048, (n + 3) / 4);
  hipLaunchKernelGGL(brams::column_moments, dim3(blocks), dim3(256), 0, nullptr, x, (long long)n, ncol,
                     sum, sumsq);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipDeviceSynchronize());
  return 0;
}

int brams_sur_normalize(const float* raw, int64_t n, int ncol, const float* mu, const float* sd,
                        void* out, int ldo, int out_bf16) {
  if (!raw || !out || n <= 0 || ncol > ldo) return fail(-2, "bad normalize arguments");
  const long long tot = (long long)n * ldo;
  // at most 2^20 workgroups of 256: the grid stays far below 2^32 work-items
  const unsigned blocks = (unsigned)std::min<long long>((tot + 255) / 256, 1LL << 20);
  if (out_bf16)
    hipLaunchKernelGGL(brams::normalize_rows<true>, dim3(blocks), dim3(256), 0, nullptr, raw,
                       (long long)n, ncol, mu, sd, out, ldo);
  else
    hipLaunchKernelGGL(brams::normalize_rows<false>, dim3(blocks), dim3(256), 0, nullptr, raw,
                       (long long)n, ncol, mu, sd, out, ldo);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipDeviceSynchronize());
  return 0;
}

int brams_sur_step(const brams_chem_dev_args* a, const brams_sur_model* m, brams_sur_stats* out) {
  if (!a) return fail(-2, "null args");
  if (int rc0 = check_surrogate_mechanism()) return rc0;
  // the surrogate writes accepted cells straight into sc_p: the SYMMETRIC /
  // SEQUENTIAL form of the write-back only (rodas3:774-798)
  if (a->split_method == BRAMS_SPLIT_PARALLEL)
    return fail(-2, "brams_sur_step supports SYMMETRIC and SEQUENTIAL split methods only");
  if (!m || m->nmembers < 1 || m->nmembers > 4) return fail(-2, "surrogate needs 1..4 members");
  brams::StepParams p;
  long long ncell = 0;
  int rc = prepare_device(a, p, ncell);
  if (rc) return rc;
  GridState& gs = *g.cur;
  hipStream_t st = (hipStream_t)a->stream;
  const long long nrows = (ncell + 63) / 64 * 64;
  const long long nchunk = (ncell + 1023) / 1024;
  const long long maxfb = (ncell + 63) / 64;
  HIP_TRY(g.sur_X.ensure(sizeof(__bf16) * nrows * brams::kSurNinPad));
  HIP_TRY(g.sur_out.ensure(sizeof(float) * nrows * brams::kSurNoutPad * m->nmembers));
  HIP_TRY(g.sur_flag.ensure(sizeof(int) * ncell));
  HIP_TRY(g.sur_list.ensure(sizeof(int) * ncell));
  HIP_TRY(g.sur_chunk.ensure(sizeof(int) * (2 * nchunk + 2)));
  HIP_TRY(g.sur_reason.ensure(sizeof(int) * 4));
  HIP_TRY(g.fb_dt.ensure(sizeof(double) * maxfb));
  HIP_TRY(g.fb_att.ensure(sizeof(int) * maxfb * 3));
  HIP_TRY(g.mask_all.ensure(sizeof(unsigned) * MECH_MASK_WORDS));
  if (gs.hist_cells != ncell) {
    // fallback memory: each cell starts from its partition block's step
    HIP_TRY(gs.cell_dt.ensure(sizeof(double) * ncell));
    hipLaunchKernelGGL(brams::cell_history_init, dim3((unsigned)((ncell + 255) / 256)), dim3(256), 0, st,
                       gs.block_start.as<int>(), (int)a->nob, p.last_dt, ncell, gs.cell_dt.as<double>());
    HIP_TRY(hipGetLastError());
    gs.hist_cells = ncell;
  }
  hipEvent_t* ev = g.sev;
  HIP_TRY(hipEventRecord(ev[0], st));
  // 1. features
  hipLaunchKernelGGL(brams::sur_features, dim3((unsigned)((nrows + 255) / 256)), dim3(256), 0, st, p,
                     ncell, nrows, m->floor_c, m->feat_mu, m->feat_sd, g.sur_X.as<__bf16>());
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipEventRecord(ev[1], st));
  // 2. ensemble inference on the matrix cores
  brams::SurApplyParams ap{};
  float* outs[4] = {nullptr, nullptr, nullptr, nullptr};
  for (int k = 0; k < m->nmembers; ++k) {
    outs[k] = g.sur_out.as<float>() + (long long)k * nrows * brams::kSurNoutPad;
    ap.out[k] = outs[k];
  }
  rc = brams_mlp_forward_ensemble(m->members, m->nmembers, g.sur_X.p, nrows, outs, st);
  if (rc) return rc;
  HIP_TRY(hipEventRecord(ev[2], st));
  // 3. checks, write-back of accepted cells, ordered list of rejected ones
  ap.nmembers = m->nmembers;
  ap.nrows = nrows;
  ap.tmu = m->targ_mu; ap.tsd = m->targ_sd;
  ap.natoms = m->natoms; ap.catoms = m->catoms;
  ap.floor_c = m->floor_c; ap.sig_c = m->sig_c;
  ap.spread_tol = m->spread_tol; ap.n_tol = m->n_tol; ap.c_tol = m->c_tol;
  ap.flag = g.sur_flag.as<int>();
  ap.reason = g.sur_reason.as<int>();
  HIP_TRY(hipMemsetAsync(ap.reason, 0, sizeof(int) * 4, st));
  hipLaunchKernelGGL(brams::sur_apply, dim3((unsigned)((ncell + 127) / 128)), dim3(128), 0, st, p, ap,
                     ncell);
  HIP_TRY(hipGetLastError());
  int* cnt = g.sur_chunk.as<int>();
  int* off = cnt + nchunk;
  hipLaunchKernelGGL(brams::compact_count, dim3((unsigned)nchunk), dim3(256), 0, st, ap.flag, ncell, cnt);
  hipLaunchKernelGGL(brams::compact_scan, dim3(1), dim3(256), 0, st, cnt, (int)nchunk, off);
  hipLaunchKernelGGL(brams::compact_write, dim3((unsigned)nchunk), dim3(256), 0, st, ap.flag, ncell, off,
                     g.sur_list.as<int>());
  HIP_TRY(hipGetLastError());
  int nfb = 0;
  int reasons[4];
  HIP_TRY(hipMemcpyAsync(&nfb, off + nchunk, sizeof(int), hipMemcpyDeviceToHost, st));
  HIP_TRY(hipMemcpyAsync(reasons, ap.reason, sizeof reasons, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipEventRecord(ev[3], st));
  HIP_TRY(hipStreamSynchronize(st));
  // 4. RODAS3 on the rejected cells in blocks of 64 consecutive list entries
  //    (ascending cell index, so a block's gathers stay within a few
  //    columns) sharing one step, with the full structural Jacobian; each
  //    block's first step is the smallest step its cells remember.  Sorting
  //    the cells by their previous cost instead was measured slower on the
  //    1.6M-cell rollout (68 against 57 ms): scattered gathers cost more than
  //    the more uniform blocks save.
  long long fb_att = 0, fb_cell_att = 0, fb_cell_sub = 0, fb_failed = 0;
  int fb_form = -1;
  if (nfb > 0) {
    const int nb = (nfb + 63) / 64;
    int* list = g.sur_list.as<int>();
    hipLaunchKernelGGL(brams::fallback_dt_hint, dim3((unsigned)nb), dim3(64), 0, st, list, nfb,
                       gs.cell_dt.as<double>(), (double)(a->dtlt * (float)a->n_dyn_chem),
                       g.fb_dt.as<double>());
    HIP_TRY(hipGetLastError());
    unsigned ones[MECH_MASK_WORDS];
    for (auto& w : ones) w = 0xffffffffu;
    HIP_TRY(hipMemcpyAsync(g.mask_all.p, ones, sizeof ones, hipMemcpyHostToDevice, st));
    brams::StepParams q = p;
    q.cell_list = list;
    q.list_n = nfb;
    q.last_dt = g.fb_dt.as<double>();
    q.mask = g.mask_all.as<unsigned>();
    q.blk_attempts = g.fb_att.as<int>();
    q.blk_substeps = q.blk_attempts + maxfb;
    q.blk_failed = q.blk_attempts + 2 * maxfb;
    q.blk_forced = nullptr;
    q.blk_thresh = nullptr;
    q.order_thresh = nullptr;
    q.block_start = nullptr;
    if (kSlabEnabled) {
      long long nslots = 0;
      if ((rc = slab_slots(brams::kSlabList, &nslots))) return rc;
      nslots = std::min<long long>(nslots, nb);
      if ((rc = ensure_slabs(nslots))) return rc;
      HIP_TRY(brams::launch_rodas3_slab(brams::kSlabList, (unsigned)nslots, q, queue(nullptr, 0, nb), st));
    } else {
      fb_form = nb >= kListManyBlocks ? 1 : 0;
      HIP_TRY(brams::launch_rodas3_list((unsigned)nb, q, st, kListWps[fb_form]));
    }
    hipLaunchKernelGGL(brams::cell_history_update, dim3((unsigned)((nfb + 255) / 256)), dim3(256), 0, st, list,
                       nfb, q.last_dt, gs.cell_dt.as<double>());
    HIP_TRY(hipGetLastError());
    std::vector<int> att(nb), sub(nb), bad(nb);
    HIP_TRY(hipMemcpyAsync(att.data(), q.blk_attempts, sizeof(int) * nb, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(sub.data(), q.blk_substeps, sizeof(int) * nb, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(bad.data(), q.blk_failed, sizeof(int) * nb, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    for (int b = 0; b < nb; ++b) {
      const long long n = std::min(64, nfb - 64 * b);
      fb_att += att[b];
      fb_cell_att += n * att[b];
      fb_cell_sub += n * sub[b];
      fb_failed += bad[b];
    }
  }
  HIP_TRY(hipEventRecord(ev[4], st));
  HIP_TRY(hipStreamSynchronize(st));
  if (out) {
    brams_sur_stats s{};
    s.cells = ncell;
    s.fallback_cells = nfb;
    for (int k = 0; k < 4; ++k) s.reasons[k] = reasons[k];
    s.fallback_block_attempts = fb_att;
    s.fallback_cell_attempts = fb_cell_att;
    s.fallback_cell_substeps = fb_cell_sub;
    s.fallback_failed_blocks = fb_failed;
    s.fallback_wps = fb_form < 0 ? 0 : kListWps[fb_form];
    float ms;
    HIP_TRY(hipEventElapsedTime(&ms, ev[0], ev[1])); s.features_ms = ms;
    HIP_TRY(hipEventElapsedTime(&ms, ev[1], ev[2])); s.mlp_ms = ms;
    HIP_TRY(hipEventElapsedTime(&ms, ev[2], ev[3])); s.apply_ms = ms;
    HIP_TRY(hipEventElapsedTime(&ms, ev[3], ev[4])); s.fallback_ms = ms;
    HIP_TRY(hipEventElapsedTime(&ms, ev[0], ev[4])); s.total_ms = ms;
    *out = s;
  }
  // a fallback block that hit a zero pivot, a non-finite state or the
  // attempt bound was left at its last accepted state: reported like the
  // plain step's failed blocks
  return (int)fb_failed;
}

void brams_sur_reset(void) {
  for (GridState* gs : g.grids) gs->hist_cells = -1;
}

}  // extern "C"

namespace {
int host_step(const brams_chem_args* a, int solver);
}  // namespace

extern "C" {

int brams_chem_rodas3_dyndt(const brams_chem_args* a) { return host_step(a, brams::kRodas3); }

int brams_chem_ros_dyndt(const brams_chem_args* a) { return host_step(a, brams::kRos2); }

int brams_chem_ros_dyndt_device(const brams_chem_dev_args* a) {
  brams::StepParams p;
  long long ncell = 0;
  int rc = prepare_device(a, p, ncell);
  if (rc) return rc;
  hipStream_t st = (hipStream_t)a->stream;
  if ((rc = launch_ros2_step(p, st))) return rc;
  return collect_stats(st);
}

}  // extern "C"

namespace {

// one chemistry step with host arrays (chem_rodas3_dyndt / chem_ros_dyndt)
int host_step(const brams_chem_args* a, int solver) {
  if (!a) return fail(-2, "null args");
  if (!g.init) {
    int rc = brams_chem_init(0);
    if (rc) return rc;
  }
  int rc = check_common(a->m1, a->m2, a->m3, a->ia, a->iz, a->ja, a->jz, a->n_dyn_chem,
                        a->split_method, a->photoj_method, a->dtlt);
  if (rc) return rc;
  if (a->nspecies != MECH_NS || a->nr != MECH_NR || a->nr_photo != MECH_NPHOT)
    return fail(-2, "mechanism sizes %d/%d/%d differ from this library's %s %d/%d/%d", a->nspecies, a->nr,
                a->nr_photo, MECH_NAME, MECH_NS, MECH_NR, MECH_NPHOT);
  if ((rc = ensure_grid(a->m1, a->m2, a->m3, a->ia, a->iz, a->ja, a->jz))) return rc;
  const int nob = (int)g.cur->sizes.size();
  if (a->nob != nob) return fail(-2, "nob %d does not match the AllocIndex partition (%d)", a->nob, nob);
  for (int b = 0; b < nob; ++b)
    if (a->block_end[b] != g.cur->sizes[b])
      return fail(-2, "block_end(%d)=%d differs from AllocIndex (%d)", b + 1, a->block_end[b],
                  g.cur->sizes[b]);
  // transport flags from the caller's index lists
  int flags[MECH_NS];
  for (int s = 0; s < MECH_NS; ++s) flags[s] = -1;
  for (int t = 0; t < a->n_transported; ++t) {
    int n = a->transp_chem_index[t];
    if (n < 1 || n > MECH_NS) return fail(-2, "transp_chem_index out of range");
    flags[n - 1] = 1;
  }
  for (int t = 0; t < a->n_not_transported; ++t) {
    int n = a->no_transp_chem_index[t];
    if (n < 1 || n > MECH_NS) return fail(-2, "no_transp_chem_index out of range");
    flags[n - 1] = 0;
  }
  for (int s = 0; s < MECH_NS; ++s)
    if (flags[s] < 0) return fail(-2, "species %d in neither transport list", s + 1);
  const bool need_t = a->split_method == BRAMS_SPLIT_PARALLEL && a->n_dyn_chem == 1;
  const bool need_td = a->split_method == BRAMS_SPLIT_PARALLEL && a->n_dyn_chem > 1;
  if (need_t && !a->sc_t) return fail(-2, "PARALLEL with N_DYN_CHEM=1 needs sc_t");
  if (need_td && !a->sc_t_dyn) return fail(-2, "PARALLEL with N_DYN_CHEM>1 needs sc_t_dyn");

  const long long n3d = (long long)a->m1 * a->m2 * a->m3;
  const size_t fsz = sizeof(float) * n3d;
  HIP_TRY(g.sc_p.ensure(fsz * MECH_NS));
  HIP_TRY(g.met.ensure(fsz * 4));
  HIP_TRY(g.jphoto.ensure(sizeof(double) * MECH_NPHOT * n3d));
  HIP_TRY(g.last_dt.ensure(sizeof(double) * nob));
  // consts: weight[MECH_NS] f32 | transported[MECH_NS] i32 | atol[MECH_NS] f64 | rtol[MECH_NS] f64
  const size_t coff_w = 0, coff_t = kConstOffT, coff_a = kConstOffA, coff_r = kConstOffR;
  HIP_TRY(g.consts.ensure(kConstBytes));
  if (need_t) HIP_TRY(g.sc_t.ensure(fsz * MECH_NS));
  if (need_td) HIP_TRY(g.sc_t_dyn.ensure(fsz * MECH_NS));
  hipStream_t st = nullptr;
  char* sc = g.sc_p.as<char>();
  for (int s = 0; s < MECH_NS; ++s)
    HIP_TRY(hipMemcpyAsync(sc + s * fsz, a->sc_p[s], fsz, hipMemcpyHostToDevice, st));
  if (need_t)
    for (int s = 0; s < MECH_NS; ++s)
      HIP_TRY(hipMemcpyAsync(g.sc_t.as<char>() + s * fsz, a->sc_t[s], fsz, hipMemcpyHostToDevice, st));
  if (need_td)
    for (int s = 0; s < MECH_NS; ++s)
      HIP_TRY(hipMemcpyAsync(g.sc_t_dyn.as<char>() + s * fsz, a->sc_t_dyn[s], fsz,
                             hipMemcpyHostToDevice, st));
  char* met = g.met.as<char>();
  HIP_TRY(hipMemcpyAsync(met + 0 * fsz, a->pp, fsz, hipMemcpyHostToDevice, st));
  HIP_TRY(hipMemcpyAsync(met + 1 * fsz, a->pi0, fsz, hipMemcpyHostToDevice, st));
  HIP_TRY(hipMemcpyAsync(met + 2 * fsz, a->theta, fsz, hipMemcpyHostToDevice, st));
  HIP_TRY(hipMemcpyAsync(met + 3 * fsz, a->rv, fsz, hipMemcpyHostToDevice, st));
  HIP_TRY(hipMemcpyAsync(g.jphoto.p, a->jphoto, sizeof(double) * MECH_NPHOT * n3d,
                         hipMemcpyHostToDevice, st));
  HIP_TRY(hipMemcpyAsync(g.last_dt.p, a->last_accepted_dt, sizeof(double) * nob,
                         hipMemcpyHostToDevice, st));
  char cbuf[kConstBytes];
  std::memcpy(cbuf + coff_w, a->weight, sizeof(float) * MECH_NS);
  std::memcpy(cbuf + coff_t, flags, sizeof(int) * MECH_NS);
  std::memcpy(cbuf + coff_a, a->atol, sizeof(double) * MECH_NS);
  std::memcpy(cbuf + coff_r, a->rtol, sizeof(double) * MECH_NS);
  HIP_TRY(hipMemcpy(g.consts.p, cbuf, sizeof cbuf, hipMemcpyHostToDevice));

  brams::StepParams p{};
  p.m1 = a->m1; p.m2 = a->m2; p.m3 = a->m3;
  p.ia = a->ia; p.iz = a->iz; p.ja = a->ja; p.jz = a->jz;
  p.nk = a->m1 - 2; p.ni = a->iz - a->ia + 1;
  p.dtlt = a->dtlt; p.ndyn = a->n_dyn_chem; p.split = a->split_method;
  p.cp = a->cp; p.cpor = a->cpor; p.p00 = a->p00;
  p.pp = (const float*)(met + 0 * fsz);
  p.pi0 = (const float*)(met + 1 * fsz);
  p.theta = (const float*)(met + 2 * fsz);
  p.rv = (const float*)(met + 3 * fsz);
  p.jphoto = g.jphoto.as<double>();
  p.weight = (const float*)(g.consts.as<char>() + coff_w);
  p.transported = (const int*)(g.consts.as<char>() + coff_t);
  p.atol = (const double*)(g.consts.as<char>() + coff_a);
  p.rtol = (const double*)(g.consts.as<char>() + coff_r);
  p.sc_p = g.sc_p.as<float>();
  p.sc_t = need_t ? g.sc_t.as<float>() : nullptr;
  p.sc_t_dyn = need_td ? g.sc_t_dyn.as<float>() : nullptr;
  p.n3d = n3d;
  p.last_dt = g.last_dt.as<double>();
  if ((rc = solver == brams::kRos2 ? launch_ros2_step(p, st) : launch_step(p, st))) return rc;
  // results back into the caller's arrays
  if (need_t) {
    for (int s = 0; s < MECH_NS; ++s)
      HIP_TRY(hipMemcpyAsync(a->sc_t[s], g.sc_t.as<char>() + s * fsz, fsz, hipMemcpyDeviceToHost, st));
  } else {
    for (int s = 0; s < MECH_NS; ++s)
      HIP_TRY(hipMemcpyAsync(a->sc_p[s], sc + s * fsz, fsz, hipMemcpyDeviceToHost, st));
  }
  HIP_TRY(hipMemcpyAsync(a->last_accepted_dt, g.last_dt.p, sizeof(double) * nob,
                         hipMemcpyDeviceToHost, st));
  return collect_stats(st);
}

// One QSSA step (chem_qssa) on device arrays already in p: the counter is
// cleared, the kernel runs between the step events, and the call waits for
// it to read the count of cells that reached 1e20 (the stats' failed count).
int qssa_step(const brams::StepParams& p, int itchim, long long ncell, hipStream_t st) {
  HIP_TRY(g.qssa_fatal.ensure(sizeof(int)));
  HIP_TRY(hipMemsetAsync(g.qssa_fatal.p, 0, sizeof(int), st));
  HIP_TRY(hipEventRecord(g.ev0, st));
  HIP_TRY(brams::launch_qssa(p, itchim, ncell, g.qssa_fatal.as<int>(), st));
  HIP_TRY(hipEventRecord(g.ev1, st));
  HIP_TRY(hipStreamSynchronize(st));
  int nfatal = 0;
  HIP_TRY(hipMemcpy(&nfatal, g.qssa_fatal.p, sizeof(int), hipMemcpyDeviceToHost));
  brams_chem_stats s{};
  s.cells = ncell;
  s.failed_blocks = nfatal;
  float ms = 0.f;
  HIP_TRY(hipEventElapsedTime(&ms, g.ev0, g.ev1));
  s.kernel_ms = ms;
  s.main_ms = ms;
  s.main_cells = ncell;
  s.kernel_form = -1;
  g.stats = s;
  return nfatal;
}

int check_qssa(int itchim, int nspecies) {
  if (itchim < 1) return fail(-2, "itchim must be >= 1 (chemistry.f90 passes 75)");
  if (nspecies != MECH_NS)
    return fail(-2, "nspecies %d differs from this library's %s %d", nspecies, MECH_NAME, MECH_NS);
  return 0;
}

}  // namespace

extern "C" {

int brams_chem_qssa_device(const brams_chem_dev_args* a, int32_t itchim) {
  if (!a) return fail(-2, "null args");
  int rc = check_grid(a->m1, a->m2, a->m3, a->ia, a->iz, a->ja, a->jz, a->photoj_method, a->dtlt);
  if (rc || (rc = check_qssa(itchim, MECH_NS))) return rc;
  if (!a->pp || !a->pi0 || !a->theta || !a->rv || !a->jphoto || !a->weight || !a->sc_p || !a->sc_t)
    return fail(-2, "null device pointer (QSSA needs pp, pi0, theta, rv, jphoto, weight, sc_p, sc_t)");
  if (!g.init && (rc = brams_chem_init(0))) return rc;
  brams::StepParams p;
  fill_params(a, p);
  const long long ncell = (long long)(a->m1 - 2) * (a->iz - a->ia + 1) * (a->jz - a->ja + 1);
  return qssa_step(p, itchim, ncell, (hipStream_t)a->stream);
}

int brams_chem_qssa(const brams_chem_args* a, int32_t itchim) {
  if (!a) return fail(-2, "null args");
  int rc = check_grid(a->m1, a->m2, a->m3, a->ia, a->iz, a->ja, a->jz, a->photoj_method, a->dtlt);
  if (rc || (rc = check_qssa(itchim, a->nspecies))) return rc;
  if (a->nr != MECH_NR || a->nr_photo != MECH_NPHOT)
    return fail(-2, "mechanism sizes %d/%d differ from this library's %s %d/%d", a->nr, a->nr_photo,
                MECH_NAME, MECH_NR, MECH_NPHOT);
  if (!a->pp || !a->pi0 || !a->theta || !a->rv || !a->jphoto || !a->weight || !a->sc_p || !a->sc_t)
    return fail(-2, "null argument (QSSA needs pp, pi0, theta, rv, jphoto, weight, sc_p, sc_t)");
  for (int s = 0; s < MECH_NS; ++s)
    if (!a->sc_p[s] || !a->sc_t[s]) return fail(-2, "species %d has no sc_p or sc_t array", s + 1);
  if (!g.init && (rc = brams_chem_init(0))) return rc;
  const long long n3d = (long long)a->m1 * a->m2 * a->m3;
  const size_t fsz = sizeof(float) * n3d;
  HIP_TRY(g.sc_p.ensure(fsz * MECH_NS));
  HIP_TRY(g.sc_t.ensure(fsz * MECH_NS));
  HIP_TRY(g.met.ensure(fsz * 4));
  HIP_TRY(g.jphoto.ensure(sizeof(double) * MECH_NPHOT * n3d));
  HIP_TRY(g.consts.ensure(kConstBytes));
  hipStream_t st = nullptr;
  char* sc = g.sc_p.as<char>();
  char* tc = g.sc_t.as<char>();
  for (int s = 0; s < MECH_NS; ++s) {
    HIP_TRY(hipMemcpyAsync(sc + s * fsz, a->sc_p[s], fsz, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(tc + s * fsz, a->sc_t[s], fsz, hipMemcpyHostToDevice, st));
  }
  char* met = g.met.as<char>();
  HIP_TRY(hipMemcpyAsync(met + 0 * fsz, a->pp, fsz, hipMemcpyHostToDevice, st));
  HIP_TRY(hipMemcpyAsync(met + 1 * fsz, a->pi0, fsz, hipMemcpyHostToDevice, st));
  HIP_TRY(hipMemcpyAsync(met + 2 * fsz, a->theta, fsz, hipMemcpyHostToDevice, st));
  HIP_TRY(hipMemcpyAsync(met + 3 * fsz, a->rv, fsz, hipMemcpyHostToDevice, st));
  HIP_TRY(hipMemcpyAsync(g.jphoto.p, a->jphoto, sizeof(double) * MECH_NPHOT * n3d, hipMemcpyHostToDevice, st));
  HIP_TRY(hipMemcpy(g.consts.p, a->weight, sizeof(float) * MECH_NS, hipMemcpyHostToDevice));
  brams::StepParams p{};
  p.m1 = a->m1; p.m2 = a->m2; p.m3 = a->m3;
  p.ia = a->ia; p.iz = a->iz; p.ja = a->ja; p.jz = a->jz;
  p.nk = a->m1 - 2; p.ni = a->iz - a->ia + 1;
  p.dtlt = a->dtlt;
  p.cp = a->cp; p.cpor = a->cpor; p.p00 = a->p00;
  p.pp = (const float*)(met + 0 * fsz);
  p.pi0 = (const float*)(met + 1 * fsz);
  p.theta = (const float*)(met + 2 * fsz);
  p.rv = (const float*)(met + 3 * fsz);
  p.jphoto = g.jphoto.as<double>();
  p.weight = g.consts.as<float>();
  p.sc_p = g.sc_p.as<float>();
  p.sc_t = g.sc_t.as<float>();
  p.n3d = n3d;
  const long long ncell = (long long)(a->m1 - 2) * (a->iz - a->ia + 1) * (a->jz - a->ja + 1);
  const int nfatal = qssa_step(p, itchim, ncell, st);
  if (nfatal < 0) return nfatal;
  for (int s = 0; s < MECH_NS; ++s)
    HIP_TRY(hipMemcpyAsync(a->sc_t[s], tc + s * fsz, fsz, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  return nfatal;
}

int brams_chem_tuv_jphoto_device(const brams_tuv_map_args* a) {
  if (!a) return fail(-2, "null args");
  if (a->m1 < 3 || a->m2 < 1 || a->m3 < 1 || a->ia < 1 || a->iz > a->m2 || a->ia > a->iz || a->ja < 1 ||
      a->jz > a->m3 || a->ja > a->jz)
    return fail(-2, "bad grid %d %d %d / interior %d..%d %d..%d", a->m1, a->m2, a->m3, a->ia, a->iz, a->ja, a->jz);
  if (a->nr_photo != MECH_NPHOT)
    return fail(-2, "nr_photo %d differs from this library's %s %d", a->nr_photo, MECH_NAME, MECH_NPHOT);
  if (a->kj < 1 || a->kz < a->m1 - 1)
    return fail(-2, "valj must hold kj >= 1 reactions and kz >= m1-1 = %d levels (kj %d, kz %d)", a->m1 - 1,
                a->kj, a->kz);
  if (!a->valj || !a->xref || !a->jphoto) return fail(-2, "null pointer");
  brams::TuvMapParams p{};
  p.m1 = a->m1; p.m2 = a->m2; p.m3 = a->m3;
  p.ia = a->ia; p.iz = a->iz; p.ja = a->ja; p.jz = a->jz;
  p.kj = a->kj; p.kz = a->kz;
  p.valj = a->valj;
  p.jphoto = a->jphoto;
  for (int n = 0; n < MECH_NPHOT; ++n) p.src[n] = -1;
  for (int k = 0; k < a->kj; ++k) {            // ascending: a later reaction overwrites an earlier one
    const int x = a->xref[k];
    if (x < 0 || x > MECH_NPHOT) return fail(-2, "xref(%d) = %d outside 0..%d", k + 1, x, MECH_NPHOT);
    if (x > 0) p.src[x - 1] = k;
  }
  int rc;
  if (!g.init && (rc = brams_chem_init(0))) return rc;
  hipStream_t st = (hipStream_t)a->stream;
  const long long ni = a->iz - a->ia + 1, nj = a->jz - a->ja + 1;
  const long long npts = (long long)a->m1 * ni * nj, n3d = (long long)a->m1 * a->m2 * a->m3;
  HIP_TRY(hipEventRecord(g.ev0, st));
  const unsigned tiles = (unsigned)((npts + 255) / 256), chunk = (tiles + brams::kXcds - 1) / brams::kXcds;
  hipLaunchKernelGGL(brams::tuv_map_interior, dim3(chunk * brams::kXcds), dim3(256), 0, st, p, npts, chunk);
  HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL(brams::jphoto_relations_outside, dim3((unsigned)((n3d + 255) / 256)), dim3(256), 0, st, p,
                     n3d);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipEventRecord(g.ev1, st));
  HIP_TRY(hipStreamSynchronize(st));
  brams_chem_stats s{};
  s.cells = npts;
  float ms = 0.f;
  HIP_TRY(hipEventElapsedTime(&ms, g.ev0, g.ev1));
  s.kernel_ms = ms;
  s.main_ms = ms;
  s.main_cells = npts;
  s.kernel_form = -1;
  g.stats = s;
  return 0;
}

int brams_chem_block_status(int32_t* attempts, int32_t* substeps, int32_t* failed,
                            float* forced_err, int32_t max_nob) {
  if (!g.cur) return fail(-2, "no chemistry step has run");
  const GridState& gs = *g.cur;
  const int nob = (int)gs.h_att.size();
  if (nob > max_nob) return fail(-2, "max_nob %d < nob %d", max_nob, nob);
  for (int b = 0; b < nob; ++b) {
    if (attempts) attempts[b] = gs.h_att[b];
    if (substeps) substeps[b] = gs.h_sub[b];
    if (failed) failed[b] = gs.h_fail[b];
    if (forced_err) forced_err[b] = gs.h_forced[b];
  }
  return nob;
}

int brams_chem_block_thresholds(int32_t* counts, int32_t max_nob) {
  if (!g.cur) return fail(-2, "no chemistry step has run");
  const GridState& gs = *g.cur;
  const int nob = (int)gs.h_thresh.size();
  if (nob > max_nob) return fail(-2, "max_nob %d < nob %d", max_nob, nob);
  for (int b = 0; b < nob && counts; ++b) counts[b] = gs.h_thresh[b];
  return nob;
}

int brams_chem_get_stats(brams_chem_stats* out) {
  if (!out) return fail(-2, "null stats");
  *out = g.stats;
  return 0;
}

const char* brams_chem_last_error(void) { return g_err.c_str(); }

void brams_chem_finalize(void) {
  for (GridState* gs : g.grids) {
    DevBuf* gb[] = {&gs->block_start, &gs->mask, &gs->attempts, &gs->substeps, &gs->failed,
                    &gs->forced, &gs->thresh, &gs->order_thresh, &gs->cell_dt};
    for (DevBuf* b : gb) b->release();
    delete gs;
  }
  g.grids.clear();
  g.cur = nullptr;
  DevBuf* bufs[] = {&g.sc_p, &g.sc_t, &g.sc_t_dyn, &g.met, &g.jphoto, &g.last_dt, &g.consts,
                    &g.sur_X, &g.sur_out, &g.sur_flag, &g.sur_list, &g.sur_chunk,
                    &g.sur_reason, &g.fb_dt, &g.fb_att, &g.mask_all, &g.sort_keys, &g.sort_ids,
                    &g.sort_count, &g.sort_off, &g.order, &g.slab_lu, &g.slab_jac, &g.slab_rk,
                    &g.slab_vec, &g.qcount, &g.tune_sc, &g.tune_dt};
  for (DevBuf* b : bufs) b->release();
  hipEvent_t* evs[] = {&g.ev0, &g.ev1, &g.evb, &g.evm, &g.sev[0], &g.sev[1], &g.sev[2], &g.sev[3],
                       &g.sev[4]};
  for (hipEvent_t* e : evs) {
    if (*e) hipEventDestroy(*e);
    *e = nullptr;
  }
  g.slab_slots = 0;
  g.init = false;
}

int brams_chem_set_form(int32_t form) {
  if (form < -1 || form >= kForms) return fail(-2, "kernel form %d out of range (-1..%d)", form, kForms - 1);
  if (!kSlabEnabled && form >= 0 && form < kFormPerLane)
    return fail(-6, "kernel form %d (slab) is switched off in this build", form);
  if (form >= kFormPerLane && !form_built(form))
    return fail(-6, "kernel form %d is not built into this %s library (lean build)", form, MECH_NAME);
  g.force_form = form;
  return 0;
}

int brams_chem_get_tuning(int32_t* form, float* ms, int32_t n) {
  if (!g.cur) return fail(-2, "no chemistry step has run");
  if (form) *form = g.cur->form;
  for (int f = 0; f < n && f < kForms && ms; ++f) ms[f] = g.cur->tune_ms[f];
  return kForms;
}

}  // extern "C"
