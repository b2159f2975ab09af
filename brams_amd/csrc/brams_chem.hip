// C-ABI of the CCATT RODAS3 chemistry step on MI355X (include/brams_chem.h).
//
// Host side of the drop-in for chem_rodas3_dyndt
// (src/brams/ccatt/chem_spack_rodas3_dyndt.f90:41-131): validates the
// arguments, keeps persistent device buffers per grid size, moves host arrays
// to HBM and back for the Fortran entry point, and launches the kernels of
// rodas3_kernel.h:
//   1. block 1 alone (its value-based sparsity mask changes every attempt and
//      fixes the mask for every later block, rodas3:470-506);
//   2. the remaining blocks, 128-lane workgroups for the leading blocks with
//      more than 64 cells and wave64 workgroups for the rest.
#include <hip/hip_runtime.h>

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
#include "sched_kernels.h"
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

struct State {
  bool init = false;
  int device = 0;
  // persistent per-call buffers
  DevBuf block_start, mask, attempts, substeps, failed, forced;
  // host-entry staging
  DevBuf sc_p, sc_t, sc_t_dyn, met, jphoto, last_dt, consts;
  // surrogate step
  DevBuf sur_X, sur_out, sur_flag, sur_list, sur_chunk, sur_reason, fb_dt, fb_att, mask_all;
  // work ordering (sched_kernels.h): sort scratch, block order, and the
  // fallback's per-cell step memory
  DevBuf sort_keys, sort_ids, sort_count, sort_off, order, cell_dt;
  bool blk_hist = false;         // attempts[] holds a previous call on this partition
  long long hist_cells = -1;     // cells of the fallback memory (-1: to be started)
  hipEvent_t ev0 = nullptr, ev1 = nullptr, evb = nullptr, evm = nullptr;
  long long main_cells = 0;
  brams_chem_stats stats{};
  std::vector<int> part_sizes;   // block sizes of the cached partition
  long long part_ncell = -1;
  std::vector<int> h_att, h_sub, h_fail;
  std::vector<float> h_forced;
};
State g;

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

int ensure_partition(long long ncell) {
  if (g.part_ncell == ncell) return 0;
  int nob = partition(64, ncell, g.part_sizes);
  if (nob < 0) return fail(-2, "empty grid");
  for (int s : g.part_sizes)
    if (s > 128) return fail(-3, "block of %d cells exceeds 128", s);
  std::vector<int> start(nob + 1, 0);
  for (int b = 0; b < nob; ++b) start[b + 1] = start[b] + g.part_sizes[b];
  HIP_TRY(g.block_start.ensure(sizeof(int) * (nob + 1)));
  HIP_TRY(hipMemcpy(g.block_start.p, start.data(), sizeof(int) * (nob + 1), hipMemcpyHostToDevice));
  HIP_TRY(g.attempts.ensure(sizeof(int) * nob));
  HIP_TRY(g.substeps.ensure(sizeof(int) * nob));
  HIP_TRY(g.failed.ensure(sizeof(int) * nob));
  HIP_TRY(g.forced.ensure(sizeof(float) * nob));
  g.part_ncell = ncell;
  g.blk_hist = false;
  g.hist_cells = -1;
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
                     g.attempts.as<int>(), b0, n, g.sort_keys.as<int>(), g.sort_ids.as<int>());
  HIP_TRY(hipGetLastError());
  return sort_desc(g.sort_keys.as<int>(), g.sort_ids.as<int>(), n, g.order.as<int>(), st);
}

// Launch the blocks of one grid; p has every pointer filled except the
// per-call bookkeeping arrays, which this sets.
int launch_step(brams::StepParams p, hipStream_t st, int nob) {
  p.block_start = g.block_start.as<int>();
  p.mask = g.mask.as<unsigned>();
  p.blk_attempts = g.attempts.as<int>();
  p.blk_substeps = g.substeps.as<int>();
  p.blk_failed = g.failed.as<int>();
  p.blk_forced = g.forced.as<float>();
  const std::vector<int>& sz = g.part_sizes;
  int nbig = 0;
  while (nbig < nob && sz[nbig] > 64) ++nbig;
  const int first64 = nbig > 1 ? nbig : 1;
  // the 64-cell blocks run longest first once a call on this partition has
  // counted their attempts (the order changes no block's result)
  brams::StepParams pm = p;
  if (nob > first64 && g.blk_hist) {
    int rc = order_blocks(first64, nob - first64, st);
    if (rc) return rc;
    pm.block_order = g.order.as<int>();
  }
  HIP_TRY(hipEventRecord(g.ev0, st));
  HIP_TRY(brams::launch_rodas3(sz[0] > 64 ? 128 : 64, true, 1, p, 0, st));
  HIP_TRY(hipEventRecord(g.evb, st));
  if (nbig > 1) HIP_TRY(brams::launch_rodas3(128, false, (unsigned)(nbig - 1), p, 1, st));
  HIP_TRY(hipEventRecord(g.evm, st));
  g.main_cells = 0;
  if (nob > first64) {
    HIP_TRY(brams::launch_rodas3(64, false, (unsigned)(nob - first64), pm, first64, st));
    for (int b = first64; b < nob; ++b) g.main_cells += sz[b];
  }
  HIP_TRY(hipEventRecord(g.ev1, st));
  g.blk_hist = true;
  return 0;
}

int collect_stats(hipStream_t st, int nob, long long ncell) {
  HIP_TRY(hipStreamSynchronize(st));
  g.h_att.resize(nob);
  g.h_sub.resize(nob);
  g.h_fail.resize(nob);
  g.h_forced.resize(nob);
  unsigned mw[MECH_MASK_WORDS];
  HIP_TRY(hipMemcpy(g.h_att.data(), g.attempts.p, sizeof(int) * nob, hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(g.h_sub.data(), g.substeps.p, sizeof(int) * nob, hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(g.h_fail.data(), g.failed.p, sizeof(int) * nob, hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(g.h_forced.data(), g.forced.p, sizeof(float) * nob, hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(mw, g.mask.p, sizeof mw, hipMemcpyDeviceToHost));
  brams_chem_stats s{};
  s.cells = ncell;
  s.blocks = nob;
  for (int b = 0; b < nob; ++b) {
    s.attempts += g.h_att[b];
    s.substeps += g.h_sub[b];
    s.failed_blocks += g.h_fail[b];
    s.cell_attempts += (long long)g.h_att[b] * g.part_sizes[b];
    s.cell_substeps += (long long)g.h_sub[b] * g.part_sizes[b];
    if (g.h_forced[b] > 1.0f) ++s.forced_blocks;
    s.max_forced_err = std::max(s.max_forced_err, g.h_forced[b]);
  }
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
  g.stats = s;
  return (int)s.failed_blocks;
}

int check_common(int m1, int m2, int m3, int ia, int iz, int ja, int jz, int ndyn, int split,
                 int photoj, float dtlt) {
  if (m1 < 3 || m2 < 1 || m3 < 1) return fail(-2, "bad grid extents %d %d %d", m1, m2, m3);
  if (ia < 1 || iz > m2 || ia > iz || ja < 1 || jz > m3 || ja > jz)
    return fail(-2, "bad interior bounds ia=%d iz=%d ja=%d jz=%d", ia, iz, ja, jz);
  if (ndyn < 1) return fail(-2, "N_DYN_CHEM must be >= 1");
  if (split < 0 || split > 2) return fail(-2, "unknown split_method %d", split);
  if (photoj == BRAMS_PHOTOJ_LUT)
    return fail(-4, "PhotojMethod LUT is not supported (RELACS_TUV uses FAST-TUV)");
  if (photoj < 0 || photoj > 2) return fail(-2, "unknown PhotojMethod %d", photoj);
  if (!(dtlt > 0.f)) return fail(-2, "dtlt must be positive");
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
  HIP_TRY(g.mask.ensure(sizeof(unsigned) * MECH_MASK_WORDS));
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

}  // extern "C"

namespace {

// validate device arguments and fill the kernel parameters (no launch)
int prepare_device(const brams_chem_dev_args* a, brams::StepParams& p, long long& ncell) {
  if (!a) return fail(-2, "null args");
  if (!g.init) {
    int rc = brams_chem_init(0);
    if (rc) return rc;
  }
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
  ncell = (long long)(a->m1 - 2) * (a->iz - a->ia + 1) * (a->jz - a->ja + 1);
  if ((rc = ensure_partition(ncell))) return rc;
  const int nob = (int)g.part_sizes.size();
  if (a->nob != nob) return fail(-2, "nob %d does not match the AllocIndex partition (%d)", a->nob, nob);
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
  return 0;
}

}  // namespace

extern "C" {

int brams_chem_rodas3_dyndt_device(const brams_chem_dev_args* a) {
  brams::StepParams p;
  long long ncell = 0;
  int rc = prepare_device(a, p, ncell);
  if (rc) return rc;
  const int nob = (int)g.part_sizes.size();
  hipStream_t st = (hipStream_t)a->stream;
  if ((rc = launch_step(p, st, nob))) return rc;
  return collect_stats(st, nob, ncell);
}

// ------------------------------------------------------------------ surrogate
int brams_sur_features_raw(const brams_chem_dev_args* a, float* feat, double floor_c) {
  brams::StepParams p;
  long long ncell = 0;
  int rc = prepare_device(a, p, ncell);
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
  brams::StepParams p;
  long long ncell = 0;
  int rc = prepare_device(a, p, ncell);
  if (rc) return rc;
  if (!m || m->nmembers < 1 || m->nmembers > 4) return fail(-2, "surrogate needs 1..4 members");
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
  if (g.hist_cells != ncell) {
    // fallback memory: each cell starts from its partition block's step
    HIP_TRY(g.cell_dt.ensure(sizeof(double) * ncell));
    hipLaunchKernelGGL(brams::cell_history_init, dim3((unsigned)((ncell + 255) / 256)), dim3(256), 0, st,
                       g.block_start.as<int>(), (int)a->nob, p.last_dt, ncell, g.cell_dt.as<double>());
    HIP_TRY(hipGetLastError());
    g.hist_cells = ncell;
  }
  hipEvent_t ev[5];
  for (auto& e : ev) HIP_TRY(hipEventCreate(&e));
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
  long long fb_att = 0, fb_cell_att = 0, fb_cell_sub = 0;
  if (nfb > 0) {
    const int nb = (nfb + 63) / 64;
    int* list = g.sur_list.as<int>();
    hipLaunchKernelGGL(brams::fallback_dt_hint, dim3((unsigned)nb), dim3(64), 0, st, list, nfb,
                       g.cell_dt.as<double>(), (double)(a->dtlt * (float)a->n_dyn_chem),
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
    q.block_start = nullptr;
    HIP_TRY(brams::launch_rodas3_list((unsigned)nb, q, st));
    hipLaunchKernelGGL(brams::cell_history_update, dim3((unsigned)((nfb + 255) / 256)), dim3(256), 0, st, list,
                       nfb, q.last_dt, g.cell_dt.as<double>());
    HIP_TRY(hipGetLastError());
    std::vector<int> att(nb), sub(nb);
    HIP_TRY(hipMemcpyAsync(att.data(), q.blk_attempts, sizeof(int) * nb, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(sub.data(), q.blk_substeps, sizeof(int) * nb, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    for (int b = 0; b < nb; ++b) {
      const long long n = std::min(64, nfb - 64 * b);
      fb_att += att[b];
      fb_cell_att += n * att[b];
      fb_cell_sub += n * sub[b];
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
    float ms;
    HIP_TRY(hipEventElapsedTime(&ms, ev[0], ev[1])); s.features_ms = ms;
    HIP_TRY(hipEventElapsedTime(&ms, ev[1], ev[2])); s.mlp_ms = ms;
    HIP_TRY(hipEventElapsedTime(&ms, ev[2], ev[3])); s.apply_ms = ms;
    HIP_TRY(hipEventElapsedTime(&ms, ev[3], ev[4])); s.fallback_ms = ms;
    HIP_TRY(hipEventElapsedTime(&ms, ev[0], ev[4])); s.total_ms = ms;
    *out = s;
  }
  for (auto& e : ev) HIP_TRY(hipEventDestroy(e));
  return 0;
}

void brams_sur_reset(void) { g.hist_cells = -1; }

int brams_chem_rodas3_dyndt(const brams_chem_args* a) {
  if (!a) return fail(-2, "null args");
  if (!g.init) {
    int rc = brams_chem_init(0);
    if (rc) return rc;
  }
  int rc = check_common(a->m1, a->m2, a->m3, a->ia, a->iz, a->ja, a->jz, a->n_dyn_chem,
                        a->split_method, a->photoj_method, a->dtlt);
  if (rc) return rc;
  if (a->nspecies != MECH_NS || a->nr != MECH_NR || a->nr_photo != MECH_NPHOT)
    return fail(-2, "mechanism sizes %d/%d/%d differ from RELACS_TUV %d/%d/%d", a->nspecies, a->nr,
                a->nr_photo, MECH_NS, MECH_NR, MECH_NPHOT);
  const long long ncell = (long long)(a->m1 - 2) * (a->iz - a->ia + 1) * (a->jz - a->ja + 1);
  if ((rc = ensure_partition(ncell))) return rc;
  const int nob = (int)g.part_sizes.size();
  if (a->nob != nob) return fail(-2, "nob %d does not match the AllocIndex partition (%d)", a->nob, nob);
  for (int b = 0; b < nob; ++b)
    if (a->block_end[b] != g.part_sizes[b])
      return fail(-2, "block_end(%d)=%d differs from AllocIndex (%d)", b + 1, a->block_end[b],
                  g.part_sizes[b]);
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
  // consts: weight[47] f32 | transported[47] i32 | atol[47] f64 | rtol[47] f64
  const size_t coff_w = 0, coff_t = 256, coff_a = 512, coff_r = 1024;
  HIP_TRY(g.consts.ensure(1536));
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
  char cbuf[1536];
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
  if ((rc = launch_step(p, st, nob))) return rc;
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
  return collect_stats(st, nob, ncell);
}

int brams_chem_block_status(int32_t* attempts, int32_t* substeps, int32_t* failed,
                            float* forced_err, int32_t max_nob) {
  const int nob = (int)g.h_att.size();
  if (nob > max_nob) return fail(-2, "max_nob %d < nob %d", max_nob, nob);
  for (int b = 0; b < nob; ++b) {
    if (attempts) attempts[b] = g.h_att[b];
    if (substeps) substeps[b] = g.h_sub[b];
    if (failed) failed[b] = g.h_fail[b];
    if (forced_err) forced_err[b] = g.h_forced[b];
  }
  return nob;
}

int brams_chem_get_stats(brams_chem_stats* out) {
  if (!out) return fail(-2, "null stats");
  *out = g.stats;
  return 0;
}

const char* brams_chem_last_error(void) { return g_err.c_str(); }

void brams_chem_finalize(void) {
  DevBuf* bufs[] = {&g.block_start, &g.mask, &g.attempts, &g.substeps, &g.failed, &g.forced, &g.sc_p,
                    &g.sc_t, &g.sc_t_dyn, &g.met, &g.jphoto, &g.last_dt, &g.consts,
                    &g.sur_X, &g.sur_out, &g.sur_flag, &g.sur_list, &g.sur_chunk,
                    &g.sur_reason, &g.fb_dt, &g.fb_att, &g.mask_all, &g.sort_keys, &g.sort_ids,
                    &g.sort_count, &g.sort_off, &g.order, &g.cell_dt};
  for (DevBuf* b : bufs) b->release();
  hipEvent_t* evs[] = {&g.ev0, &g.ev1, &g.evb, &g.evm};
  for (hipEvent_t* e : evs) {
    if (*e) hipEventDestroy(*e);
    *e = nullptr;
  }
  g.part_ncell = -1;
  g.blk_hist = false;
  g.hist_cells = -1;
  g.init = false;
}

}  // extern "C"
