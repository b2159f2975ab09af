// Work ordering for the RODAS3 launches (no reference counterpart; block
// results do not depend on the order blocks run in).
//
// Block costs are very uneven (a warm 1.6M-cell step: median 1 attempt per
// block, p99 24, max 39), so a launch in partition order can end with a few
// expensive blocks running alone.  The plain step therefore launches its
// blocks longest first, by the attempts each block needed in the previous
// call.  The surrogate's fallback keeps a per-cell memory of the last
// accepted step instead of the partition block's.
//
// The sort is a stable counting sort on small keys (0..kSortBins-1, larger
// costs clamped), descending, deterministic: one thread owns a chunk of
// kSortChunk consecutive items, counts it, and scatters it in order after
// a single-workgroup scan of the per-(key, chunk) counts (compact_scan).
#pragma once
#include <hip/hip_runtime.h>

#include "mechanism.h"

namespace brams {

constexpr int kSortBins = 64;
constexpr int kSortChunk = 256;

__device__ __forceinline__ int sort_key(int cost) { return cost < 0 ? 0 : (cost >= kSortBins ? kSortBins - 1 : cost); }

// count[(kSortBins-1-k) * nchunk + chunk]: items of the chunk with key k, laid
// out in the scan order (descending key, ascending chunk).  One wave per 64
// chunks; each thread keeps its chunk's counters in its own LDS column.
__global__ __launch_bounds__(64) void sort_count(const int* __restrict__ keys, int n, int nchunk,
                                                 int* __restrict__ count) {
  __shared__ int h[kSortBins][64];
  const int t = threadIdx.x;
  const int ch = blockIdx.x * 64 + t;
#pragma unroll
  for (int k = 0; k < kSortBins; ++k) h[k][t] = 0;
  if (ch >= nchunk) return;
  const int end = min(n, (ch + 1) * kSortChunk);
  for (int i = ch * kSortChunk; i < end; ++i) h[keys[i]][t] += 1;
#pragma unroll
  for (int k = 0; k < kSortBins; ++k) count[(kSortBins - 1 - k) * nchunk + ch] = h[k][t];
}

// out[off(key, chunk) + rank] = items[i] (items == nullptr: the index i itself)
__global__ __launch_bounds__(64) void sort_scatter(const int* __restrict__ keys, const int* __restrict__ items,
                                                   int n, int nchunk, const int* __restrict__ off,
                                                   int* __restrict__ out) {
  __shared__ int pos[kSortBins][64];
  const int t = threadIdx.x;
  const int ch = blockIdx.x * 64 + t;
  if (ch >= nchunk) return;
#pragma unroll
  for (int k = 0; k < kSortBins; ++k) pos[k][t] = off[(kSortBins - 1 - k) * nchunk + ch];
  const int end = min(n, (ch + 1) * kSortChunk);
  for (int i = ch * kSortChunk; i < end; ++i) {
    const int k = keys[i];
    const int q = pos[k][t];
    pos[k][t] = q + 1;
    out[q] = items ? items[i] : i;
  }
}

// keys of the blocks b0..b0+n-1 from their attempts in the previous call
__global__ void block_keys(const int* __restrict__ attempts, int b0, int n, int* __restrict__ keys,
                           int* __restrict__ ids) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  keys[i] = sort_key(attempts[b0 + i]);
  ids[i] = b0 + i;
}

// per-cell step memory of the fallback, started from the partition blocks'
// last_accepted_dt (the block of cell g found by bisection of block_start)
__global__ void cell_history_init(const int* __restrict__ block_start, int nob, const double* __restrict__ last_dt,
                                  long long ncell, double* __restrict__ cell_dt) {
  const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= ncell) return;
  int lo = 0, hi = nob - 1;                    // largest b with block_start[b] <= g
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (block_start[mid] <= g) lo = mid;
    else hi = mid - 1;
  }
  cell_dt[g] = last_dt[lo];
}

// first step of each fallback block: the smallest remembered step of its cells
__global__ __launch_bounds__(64) void fallback_dt_hint(const int* __restrict__ list, int nfb,
                                                       const double* __restrict__ cell_dt, double dt_max,
                                                       double* __restrict__ fb_dt) {
  const int i = blockIdx.x * 64 + threadIdx.x;
  double d = dt_max;
  if (i < nfb) d = fmin(d, cell_dt[list[i]]);
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) d = fmin(d, __shfl_xor(d, off, 64));
  if (threadIdx.x == 0) fb_dt[blockIdx.x] = d;
}

// after the fallback: every listed cell remembers its block's last accepted step
__global__ void cell_history_update(const int* __restrict__ list, int nfb, const double* __restrict__ fb_dt,
                                    double* __restrict__ cell_dt) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nfb) return;
  cell_dt[list[i]] = fb_dt[i / 64];
}

// AllocIndex block starts in closed form (chem_spack_utils.f90:93-125): the
// remainder ncell - 64*nob is spread one cell at a time over the blocks from
// the first on, so block b holds base + (b < r) cells with base = 64 + the
// whole rounds and r the last partial round.  block_start[0..nob].
__global__ void partition_starts(int nob, int base, int r, int* __restrict__ block_start) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b <= nob) block_start[b] = b * base + (b < r ? b : r);
}

// chem_accum (chemistry.f90:528-604) on the transported species of a
// species-major [nspecies][n3d] pair: ntask 1 adds ni * sc_t to sc_t_dyn,
// otherwise sets sc_t_dyn to zero.  REAL arithmetic with the product and the
// sum rounded separately, as the reference's non-contracted build does.
struct AccumSpecies {
  int n;
  int idx[MECH_NS];          // 0-based species of the transported list
};
__global__ __launch_bounds__(256) void chem_accum(long long n3d, float ni, int accumulate, AccumSpecies sp,
                                                  float* __restrict__ sc_t_dyn, const float* __restrict__ sc_t) {
  const long long base = (long long)sp.idx[blockIdx.y] * n3d;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n3d;
       i += (long long)gridDim.x * blockDim.x) {
    sc_t_dyn[base + i] = accumulate ? __fadd_rn(sc_t_dyn[base + i], __fmul_rn(ni, sc_t[base + i])) : 0.0f;
  }
}

// out[0] = how many of flags[0..n) are nonzero; one workgroup of 256
__global__ __launch_bounds__(256) void count_nonzero(const int* __restrict__ flags, int n,
                                                     int* __restrict__ out) {
  __shared__ int part[4];
  int v = 0;
  for (int i = threadIdx.x; i < n; i += 256) v += flags[i] != 0 ? 1 : 0;
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) out[0] = part[0] + part[1] + part[2] + part[3];
}

}  // namespace brams
