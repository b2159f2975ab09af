// Kernels of the surrogate chemistry step (no reference counterpart).
//
// Cells are addressed by their global index g in the reference's j -> i -> k
// order (chem_spack_utils.f90:144-160), the same order the RODAS3 kernels
// use; the state is gathered from and scattered to the Fortran (k,i,j)
// arrays with the unit conversions of chem_spack_rodas3_dyndt.f90:226-321 and
// 713-798 (SYMMETRIC/SEQUENTIAL write-back).
//
// Features of one cell (68 for RELACS_TUV): log10 of the MECH_NS
// concentrations (molec/cm^3, floored), T, ln p, log10 rv, the MECH_NPHOT
// photolysis rates, log10 of the step.  Target (MECH_NS): log10 increment of
// each concentration over the step.
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>
#include <stdint.h>

#include "rodas3_kernel.h"

namespace brams {

constexpr int kSurNin = MECH_NS + 3 + MECH_NPHOT + 1;
constexpr int kSurNinPad = 128;
constexpr int kSurNout = MECH_NS;
// member outputs are padded to 64 columns; a mechanism with more species has
// no surrogate (the step refuses every model, whose sizes cannot match)
constexpr int kSurNoutPad = 64;
constexpr bool kSurrogateBuilt = kSurNout <= kSurNoutPad && kSurNin <= kSurNinPad;

struct CellState {
  long long o;        // offset in the (k,i,j) arrays
  double temp, press, vapp, volmol, volmol_i;
};

__device__ __forceinline__ CellState cell_state(const StepParams& p, long long g) {
  CellState c;
  const long long kk = 2 + g % p.nk;
  const long long t = g / p.nk;
  const long long ii = p.ia + t % p.ni;
  const long long jj = p.ja + t / p.ni;
  c.o = (kk - 1) + (long long)p.m1 * (ii - 1) + (long long)p.m1 * p.m2 * (jj - 1);
  const float exner = (p.pp[c.o] + p.pi0[c.o]) / p.cp;
  double press = (double)exner;
  c.temp = (double)p.theta[c.o] * press;
  c.press = pow(press, (double)p.cpor) * (double)p.p00;
  c.vapp = (double)p.rv[c.o];
  c.volmol = (6.02e23 * 1e-15 * 28.96) * c.press / (8.314 * c.temp);
  c.volmol_i = 1.0 / c.volmol;
  return c;
}

// concentration of species s in molec/cm^3 from a species-major state array
__device__ __forceinline__ double cell_conc(const StepParams& p, const float* sc, const CellState& c,
                                            int s) {
  const float v = sc[(long long)s * p.n3d + c.o];
  if (!p.transported[s]) return (double)v;
  const double x = (double)v * c.volmol / (double)p.weight[s];
  return x > 0.0 ? x : 0.0;
}

// raw (unnormalised) features, one thread per cell: feat[g][kSurNin]
__global__ void sur_raw_features(StepParams p, long long ncell, double floor_c, float* feat) {
  const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= ncell) return;
  const CellState c = cell_state(p, g);
  float* f = feat + g * kSurNin;
  for (int s = 0; s < MECH_NS; ++s) f[s] = (float)log10(fmax(cell_conc(p, p.sc_p, c, s), floor_c));
  f[MECH_NS] = (float)c.temp;
  f[MECH_NS + 1] = (float)log(c.press);
  f[MECH_NS + 2] = (float)log10(fmax(c.vapp, 1e-12));
  for (int n = 0; n < MECH_NPHOT; ++n) f[MECH_NS + 3 + n] = (float)p.jphoto[n + (long long)MECH_NPHOT * c.o];
  f[kSurNin - 1] = (float)log10((double)(p.dtlt * (float)p.ndyn));
}

// raw targets: log10 increment between two states of the same cells
__global__ void sur_raw_targets(StepParams p, long long ncell, double floor_c, const float* before,
                                float* targ) {
  const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= ncell) return;
  const CellState c = cell_state(p, g);
  float* t = targ + g * kSurNout;
  for (int s = 0; s < MECH_NS; ++s) {
    const double y0 = fmax(cell_conc(p, before, c, s), floor_c);
    const double y1 = fmax(cell_conc(p, p.sc_p, c, s), floor_c);
    t[s] = (float)(log10(y1) - log10(y0));
  }
}

// per-column sum and sum of squares (FP64), rows split over blocks
__global__ __launch_bounds__(256) void column_moments(const float* X, long long n, int ncol,
                                                      double* sum, double* sumsq) {
  const int col = threadIdx.x % 64;
  const int ph = threadIdx.x / 64;
  for (int c0 = 0; c0 < ncol; c0 += 64) {
    const int c = c0 + col;
    double s = 0.0, q = 0.0;
    if (c < ncol)
      for (long long r = (long long)blockIdx.x * 4 + ph; r < n; r += (long long)gridDim.x * 4) {
        const double v = X[r * ncol + c];
        s += v;
        q += v * v;
      }
    if (c < ncol) {
      atomicAdd(&sum[c], s);
      atomicAdd(&sumsq[c], q);
    }
  }
}

// z-score columns into a padded row: bf16 (features) or f32 (targets).
// Grid-stride: a dispatch's size in work-items is a 32-bit quantity, and a
// dataset of 37.7M rows x 128 columns (4.8e9 elements) would exceed it.
template <bool BF16>
__global__ void normalize_rows(const float* raw, long long n, int ncol, const float* mu,
                               const float* sd, void* out, int ldo) {
  const long long tot = n * ldo;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < tot;
       i += (long long)gridDim.x * blockDim.x) {
    const long long r = i / ldo;
    const int c = (int)(i % ldo);
    const float v = c < ncol ? (raw[r * ncol + c] - mu[c]) / sd[c] : 0.0f;
    if constexpr (BF16)
      static_cast<__bf16*>(out)[i] = (__bf16)v;
    else
      static_cast<float*>(out)[i] = v;
  }
}

// normalised features straight from the state (inference): X[g][128] bf16
__global__ __launch_bounds__(256) void sur_features(StepParams p, long long ncell, long long nrows, double floor_c,
                                                   const float* mu, const float* sd, __bf16* X) {
  const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= nrows) return;
  __bf16* x = X + g * kSurNinPad;
  if (g >= ncell) {                          // padding rows
    for (int k = 0; k < kSurNinPad; ++k) x[k] = (__bf16)0.0f;
    return;
  }
  const CellState c = cell_state(p, g);
  float f[kSurNin];
#pragma unroll
  for (int s = 0; s < MECH_NS; ++s) f[s] = (float)log10(fmax(cell_conc(p, p.sc_p, c, s), floor_c));
  f[MECH_NS] = (float)c.temp;
  f[MECH_NS + 1] = (float)log(c.press);
  f[MECH_NS + 2] = (float)log10(fmax(c.vapp, 1e-12));
#pragma unroll
  for (int n = 0; n < MECH_NPHOT; ++n) f[MECH_NS + 3 + n] = (float)p.jphoto[n + (long long)MECH_NPHOT * c.o];
  f[kSurNin - 1] = (float)log10((double)(p.dtlt * (float)p.ndyn));
#pragma unroll
  for (int k = 0; k < kSurNin; ++k) x[k] = (__bf16)((f[k] - mu[k]) / sd[k]);
#pragma unroll
  for (int k = kSurNin; k < kSurNinPad; ++k) x[k] = (__bf16)0.0f;
}

// Deterministic compaction of flagged cells, ascending cell index:
// count per 1024-cell chunk, scan of the chunk counts, ordered write.
__global__ __launch_bounds__(256) void compact_count(const int* flag, long long n, int* chunk_count) {
  __shared__ int sh[4];
  const long long base = (long long)blockIdx.x * 1024;
  int cnt = 0;
  for (int k = 0; k < 4; ++k) {
    const long long i = base + k * 256 + threadIdx.x;
    cnt += (i < n && flag[i]) ? 1 : 0;
  }
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) cnt += __shfl_xor(cnt, off, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = cnt;
  __syncthreads();
  if (threadIdx.x == 0) chunk_count[blockIdx.x] = sh[0] + sh[1] + sh[2] + sh[3];
}

// exclusive scan of nchunk counts by one workgroup; total in off[nchunk]
__global__ __launch_bounds__(256) void compact_scan(const int* cnt, int nchunk, int* off) {
  __shared__ int carry;
  __shared__ int wsum[4];
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (int b0 = 0; b0 < nchunk; b0 += 256) {
    const int i = b0 + threadIdx.x;
    const int v = i < nchunk ? cnt[i] : 0;
    // inclusive scan inside the wave
    int x = v;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const int y = __shfl_up(x, d, 64);
      if (lane >= d) x += y;
    }
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    int before = carry;
    for (int k = 0; k < w; ++k) before += wsum[k];
    if (i < nchunk) off[i] = before + x - v;
    __syncthreads();
    if (threadIdx.x == 0) carry += wsum[0] + wsum[1] + wsum[2] + wsum[3];
    __syncthreads();
  }
  if (threadIdx.x == 0) off[nchunk] = carry;
}

__global__ __launch_bounds__(256) void compact_write(const int* flag, long long n, const int* off,
                                                     int* list) {
  __shared__ int wsum[4];
  const long long base = (long long)blockIdx.x * 1024;
  int pos = off[blockIdx.x];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int k = 0; k < 4; ++k) {
    const long long i = base + k * 256 + threadIdx.x;
    const int f = (i < n && flag[i]) ? 1 : 0;
    int x = f;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const int y = __shfl_up(x, d, 64);
      if (lane >= d) x += y;
    }
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    int before = pos;
    for (int q = 0; q < w; ++q) before += wsum[q];
    if (f) list[before + x - 1] = (int)i;
    const int tot = wsum[0] + wsum[1] + wsum[2] + wsum[3];
    __syncthreads();
    pos += tot;
  }
}

struct SurApplyParams {
  const float* out[4];        // member outputs [64][nrows] (normalised increments)
  long long nrows;
  int nmembers;
  const float* tmu;           // [MECH_NS] target mean / std
  const float* tsd;
  const float* natoms;        // [MECH_NS] nitrogen atoms per molecule
  const float* catoms;        // [MECH_NS] carbon atoms per molecule
  double floor_c;             // concentration floor (molec/cm^3) of the log features
  double sig_c;               // species below this (before and after) skip the spread check
  float spread_tol;           // max relative ensemble spread (e.g. 0.02)
  float n_tol, c_tol;         // max relative N / C budget change
  int* flag;                  // [ncell] 1 = rejected (integrated by RODAS3)
  int* reason;                // [4] counts: non-finite, budget N, budget C, spread
};

// Ensemble mean of the predicted log10 concentrations, acceptance checks,
// write-back of accepted cells; rejected cells are flagged and keep their
// state for the RODAS3 kernels.
// predicted concentration of species s (molec/cm^3) and its ensemble spread
__device__ __forceinline__ double sur_predict(const StepParams& p, const SurApplyParams& a, const CellState& c,
                                             long long g, int s, double& y0, float& spread) {
  y0 = cell_conc(p, p.sc_p, c, s);
  float lm[4];
  float mean = 0.f;
#pragma unroll
  for (int m = 0; m < 4; ++m) {
    lm[m] = m < a.nmembers ? a.out[m][s * a.nrows + g] * a.tsd[s] + a.tmu[s] : 0.f;
    mean += lm[m];
  }
  mean /= (float)a.nmembers;
  spread = 0.f;
#pragma unroll
  for (int m = 0; m < 4; ++m)
    if (m < a.nmembers) spread = fmaxf(spread, fabsf(lm[m] - mean));
  if (a.tsd[s] == 0.0f) return y0;              // species the mechanism never changes
  // 10**(log10 max(y0, floor) + mean), as a multiplicative update
  return fmax(y0, a.floor_c) * (double)exp10f(mean);
}

// Two passes over the species keep no per-species arrays: the checks, then
// (accepted cells only) the prediction again and the write-back.
__global__ __launch_bounds__(128) void sur_apply(StepParams p, SurApplyParams a, long long ncell) {
  const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= ncell) return;
  const CellState c = cell_state(p, g);
  const float log_spread = log10f(1.0f + a.spread_tol);
  int bad_f = 0, bad_s = 0;
  double n0 = 0.0, n1 = 0.0, c0 = 0.0, c1 = 0.0;
  for (int s = 0; s < MECH_NS; ++s) {
    double y0;
    float spread;
    const double v = sur_predict(p, a, c, g, s, y0, spread);
    if (!isfinite(v) || v < 0.0) bad_f = 1;
    if (spread > log_spread && (v > a.sig_c || y0 > a.sig_c)) bad_s = 1;
    n0 += a.natoms[s] * y0;
    n1 += a.natoms[s] * v;
    c0 += a.catoms[s] * y0;
    c1 += a.catoms[s] * v;
  }
  const int bad_n = fabs(n1 - n0) > a.n_tol * n0;
  const int bad_c = fabs(c1 - c0) > a.c_tol * c0;
  const int why = bad_f ? 0 : bad_n ? 1 : bad_c ? 2 : bad_s ? 3 : -1;
  a.flag[g] = why >= 0 ? 1 : 0;
  // reason counts: one atomic per wave and reason, not one per cell
  const unsigned long long active = __ballot(1);
  const int leader = __ffsll((long long)active) - 1;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const unsigned long long m = __ballot(why == r);
    if (m && (int)(threadIdx.x & 63) == leader) atomicAdd(&a.reason[r], __popcll(m));
  }
  if (why >= 0) return;
  for (int s = 0; s < MECH_NS; ++s) {
    double y0;
    float spread;
    const double v = sur_predict(p, a, c, g, s, y0, spread);
    const float w = p.transported[s] ? (float)(v * (double)p.weight[s] * c.volmol_i) : (float)v;
    p.sc_p[(long long)s * p.n3d + c.o] = w > 0.0f ? w : 0.0f;
  }
}

}  // namespace brams
