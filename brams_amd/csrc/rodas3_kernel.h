// Batched FP64 RODAS3 integration of the RELACS_TUV mechanism on CDNA4.
//
// Mapping: one workgroup = one reference block (chem_spack_utils.f90:93-160),
// one lane = one cell.  Blocks of 64 cells are exactly one wave64; blocks of
// 65..127 cells run as a 128-lane workgroup.  The block-shared step size of
// the reference (rodas3:634-700) is a workgroup reduction of the per-cell
// error, so every lane takes the same accept/reject branch.
//
// State never leaves the lane: gather from the Fortran (k,i,j) arrays, rate
// coefficients, the sub-step loop (Jacobian, masked LU, four stages, error)
// and the write-back all happen in one kernel.  The Jacobian is not stored:
// each attempt rebuilds 1/(gamma h) I - J from y0 and rk inside the
// factorisation, which yields the same values the reference keeps in DLdrdc
// (rejections recompute neither J nor f(y0), rodas3:424-457).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mechanism.h"

namespace brams {

struct StepParams {
  int m1, m2, m3, ia, iz, ja, jz;
  int nk, ni;                    // interior levels (m1-2) and columns in i
  float dtlt;
  int ndyn, split;
  float cp, cpor, p00;
  const float *pp, *pi0, *theta, *rv;
  const double* jphoto;          // (17, m1, m2, m3)
  const float* weight;           // [47]
  const int* transported;        // [47] 1/0
  float* sc_p;                   // [47][n3d]
  float* sc_t;                   // [47][n3d] or null
  const float* sc_t_dyn;         // [47][n3d] or null
  long long n3d;
  const int* block_start;        // [nob+1]
  double* last_dt;               // [nob]
  const double *atol, *rtol;     // [47]
  unsigned* mask;                // [MECH_MASK_WORDS]: block-1 mask, read by later blocks
  int* blk_attempts;             // [nob]
  int* blk_substeps;             // [nob]
  int* blk_failed;               // [nob]
  float* blk_forced;             // [nob] largest error of an acceptance forced at dt_min (0: none)
  int* blk_thresh;               // [nob] or null: pivots failing Sparse1.3's threshold rule, summed
                                 // over the block's cells and attempts
  int* order_thresh;             // [1] or null: the same count for cell 1 of block 1 (DYN kernel),
                                 // the matrix the reference orders on (rodas3:470-506)
  const int* cell_list;          // optional: blocks are consecutive WG-sized runs of this list
  const int* block_order;        // optional: workgroup w integrates block block_order[w]
  int list_n;                    // length of cell_list
};

enum { kSplitSymmetric = 0, kSplitSequential = 1, kSplitParallel = 2 };
// Upper bound on attempts per block and call.  Each rejection shrinks dt by
// at least 10% until dt_min forces acceptance (rodas3:650-676), so a block
// needs far fewer (at most dtlt/dt_min accepted steps plus a few dozen
// rejections each); the bound only guards against a non-terminating wave.
constexpr int kMaxAttempts = 20000;

template <int WG>
__device__ __forceinline__ void block_reduce(double& errmax, int& rej, int& fail,
                                             double* s_err, int* s_rej, int* s_fail) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    errmax = fmax(errmax, __shfl_xor(errmax, off, 64));
    rej |= __shfl_xor(rej, off, 64);
    fail |= __shfl_xor(fail, off, 64);
  }
  if constexpr (WG > 64) {
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
      s_err[wave] = errmax;
      s_rej[wave] = rej;
      s_fail[wave] = fail;
    }
    __syncthreads();
#pragma unroll
    for (int w = 0; w < WG / 64; ++w) {
      errmax = fmax(errmax, s_err[w]);
      rej |= s_rej[w];
      fail |= s_fail[w];
    }
    __syncthreads();
  }
}

// DYN: block 1 of the call.  Its sparsity mask is recomputed from cell 1 on
// every attempt (rodas3:470-482) and the last one is published for the other
// blocks.  Otherwise the mask is read from p.mask.
// A lane's state (about 1000 doubles: LU, rate constants, stage vectors)
// never fits the register file, so the compiler spills whatever is asked
// of it; what matters is how many waves hide the spill traffic.  WPS sets
// the register budget by the waves per SIMD it must allow.  Measured on the
// 1.6M-cell workload (25 000 blocks; cold step / warm step, ms): WPS 1:
// 1294 / 120, 2: 1059 / 102, 3: 998 / 98, 4: 939 / 95, 6: 1023 / 112,
// 8: 1082 / 118, the scratch per lane growing only from 8.0 to 9.2 KB.  A
// launch of a few thousand blocks has too few waves to use the occupancy
// and runs faster with the whole register file.  Crossover measurements
// (WPS 4 / 1, ms): 6002 blocks (98x98x40) 255 / 285; 2300 blocks (the
// surrogate's fallback) 44 / 31; 1562 blocks (50x50x40) 93 / 74.  The
// library times WPS 1, 2 and 4 on the first call of each grid and keeps the
// fastest; launches without a timed choice pick by block count.
constexpr int kRodas3ManyWaves = 4;        // WPS of large launches
constexpr int kRodas3ManyBlocks = 4096;    // from this many blocks on

// The two Rosenbrock solvers of the reference on the same mechanism code:
// RODAS3 (chem_spack_rodas3_dyndt.f90, CHEMISTRY=4) and the two-stage ROS2
// (chem_spack_ros_dyndt.f90, CHEMISTRY=3).
enum { kRodas3 = 0, kRos2 = 1 };
// ROS2's gamma = 1 + 1/sqrt(2), with sqrt(2) written as the double value of
// its single-precision rounding, 1.4142135381698608 (ros_dyndt:130)
constexpr double kRos2Gamma = 1.0 + 1.0 / 1.4142135381698608;
constexpr double kRos2GammaInv = 1.0 / kRos2Gamma;

// LIST: the blocks are consecutive WG-sized runs of p.cell_list (fallback of
// the surrogate step) instead of the AllocIndex partition.
// SOLVER kRos2: chem_ros_dyndt's integration, which uses the full structural
// Jacobian (no value mask, so no block-1 form) and no clipping.
template <int WG, bool DYN, bool LIST, int SOLVER = kRodas3>
__device__ __forceinline__ void integrate_block(const StepParams& p, int b) {
  static_assert(SOLVER == kRodas3 || (!DYN && !LIST), "ROS2 has no block-1 or list form");
  // Sparse1.3 applies its threshold rule only where it orders, on cell 1 of
  // block 1 (rodas3:470-506); the other blocks reuse that order unchecked.
  // Block 1 and the 65..127-cell blocks count failing pivots over all their
  // cells; the 64-cell and list blocks skip the count (it costs them a few
  // percent of their time).
  constexpr bool kCountThreshold = DYN || (WG > 64);
  int start, ncell;
  if constexpr (LIST) {                             // fallback blocks of listed cells
    start = b * WG;
    ncell = min(WG, p.list_n - start);
  } else {
    start = p.block_start[b];
    ncell = p.block_start[b + 1] - start;
  }
  const int c = threadIdx.x;
  const bool act = c < ncell;
  const long long g = LIST ? (long long)p.cell_list[start + (act ? c : 0)]
                                  : (long long)(start + (act ? c : 0));  // idle lanes shadow cell 1

  // cell position, j -> i -> k order (chem_spack_utils.f90:144-160)
  const long long kk = 2 + g % p.nk;
  const long long t = g / p.nk;
  const long long ii = p.ia + t % p.ni;
  const long long jj = p.ja + t / p.ni;
  const long long o = (kk - 1) + (long long)p.m1 * (ii - 1) + (long long)p.m1 * p.m2 * (jj - 1);

  // ---- gather and unit conversion, rodas3:226-354
  const float exner = (p.pp[o] + p.pi0[o]) / p.cp;   // REAL arithmetic
  double press = (double)exner;
  const double temp = (double)p.theta[o] * press;
  press = pow(press, (double)p.cpor) * (double)p.p00;
  const double vapp = (double)p.rv[o];
  const double volmol = (6.02e23 * 1e-15 * 28.96) * press / (8.314 * temp);
  const double volmol_i = 1.0 / volmol;
  const bool par_n = (p.split == kSplitParallel) && (p.ndyn > 1);

  double y[MECH_NS];
#pragma unroll
  for (int s = 0; s < MECH_NS; ++s) {
    const float v = p.sc_p[(long long)s * p.n3d + o];
    if (!p.transported[s]) {
      y[s] = (double)v;
    } else if (par_n) {
      const float st = v - p.sc_t_dyn[(long long)s * p.n3d + o] * (float)p.ndyn * p.dtlt;
      y[s] = (double)st * volmol / (double)p.weight[s];
    } else {
      const double x = (double)v * volmol / (double)p.weight[s];
      y[s] = x > 0.0 ? x : 0.0;
    }
  }
  double rk[MECH_NR];
  {
    double J[MECH_NPHOT];
#pragma unroll
    for (int n = 0; n < MECH_NPHOT; ++n) J[n] = p.jphoto[n + (long long)MECH_NPHOT * o];
    mech_kinetic(temp, press, vapp, J, rk);
  }

  // ---- RODAS3 sub-steps, rodas3:413-705
  const double c83 = 8.0 / 3.0;
  const float tnf = p.dtlt * (float)p.ndyn;
  const double dt_min = fmax(1.0, 1.0e-2 * (double)tnf);
  const double dt_max = (double)tnf;
  const double time_f = (double)tnf;
  double dt = p.last_dt[b];
  double last_dt = dt;
  double time_c = 0.0;
  int natt = 0, nsub = 0, failed = 0, nviol = 0;
  float forced = 0.f;

  __shared__ unsigned s_mask[MECH_MASK_WORDS];
  __shared__ double s_err[WG / 64 + 1];
  __shared__ int s_rej[WG / 64 + 1], s_fail[WG / 64 + 1];
  unsigned mw[MECH_MASK_WORDS];
  if constexpr (SOLVER == kRos2) {
#pragma unroll
    for (int w = 0; w < MECH_MASK_WORDS; ++w) mw[w] = 0xffffffffu;   // every structural entry
  } else if constexpr (!DYN) {
#pragma unroll
    for (int w = 0; w < MECH_MASK_WORDS; ++w) mw[w] = p.mask[w];
  }

  while (time_c + 1.0e-8 < time_f) {
    double f0[MECH_NS];
    mech_fun(rk, y, f0);
    for (;;) {
      ++natt;
      // 1/(gamma h): RODAS3 gamma = 1/2 (rodas3:462), ROS2 ros_dyndt:387
      const double igdt = SOLVER == kRos2 ? 1.0 / (kRos2Gamma * dt) : 1.0 / (0.5 * dt);
      if constexpr (DYN) {
        // every lane evaluates the mask of its own cell, so the evaluation
        // is no divergent region; cell 1's is the one published
        unsigned m[MECH_MASK_WORDS];
        mech_mask(rk, y, igdt, m);
        if (c == 0) {
#pragma unroll
          for (int w = 0; w < MECH_MASK_WORDS; ++w) s_mask[w] = m[w];
        }
        __syncthreads();
#pragma unroll
        for (int w = 0; w < MECH_MASK_WORDS; ++w) mw[w] = s_mask[w];
      }
      double lu[MECH_NLU];
      const int bad = mech_factor<kCountThreshold>(rk, y, igdt, mw, lu, &nviol);
      const double dti = 1.0 / dt;
      double yn[MECH_NS];
      double e = 0.0;
      if constexpr (SOLVER == kRos2) {
        // stage 1, ros_dyndt:404-424: k1 = (1/(gamma h) - J)^-1 f(y0)
        double k1[MECH_NS], k2[MECH_NS];
#pragma unroll
        for (int s = 0; s < MECH_NS; ++s) k1[s] = f0[s];
        mech_solve(lu, k1);
#pragma unroll
        for (int s = 0; s < MECH_NS; ++s) yn[s] = y[s] + kRos2GammaInv * k1[s];
        // stage 2, ros_dyndt:433-460: f(y0 + k1/gamma) - 2/(gamma h) k1
        mech_fun(rk, yn, k2);
#pragma unroll
        for (int s = 0; s < MECH_NS; ++s) k2[s] = k2[s] - 2.0 * kRos2GammaInv * dti * k1[s];
        mech_solve(lu, k2);
        // solution and error, ros_dyndt:463-498 (no clipping)
#pragma unroll
        for (int s = 0; s < MECH_NS; ++s) {
          const double v = y[s] + kRos2GammaInv * (1.5 * k1[s] + 0.5 * k2[s]);
          yn[s] = v;
          const double tol = p.atol[s] + p.rtol[s] * fmax(fabs(y[s]), fabs(v));
          const double q = (0.5 * kRos2GammaInv * (k1[s] + k2[s])) / tol;
          e = e + q * q;
        }
      } else {
      // stage 1, rodas3:523-527
      double k1[MECH_NS];
#pragma unroll
      for (int s = 0; s < MECH_NS; ++s) k1[s] = f0[s];
      mech_solve(lu, k1);
      // stage 2, rodas3:534-538
      double k2[MECH_NS];
#pragma unroll
      for (int s = 0; s < MECH_NS; ++s) k2[s] = (4.0 * dti) * k1[s] + f0[s];
      mech_solve(lu, k2);
      // stage 3 with clipping, rodas3:542-566
      double ft[MECH_NS];
#pragma unroll
      for (int s = 0; s < MECH_NS; ++s) {
        yn[s] = y[s] + 2.0 * k1[s];
        if (yn[s] < 0.0) {
          yn[s] = 0.0;
          k1[s] = 0.5 * (yn[s] - y[s]);
        }
      }
      mech_fun(rk, yn, ft);
      double k3[MECH_NS];
#pragma unroll
      for (int s = 0; s < MECH_NS; ++s) {
        k2[s] = k1[s] - k2[s];                    // k1 - k2, reused below
        k3[s] = ft[s] + dti * k2[s];
      }
      mech_solve(lu, k3);
      // stage 4 with clipping, rodas3:573-599
#pragma unroll
      for (int s = 0; s < MECH_NS; ++s) {
        yn[s] = y[s] + 2.0 * k1[s] + k3[s];
        if (yn[s] < 0.0) {
          yn[s] = 0.0;
          k3[s] = (yn[s] - y[s]) - 2.0 * k1[s];
        }
      }
      mech_fun(rk, yn, ft);
#pragma unroll
      for (int s = 0; s < MECH_NS; ++s) ft[s] = ft[s] + dti * (k2[s] - c83 * k3[s]);
      mech_solve(lu, ft);                          // ft now holds k4
      // solution and error, rodas3:609-648
#pragma unroll
      for (int s = 0; s < MECH_NS; ++s) {
        double v = y[s] + 2.0 * k1[s] + k3[s] + ft[s];
        v = fmax(v, 0.0);
        yn[s] = v;
        const double tol = p.atol[s] + p.rtol[s] * fmax(fabs(y[s]), fabs(v));
        const double q = ft[s] / tol;
        e = e + q * q;
      }
      }
      const double err = fmax(1.0e-15, sqrt(e / (double)MECH_NS));
      double errmax = act ? err : 0.0;
      int rej = (act && (err - 1.0e-8 > 1.0)) ? 1 : 0;
      // a zero pivot or a non-finite error sum ends the block (the reference
      // aborts the process there); kMaxAttempts bounds every wave's work
      int fail = (act && (bad || !isfinite(e))) ? 1 : 0;
      if (natt >= kMaxAttempts) fail = 1;
      block_reduce<WG>(errmax, rej, fail, s_err, s_rej, s_fail);
      // controller, rodas3:650-700 (error order 3) and ros_dyndt:501-551 (order 2)
      const double fac = fmin(6.0, fmax(0.2, 0.9 / pow(errmax, SOLVER == kRos2 ? 1.0 / 2.0 : 1.0 / 3.0)));
      double dt_new = dt * fac;
      dt_new = fmax(dt_min, fmin(dt_max, dt_new));
      if (fail) {
        failed = 1;
        time_c = time_f;
        break;
      }
      if (rej && dt_new > dt_min) {
        dt = dt_new;
        continue;
      }
      // an error above 1 accepted anyway: dt reached dt_min (rodas3:671-676)
      if (rej) forced = fmaxf(forced, (float)errmax);
      time_c = time_c + dt;
      dt = fmin(dt_new, time_f - time_c);
      if (time_c < time_f) last_dt = dt_new;
#pragma unroll
      for (int s = 0; s < MECH_NS; ++s) y[s] = yn[s];
      ++nsub;
      break;
    }
  }

  // ---- write-back, rodas3:713-798
  if (act) {
#pragma unroll
    for (int s = 0; s < MECH_NS; ++s) {
      const long long idx = (long long)s * p.n3d + o;
      if (!p.transported[s]) {
        const float v = (float)y[s];
        p.sc_p[idx] = v > 0.0f ? v : 0.0f;
        continue;
      }
      const double conv = y[s] * (double)p.weight[s] * volmol_i;
      if (par_n) {
        const float base = p.sc_t_dyn[idx] * (float)p.ndyn * p.dtlt;
        const float v = (float)((double)base + conv);
        p.sc_p[idx] = v > 0.0f ? v : 0.0f;
      } else if (p.split == kSplitParallel) {
        const double dti = 1.0 / (double)p.dtlt;
        p.sc_t[idx] = (float)((double)p.sc_t[idx] + (conv - (double)p.sc_p[idx]) * dti);
      } else {
        const float v = (float)conv;
        p.sc_p[idx] = v > 0.0f ? v : 0.0f;
      }
    }
  }
  // threshold-rule count of the block: sum over its cells (0 where not counted)
  if constexpr (!kCountThreshold) {
    if (c == 0 && p.blk_thresh) p.blk_thresh[b] = 0;
  } else {
    int v = act ? nviol : 0;
    if constexpr (DYN) {
      if (c == 0 && p.order_thresh) p.order_thresh[0] = nviol;
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    if constexpr (WG > 64) {
      __syncthreads();
      if ((threadIdx.x & 63) == 0) s_rej[threadIdx.x >> 6] = v;
      __syncthreads();
      v = 0;
#pragma unroll
      for (int w = 0; w < WG / 64; ++w) v += s_rej[w];
    }
    if (c == 0 && p.blk_thresh) p.blk_thresh[b] = v;
  }
  if (c == 0) {
    p.last_dt[b] = last_dt;
    p.blk_attempts[b] = natt;
    p.blk_substeps[b] = nsub;
    p.blk_failed[b] = failed;
    if (p.blk_forced) p.blk_forced[b] = forced;
    if constexpr (DYN) {
#pragma unroll
      for (int w = 0; w < MECH_MASK_WORDS; ++w) p.mask[w] = s_mask[w];
    }
  }
}

template <int WG, bool DYN, bool LIST, int WPS>
__global__ __launch_bounds__(WG, WPS) void rodas3_blocks(StepParams p, int block0) {
  const int b = (!LIST && p.block_order) ? p.block_order[blockIdx.x] : block0 + blockIdx.x;
  integrate_block<WG, DYN, LIST>(p, b);
}

// CHEMISTRY=3: chem_ros_dyndt's blocks, all of one launch alike (no block-1 form)
template <int WG, int WPS>
__global__ __launch_bounds__(WG, WPS) void ros2_blocks(StepParams p, int block0) {
  const int b = p.block_order ? p.block_order[blockIdx.x] : block0 + blockIdx.x;
  integrate_block<WG, false, false, kRos2>(p, b);
}

// Several grids in one launch (the nested grids of one model step, which the
// reference integrates one call per grid, chemistry.f90:313-320): workgroup
// w integrates block (work[w] & kGridBlockMask) of the grid whose parameters
// are grids[work[w] >> kGridShift].  Blocks never exchange data, so a block
// gives the same result here as in its own grid's launch.
constexpr int kGridShift = 24;
constexpr int kGridBlockMask = (1 << kGridShift) - 1;
template <int WG, bool DYN, int WPS>
__global__ __launch_bounds__(WG, WPS) void rodas3_grids(const StepParams* __restrict__ grids,
                                                        const int* __restrict__ work) {
  const int w = work[blockIdx.x];
  integrate_block<WG, DYN, false>(grids[w >> kGridShift], w & kGridBlockMask);
}

}  // namespace brams
