// FP64 RODAS3 for 64-cell blocks with explicit placement of the per-cell
// working set ("slab kernel").
//
// The per-lane kernel of rodas3_kernel.h keeps a cell's whole working set
// (LU 534, rate constants 128, stage vectors) as locals, about 1000 doubles,
// which the compiler spills to scratch however the registers are budgeted.
// Here each piece has a fixed home, chosen by its reuse:
//
//   registers  the stage vectors k1, k1-k2, k3, the stage state and f(stage)
//              (at most four 47-vectors live), and the active entries of
//              the factorisation: mech_factor_rl eliminates right-looking
//              with lazy assembly, so an entry is a register from its first
//              update until it is final and at most 169 are live at once;
//   LDS        the cell state y (47 doubles per lane, read by the Jacobian,
//              the rate terms and every stage) and the first NLDS slots of
//              the LU;
//   slab       a lane-interleaved scratch area per workgroup in HBM
//              (element s of lane l at slab[s * 64 + l], so a wave's access
//              is one coalesced 512-byte line): the rest of the LU (written
//              once per attempt, read by the four solves), the Jacobian
//              entries of the sub-step (written once per sub-step, read by
//              every attempt's factorisation), the rate constants, f(y0)
//              and the stage-4 right-hand-side part.
//
// The workgroups are persistent: a grid of slots (one slab each) pops block
// indices from a device counter, in the order given (longest first from the
// previous call's attempts, see sched_kernels.h), until the queue is empty.
// Every block's arithmetic is that of rodas3_blocks operation for operation
// (the generated functions of relacs_tuv_slab.h are checked bit for bit
// against those of relacs_tuv_gen.h), so the results are bit-identical and
// independent of which slot integrates which block.
#pragma once
#include <hip/hip_runtime.h>

#include "mechanism.h"
#include BRAMS_MECH_SLAB
#include "rodas3_kernel.h"
#include "rodas3_slab_launch.h"

// Passes a value through an SGPR by an empty asm, so that the compiler
// cannot hoist what is computed from it out of the enclosing loop (host
// emulations of this code for CPU tests define it away).
#ifndef BRAMS_OPAQUE_SGPR
#define BRAMS_OPAQUE_SGPR(x) asm volatile("" : "+s"(x))
#endif

namespace brams {

// A lane's view of a 64-lane-interleaved array in LDS: element s at
// p[s * 64], a ds access with the offset in its immediate field.
struct Lanes {
  double* p;
  __device__ double& operator[](int s) const { return p[s * 64]; }
};

// The same in a slab region, through a buffer resource: element s of lane l
// at byte l * 8 + s * 512.  The per-slot offset is a constant (soffset and
// the immediate), the lane offset one VGPR, the region base four SGPRs, so
// that no per-slot address is ever kept in registers; plain pointers made
// the compiler hoist some 250 distinct 64-bit bases out of the attempt loop
// and spill them.
typedef unsigned bv2 __attribute__((ext_vector_type(2)));
struct BufLanes {
  __amdgpu_buffer_rsrc_t r;
  int lane8;
  struct Ref {
    const BufLanes* b;
    int s;
    __device__ operator double() const {
      return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(b->r, b->lane8, s * 512, 0));
    }
    __device__ const Ref& operator=(double v) const {
      __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(bv2, v), b->r, b->lane8, s * 512, 0);
      return *this;
    }
    __device__ const Ref& operator=(const Ref& o) const { return *this = (double)o; }
  };
  __device__ Ref operator[](int s) const { return Ref{this, s}; }
};

// buffer resource over n bytes at p (raw, untyped accesses)
__device__ __forceinline__ __amdgpu_buffer_rsrc_t slab_rsrc(double* p, long long nbytes) {
  return __builtin_amdgcn_make_buffer_rsrc(p, (short)0, (int)nbytes, 0x00020000);
}

// LU slots: the first NLDS in LDS, the others in the slab
template <int NLDS>
struct LuLanes {
  double* lds;
  BufLanes glob;
  struct Ref {
    const LuLanes* l;
    int s;
    __device__ operator double() const { return s < NLDS ? l->lds[s * 64] : (double)l->glob[s]; }
    __device__ const Ref& operator=(double v) const {
      if (s < NLDS)
        l->lds[s * 64] = v;
      else
        l->glob[s] = v;
      return *this;
    }
  };
  __device__ Ref operator[](int s) const { return Ref{this, s}; }
};


// One 64-cell block, rodas3_blocks<64, DYN, LIST, *> on the slab layout.
// DYN: block 1 of the call, whose value-based mask is taken from cell 1
// (lane 0) on every attempt and published at the end (rodas3:470-506): the
// off-diagonal bits once per sub-step from its Jacobian entries (they do not
// depend on the step size), the diagonal bits per attempt from
// 1/(gamma h) - J_ii; every lane evaluates them and lane 0's words are
// broadcast, so no lane diverges.
template <int NLDS, bool LIST, bool DYN>
__device__ __forceinline__ void integrate_slab(const StepParams& p, int b, Lanes ys, const LuLanes<NLDS>& lu,
                                               const BufLanes& jac, const BufLanes& rkx,
                                               const BufLanes& f0, const BufLanes& p4) {
  int start, ncell;
  if constexpr (LIST) {
    start = b * 64;
    ncell = min(64, p.list_n - start);
  } else {
    start = p.block_start[b];
    ncell = p.block_start[b + 1] - start;
  }
  const int c = threadIdx.x;
  const bool act = c < ncell;
  const long long g = LIST ? (long long)p.cell_list[start + (act ? c : 0)]
                           : (long long)(start + (act ? c : 0));   // idle lanes shadow cell 1

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

  for (int s = 0; s < MECH_NS; ++s) {
    const float v = p.sc_p[(long long)s * p.n3d + o];
    double x;
    if (!p.transported[s]) {
      x = (double)v;
    } else if (par_n) {
      const float st = v - p.sc_t_dyn[(long long)s * p.n3d + o] * (float)p.ndyn * p.dtlt;
      x = (double)st * volmol / (double)p.weight[s];
    } else {
      const double z = (double)v * volmol / (double)p.weight[s];
      x = z > 0.0 ? z : 0.0;
    }
    ys[s] = x;
  }
  {
    double J[MECH_NPHOT];
#pragma unroll
    for (int n = 0; n < MECH_NPHOT; ++n) J[n] = p.jphoto[n + (long long)MECH_NPHOT * o];
    mech_kinetic_t(temp, press, vapp, J, rkx);
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
  unsigned mw[MECH_MASK_WORDS], moff[MECH_MASK_WORDS];
#pragma unroll
  for (int w = 0; w < MECH_MASK_WORDS; ++w) mw[w] = DYN ? 0u : p.mask[w];

  while (time_c + 1.0e-8 < time_f) {
    // The mask words are made opaque here and before every factorisation:
    // the hundreds of conditions derived from them are loop-invariant, and
    // hoisted out of the loops they would each hold an SGPR pair for the
    // whole block (the compiler then spills them).
#pragma unroll
    for (int w = 0; w < MECH_MASK_WORDS; ++w) {
      mw[w] = __builtin_amdgcn_readfirstlane(mw[w]);
      BRAMS_OPAQUE_SGPR(mw[w]);
    }
    // J and f(y0) once per sub-step: rejections reuse them (rodas3:424-457)
    if constexpr (DYN) {
      unsigned mine[MECH_MASK_WORDS];
#pragma unroll
      for (int w = 0; w < MECH_MASK_WORDS; ++w) mine[w] = 0xffffffffu;
      mech_jac_t(rkx, ys, mine, jac);            // unmasked
      mech_mask_offdiag(jac, mine);
#pragma unroll
      for (int w = 0; w < MECH_MASK_WORDS; ++w) moff[w] = __shfl(mine[w], 0, 64);
#pragma unroll
      for (int w = 0; w < MECH_MASK_WORDS; ++w) mw[w] = moff[w];
    }
    mech_jac_t(rkx, ys, mw, jac);
      __builtin_amdgcn_sched_barrier(0);
    mech_fun_t(rkx, ys, f0);
      __builtin_amdgcn_sched_barrier(0);
    for (;;) {
      ++natt;
      const double igdt = 1.0 / (0.5 * dt);
      if constexpr (DYN) {
        unsigned mine[MECH_MASK_WORDS];
        mech_mask_diag(jac, igdt, mine);
#pragma unroll
        for (int w = 0; w < MECH_MASK_WORDS; ++w) mw[w] = moff[w] | __shfl(mine[w], 0, 64);
      }
#pragma unroll
      for (int w = 0; w < MECH_MASK_WORDS; ++w) {
      mw[w] = __builtin_amdgcn_readfirstlane(mw[w]);
      BRAMS_OPAQUE_SGPR(mw[w]);
    }
      const int bad = mech_factor_rl(jac, igdt, mw, lu, &nviol);
      __builtin_amdgcn_sched_barrier(0);
      const double dti = 1.0 / dt;
      // stage 1, rodas3:523-527
      double k1[MECH_NS];
#pragma unroll
      for (int s = 0; s < MECH_NS; ++s) k1[s] = f0[s];
      mech_solve_t(lu, k1);
      __builtin_amdgcn_sched_barrier(0);
      // stage 2, rodas3:534-538
      double k2[MECH_NS];
#pragma unroll
      for (int s = 0; s < MECH_NS; ++s) k2[s] = (4.0 * dti) * k1[s] + f0[s];
      mech_solve_t(lu, k2);
      __builtin_amdgcn_sched_barrier(0);
      // stage 3 with clipping, rodas3:542-566
      double yn[MECH_NS], ft[MECH_NS];
#pragma unroll
      for (int s = 0; s < MECH_NS; ++s) {
        const double y0 = ys[s];
        yn[s] = y0 + 2.0 * k1[s];
        if (yn[s] < 0.0) {
          yn[s] = 0.0;
          k1[s] = 0.5 * (yn[s] - y0);
        }
      }
      mech_fun_t(rkx, yn, ft);
      __builtin_amdgcn_sched_barrier(0);
      double k3[MECH_NS];
#pragma unroll
      for (int s = 0; s < MECH_NS; ++s) {
        k2[s] = k1[s] - k2[s];                    // k1 - k2, reused below
        k3[s] = ft[s] + dti * k2[s];
      }
      mech_solve_t(lu, k3);
      __builtin_amdgcn_sched_barrier(0);
      // stage 4 with clipping, rodas3:573-599; the k1-k2 part of its
      // right-hand side goes to the slab while f(y4) is evaluated
#pragma unroll
      for (int s = 0; s < MECH_NS; ++s) {
        const double y0 = ys[s];
        yn[s] = y0 + 2.0 * k1[s] + k3[s];
        if (yn[s] < 0.0) {
          yn[s] = 0.0;
          k3[s] = (yn[s] - y0) - 2.0 * k1[s];
        }
        p4[s] = dti * (k2[s] - c83 * k3[s]);
      }
      mech_fun_t(rkx, yn, ft);
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int s = 0; s < MECH_NS; ++s) ft[s] = ft[s] + p4[s];
      mech_solve_t(lu, ft);                        // ft now holds k4
      __builtin_amdgcn_sched_barrier(0);
      // solution and error, rodas3:609-648
      double e = 0.0;
#pragma unroll
      for (int s = 0; s < MECH_NS; ++s) {
        const double y0 = ys[s];
        double v = y0 + 2.0 * k1[s] + k3[s] + ft[s];
        v = fmax(v, 0.0);
        yn[s] = v;
        const double tol = p.atol[s] + p.rtol[s] * fmax(fabs(y0), fabs(v));
        const double q = ft[s] / tol;
        e = e + q * q;
      }
      const double err = fmax(1.0e-15, sqrt(e / (double)MECH_NS));
      double errmax = act ? err : 0.0;
      int rej = (act && (err - 1.0e-8 > 1.0)) ? 1 : 0;
      // a zero pivot or a non-finite error sum ends the block (the reference
      // aborts the process there); kMaxAttempts bounds every wave's work
      int fail = (act && (bad || !isfinite(e))) ? 1 : 0;
      if (natt >= kMaxAttempts) fail = 1;
      block_reduce<64>(errmax, rej, fail, nullptr, nullptr, nullptr);
      // controller, rodas3:650-700
      const double fac = fmin(6.0, fmax(0.2, 0.9 / pow(errmax, 1.0 / 3.0)));
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
      for (int s = 0; s < MECH_NS; ++s) ys[s] = yn[s];
      ++nsub;
      break;
    }
  }

  // ---- write-back, rodas3:713-798
  if (act) {
    for (int s = 0; s < MECH_NS; ++s) {
      const long long idx = (long long)s * p.n3d + o;
      const double y = ys[s];
      if (!p.transported[s]) {
        const float v = (float)y;
        p.sc_p[idx] = v > 0.0f ? v : 0.0f;
        continue;
      }
      const double conv = y * (double)p.weight[s] * volmol_i;
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
  int v = act ? nviol : 0;
  if constexpr (DYN) {
    if (c == 0 && p.order_thresh) p.order_thresh[0] = nviol;
  }
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  if (c == 0) {
    if constexpr (DYN) {
#pragma unroll
      for (int w = 0; w < MECH_MASK_WORDS; ++w) p.mask[w] = mw[w];
    }
    p.last_dt[b] = last_dt;
    p.blk_attempts[b] = natt;
    p.blk_substeps[b] = nsub;
    p.blk_failed[b] = failed;
    if (p.blk_forced) p.blk_forced[b] = forced;
    if (p.blk_thresh) p.blk_thresh[b] = v;
  }
}

// Persistent slots popping 64-cell blocks (LIST: 64-entry runs of
// p.cell_list) from q.counter until q.nitems are taken.
template <int NLDS, bool LIST, bool DYN>
__global__ __launch_bounds__(64, 1) void rodas3_slab(StepParams p, QueueParams q) {
  __shared__ double s_y[MECH_NS * 64];
  __shared__ double s_lu[(NLDS > 0 ? NLDS : 1) * 64];
  const int c = threadIdx.x;
  const long long slot = blockIdx.x;
  const Lanes ys{s_y + c};
  const int lane8 = c * 8;
  const LuLanes<NLDS> lu{s_lu + c,
                         BufLanes{slab_rsrc(q.lu + slot * kSlabLuRows * 64, kSlabLuRows * 512LL), lane8}};
  const BufLanes jac{slab_rsrc(q.jac + slot * kSlabJacRows * 64, kSlabJacRows * 512LL), lane8};
  const BufLanes rkx{slab_rsrc(q.rk + slot * kSlabRkRows * 64, kSlabRkRows * 512LL), lane8};
  const BufLanes f0{slab_rsrc(q.vec + slot * kSlabVecRows * 64, MECH_NS * 512LL), lane8};
  const BufLanes p4{slab_rsrc(q.vec + slot * kSlabVecRows * 64 + MECH_NS * 64, MECH_NS * 512LL), lane8};
  for (;;) {
    int item = 0;
    if (c == 0) item = atomicAdd(q.counter, 1);
    item = __shfl(item, 0, 64);
    if (item >= q.nitems) break;
    const int b = q.order ? q.order[item] : q.b0 + item;
    integrate_slab<NLDS, LIST, DYN>(p, b, ys, lu, jac, rkx, f0, p4);
  }
}

// Sets the queue counter to zero (a kernel, not a memset, so that the
// sequence can be captured in a graph).
__global__ void queue_reset(int* counter) {
  if (threadIdx.x == 0) counter[0] = 0;
}

}  // namespace brams
