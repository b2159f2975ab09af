// Host entry of the slab kernels (defined in rodas3_slab.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "mechanism.h"
#include BRAMS_MECH_SLAB

namespace brams {

struct StepParams;

// Slab regions per slot, in 64-lane rows of doubles.  Each region is its
// own allocation, reached through its own kernel argument: the compiler then
// cannot prove two regions apart and keeps loads of one behind earlier
// stores of another, in the order the generated code gives.  (With one base
// pointer it hoisted every Jacobian load of the factorisation to the top and
// spilled.)
constexpr int kSlabLuRows = MECH_NLU;
constexpr int kSlabJacRows = MECH_NJAC;
constexpr int kSlabRkRows = MECH_NR;
constexpr int kSlabVecRows = 2 * MECH_NS;     // f(y0), stage-4 right-hand-side part

struct QueueParams {
  int* counter;          // next item to pop (set to 0 before the launch)
  const int* order;      // block id of each item, or null: item i is block b0 + i
  int b0;
  int nitems;
  // per slot (gridDim.x of them): kSlab*Rows x 64 doubles each
  double* lu;
  double* jac;
  double* rk;
  double* vec;
};


// Slab kernels (rodas3_slab.h).  Variants 0..kSlabVariants-1 integrate the
// 64-cell partition blocks with different LDS layouts; kSlabList is the
// surrogate fallback's list form, kSlabBlock1 the block-1 form (one slot).
constexpr int kSlabVariants = 3;
constexpr int kSlabList = 100;
constexpr int kSlabBlock1 = 101;
// LU slots the variant keeps in LDS (-1: not a partition variant)
int slab_variant_nlds(int variant);
// workgroups of the variant resident per CU (0 on error)
int slab_resident_per_cu(int variant);
// q.counter is reset by a kernel on st before the launch of nslots slots
hipError_t launch_rodas3_slab(int variant, unsigned nslots, const StepParams& p, const QueueParams& q,
                              hipStream_t st);

}  // namespace brams
