// Launch wrappers of the slab kernels (rodas3_slab.h): the layout variants
// the autotuner chooses among, the block-1 form and the list form of the
// surrogate's fallback.  Instantiated in this translation unit only.
#include <hip/hip_runtime.h>

#include "rodas3_slab_launch.h"
#include "rodas3_slab.h"

namespace brams {

#ifdef BRAMS_MECH_LEAN
// a lean library (rodas3_launch.h) builds no slab kernel
int slab_variant_nlds(int v) { return (v >= 0 && v < kSlabVariants) ? 0 : -1; }
int slab_resident_per_cu(int) { return 0; }
hipError_t launch_rodas3_slab(int, unsigned, const StepParams&, const QueueParams&, hipStream_t) {
  return hipErrorInvalidValue;
}
#else

namespace {

// LU slots held in LDS per partition variant.  With 24 KB of cell state in
// LDS as well, 32 slots keep a workgroup within 40 KB (four per CU, one per
// SIMD, as the 512-register budget allows); 96 slots take 73 KB (two per
// CU); 0 leaves the whole LU in the slab.
constexpr int kVariantNlds[kSlabVariants] = {32, 96, 0};

template <int NLDS, bool LIST, bool DYN>
hipError_t launch_one(unsigned nslots, const StepParams& p, const QueueParams& q, hipStream_t st) {
  hipLaunchKernelGGL(queue_reset, dim3(1), dim3(64), 0, st, q.counter);
  hipLaunchKernelGGL((rodas3_slab<NLDS, LIST, DYN>), dim3(nslots), dim3(64), 0, st, p, q);
  return hipGetLastError();
}

template <int NLDS, bool LIST, bool DYN>
int resident_per_cu() {
  int n = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, rodas3_slab<NLDS, LIST, DYN>, 64, 0) != hipSuccess)
    return 0;
  return n;
}

}  // namespace

int slab_variant_nlds(int v) { return (v >= 0 && v < kSlabVariants) ? kVariantNlds[v] : -1; }

int slab_resident_per_cu(int v) {
  switch (v) {
    case 0: return resident_per_cu<32, false, false>();
    case 1: return resident_per_cu<96, false, false>();
    case 2: return resident_per_cu<0, false, false>();
    case kSlabList: return resident_per_cu<32, true, false>();
    default: return 0;
  }
}

hipError_t launch_rodas3_slab(int variant, unsigned nslots, const StepParams& p, const QueueParams& q,
                              hipStream_t st) {
  if (q.nitems <= 0) return hipSuccess;
  if (nslots == 0) return hipErrorInvalidValue;
  switch (variant) {
    case 0: return launch_one<32, false, false>(nslots, p, q, st);
    case 1: return launch_one<96, false, false>(nslots, p, q, st);
    case 2: return launch_one<0, false, false>(nslots, p, q, st);
    case kSlabList:
      if (!p.cell_list) return hipErrorInvalidValue;
      return launch_one<32, true, false>(nslots, p, q, st);
    case kSlabBlock1:
      return launch_one<32, false, true>(1, p, q, st);
    default: return hipErrorInvalidValue;
  }
}

#endif  // BRAMS_MECH_LEAN

}  // namespace brams
