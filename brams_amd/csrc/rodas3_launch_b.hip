// Launch wrappers of the RODAS3 block kernels (rodas3_kernel.h): the
// 128-lane forms (blocks of 65..127 cells), the 64-lane form at two waves
// per SIMD, and the list forms of the surrogate's fallback.
#include <hip/hip_runtime.h>

#include "rodas3_kernel.h"
#include "rodas3_launch.h"

namespace brams {

hipError_t launch_rodas3_128(bool dyn, unsigned nblocks, const StepParams& p, int block0, hipStream_t st) {
  if (nblocks == 0) return hipSuccess;
  if (dyn)
    hipLaunchKernelGGL((rodas3_blocks<128, true, false, 1>), dim3(nblocks), dim3(128), 0, st, p, block0);
  else
    hipLaunchKernelGGL((rodas3_blocks<128, false, false, 1>), dim3(nblocks), dim3(128), 0, st, p, block0);
  return hipGetLastError();
}

hipError_t launch_rodas3_64w2(unsigned nblocks, const StepParams& p, int block0, hipStream_t st) {
  if (nblocks == 0) return hipSuccess;
#ifdef BRAMS_MECH_LEAN
  (void)p, (void)block0, (void)st;
  return hipErrorInvalidValue;
#else
  hipLaunchKernelGGL((rodas3_blocks<64, false, false, 2>), dim3(nblocks), dim3(64), 0, st, p, block0);
  return hipGetLastError();
#endif
}

hipError_t launch_rodas3_list(unsigned nblocks, const StepParams& p, hipStream_t st, int wps) {
  if (nblocks == 0) return hipSuccess;
  if (!p.cell_list) return hipErrorInvalidValue;
#ifdef BRAMS_MECH_LEAN
  (void)st, (void)wps;
  return hipErrorInvalidValue;
#else
  if (wps == 2)
    hipLaunchKernelGGL((rodas3_blocks<64, false, true, 2>), dim3(nblocks), dim3(64), 0, st, p, 0);
  else if (wps == 1)
    hipLaunchKernelGGL((rodas3_blocks<64, false, true, 1>), dim3(nblocks), dim3(64), 0, st, p, 0);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
#endif
}

}  // namespace brams
