// Launch wrapper of the multi-grid RODAS3 kernel (rodas3_grids in
// rodas3_kernel.h): the nested grids of one model step in one launch per
// block form.
#include <hip/hip_runtime.h>

#include "rodas3_kernel.h"
#include "rodas3_launch.h"

namespace brams {

hipError_t launch_rodas3_grids(int wg, bool dyn, int wps, unsigned nwork, const StepParams* grids,
                               const int* work, hipStream_t st) {
  if (nwork == 0) return hipSuccess;
  if (!grids || !work) return hipErrorInvalidValue;
#ifdef BRAMS_MECH_LEAN
  // a lean library (rodas3_launch.h) builds no rodas3_grids kernel
  (void)wg, (void)dyn, (void)wps, (void)st;
  return hipErrorInvalidValue;
#else
  if (wg == 128 && dyn)
    hipLaunchKernelGGL((rodas3_grids<128, true, 1>), dim3(nwork), dim3(128), 0, st, grids, work);
  else if (wg == 128)
    hipLaunchKernelGGL((rodas3_grids<128, false, 1>), dim3(nwork), dim3(128), 0, st, grids, work);
  else if (wg == 64 && !dyn && wps == 4)
    hipLaunchKernelGGL((rodas3_grids<64, false, 4>), dim3(nwork), dim3(64), 0, st, grids, work);
  else if (wg == 64 && !dyn && wps == 1)
    hipLaunchKernelGGL((rodas3_grids<64, false, 1>), dim3(nwork), dim3(64), 0, st, grids, work);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
#endif
}

}  // namespace brams
