// Launch wrappers of the RODAS3 block kernels (rodas3_kernel.h), 64-lane
// partition forms.
//
// The kernels are instantiated in the rodas3_launch*.hip and rodas3_grids.hip
// units only, so the rest of the host library compiles without them; wg is
// the workgroup size (64 or 128 cells per block), dyn selects the block-1
// form that recomputes and publishes the sparsity mask, wps the register
// budget (waves per SIMD) of the 64-lane form.
#include <hip/hip_runtime.h>

#include "rodas3_kernel.h"
#include "rodas3_launch.h"

namespace brams {

hipError_t launch_rodas3(int wg, bool dyn, unsigned nblocks, const StepParams& p, int block0,
                         hipStream_t st, int wps) {
  if (nblocks == 0) return hipSuccess;
  if (wg == 128) return launch_rodas3_128(dyn, nblocks, p, block0, st);
  if (wg != 64) return hipErrorInvalidValue;
  if (dyn) {
    hipLaunchKernelGGL((rodas3_blocks<64, true, false, 1>), dim3(nblocks), dim3(64), 0, st, p, block0);
    return hipGetLastError();
  }
  if (wps == 0) wps = nblocks >= (unsigned)kRodas3ManyBlocks ? kRodas3ManyWaves : 1;
  if (wps == 2) return launch_rodas3_64w2(nblocks, p, block0, st);
  if (wps == 4)
    hipLaunchKernelGGL((rodas3_blocks<64, false, false, 4>), dim3(nblocks), dim3(64), 0, st, p, block0);
  else if (wps == 1)
    hipLaunchKernelGGL((rodas3_blocks<64, false, false, 1>), dim3(nblocks), dim3(64), 0, st, p, block0);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

}  // namespace brams
