// Launch wrappers of the RODAS3 block kernels (rodas3_kernel.h).
//
// The kernels are instantiated in this translation unit only, so the rest of
// the host library compiles without them; wg is the workgroup size (64 or
// 128 cells per block), dyn selects the block-1 form that recomputes and
// publishes the sparsity mask.  Large partition launches take the
// kRodas3ManyWaves register budget, every other launch the whole register
// file (rodas3_kernel.h).
#include <hip/hip_runtime.h>

#include "rodas3_kernel.h"
#include "rodas3_launch.h"

namespace brams {

hipError_t launch_rodas3(int wg, bool dyn, unsigned nblocks, const StepParams& p, int block0,
                         hipStream_t st) {
  if (nblocks == 0) return hipSuccess;
  if (wg == 64 && dyn)
    hipLaunchKernelGGL((rodas3_blocks<64, true, false, 1>), dim3(nblocks), dim3(64), 0, st, p, block0);
  else if (wg == 64 && nblocks >= (unsigned)kRodas3ManyBlocks)
    hipLaunchKernelGGL((rodas3_blocks<64, false, false, kRodas3ManyWaves>), dim3(nblocks), dim3(64), 0, st, p,
                       block0);
  else if (wg == 64)
    hipLaunchKernelGGL((rodas3_blocks<64, false, false, 1>), dim3(nblocks), dim3(64), 0, st, p, block0);
  else if (wg == 128 && dyn)
    hipLaunchKernelGGL((rodas3_blocks<128, true, false, 1>), dim3(nblocks), dim3(128), 0, st, p, block0);
  else if (wg == 128)
    hipLaunchKernelGGL((rodas3_blocks<128, false, false, 1>), dim3(nblocks), dim3(128), 0, st, p, block0);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t launch_rodas3_list(unsigned nblocks, const StepParams& p, hipStream_t st) {
  if (nblocks == 0) return hipSuccess;
  if (!p.cell_list) return hipErrorInvalidValue;
  hipLaunchKernelGGL((rodas3_blocks<64, false, true, 1>), dim3(nblocks), dim3(64), 0, st, p, 0);
  return hipGetLastError();
}

}  // namespace brams
