// Launch wrapper of the ROS2 block kernels (ros2_blocks in rodas3_kernel.h),
// the CHEMISTRY=3 solver of the reference (chem_spack_ros_dyndt.f90).
#include <hip/hip_runtime.h>

#include "rodas3_kernel.h"
#include "rodas3_launch.h"

namespace brams {

hipError_t launch_ros2(int wg, unsigned nblocks, const StepParams& p, int block0, hipStream_t st, int wps) {
  if (nblocks == 0) return hipSuccess;
  if (wg == 64 && wps == 0) wps = nblocks >= (unsigned)kRodas3ManyBlocks ? kRodas3ManyWaves : 1;
  if (wg == 128)
    hipLaunchKernelGGL((ros2_blocks<128, 1>), dim3(nblocks), dim3(128), 0, st, p, block0);
  else if (wg == 64 && wps == kRodas3ManyWaves)
    hipLaunchKernelGGL((ros2_blocks<64, kRodas3ManyWaves>), dim3(nblocks), dim3(64), 0, st, p, block0);
  else if (wg == 64 && wps == 1)
    hipLaunchKernelGGL((ros2_blocks<64, 1>), dim3(nblocks), dim3(64), 0, st, p, block0);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

}  // namespace brams
