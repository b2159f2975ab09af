// Host entry of the RODAS3 block kernels (defined in rodas3_launch.hip).
#pragma once
#include <hip/hip_runtime.h>

namespace brams {

struct StepParams;

// nblocks workgroups of wg (64 or 128) lanes over blocks block0.. of p;
// dyn: block-1 form (mask recomputed on every attempt and published).
hipError_t launch_rodas3(int wg, bool dyn, unsigned nblocks, const StepParams& p, int block0,
                         hipStream_t st);
// nblocks workgroups of 64 lanes over consecutive 64-entry runs of p.cell_list.
hipError_t launch_rodas3_list(unsigned nblocks, const StepParams& p, hipStream_t st);

}  // namespace brams
