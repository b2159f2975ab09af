// Host entry of the RODAS3 and ROS2 block kernels (defined in
// rodas3_launch.hip, rodas3_launch_b.hip, rodas3_grids.hip and
// ros2_launch.hip, split so the units compile in parallel).
#pragma once
#include <hip/hip_runtime.h>

namespace brams {

struct StepParams;

// Kernel forms built into the library.  A library built with
// BRAMS_MECH_LEAN (the mechanisms other than RELACS_TUV) instantiates the
// per-lane RODAS3 and ROS2 kernels at one wave per SIMD for 64 and 128
// lanes, and the 64-lane ones also at four (RODAS3: the autotuner keeps the
// faster; ROS2: from kRodas3ManyBlocks blocks on); neither the two-wave, list
// (surrogate fallback), several-grid (rodas3_grids; the entry launches grid
// by grid instead) nor slab kernels: each instantiation of a large mechanism
// takes minutes to compile.  The launchers below return hipErrorInvalidValue
// for a form not built.
#ifdef BRAMS_MECH_LEAN
constexpr bool kLeanBuild = true;
#else
constexpr bool kLeanBuild = false;
#endif

// nblocks workgroups of wg (64 or 128) lanes over blocks block0.. of p;
// dyn: block-1 form (mask recomputed on every attempt and published).
// wps: waves per SIMD of the 64-lane form (1, 2 or 4; 0: by block count).
hipError_t launch_rodas3(int wg, bool dyn, unsigned nblocks, const StepParams& p, int block0,
                         hipStream_t st, int wps = 0);
// nblocks workgroups of 64 lanes over consecutive 64-entry runs of p.cell_list
// (wps 1 or 2).
hipError_t launch_rodas3_list(unsigned nblocks, const StepParams& p, hipStream_t st, int wps = 1);

// the parts of the above that live in rodas3_launch_b.hip
hipError_t launch_rodas3_128(bool dyn, unsigned nblocks, const StepParams& p, int block0, hipStream_t st);
hipError_t launch_rodas3_64w2(unsigned nblocks, const StepParams& p, int block0, hipStream_t st);

// The ROS2 solver (ros2_launch.hip): nblocks workgroups of wg (64 or 128)
// lanes over blocks block0.. of p; wps: waves per SIMD of the 64-lane form
// (1 or 4; 0: four from kRodas3ManyBlocks blocks on, else one).
hipError_t launch_ros2(int wg, unsigned nblocks, const StepParams& p, int block0, hipStream_t st, int wps = 0);

// The QSSA solver (qssa_launch.hip): one lane per interior cell, itchim
// sub-steps; nfatal (device int, zeroed by the caller) counts the cells where
// a species reached 1e20.
hipError_t launch_qssa(const StepParams& p, int itchim, long long ncell, int* nfatal, hipStream_t st);

// Several grids in one launch (rodas3_grids): nwork workgroups of wg lanes,
// workgroup w on block work[w] & kGridBlockMask of grid work[w] >> kGridShift,
// whose parameters are grids[] (device memory).  wps 1 or 4 (64 lanes only).
hipError_t launch_rodas3_grids(int wg, bool dyn, int wps, unsigned nwork, const StepParams* grids,
                               const int* work, hipStream_t st);

}  // namespace brams
