// Host emulation of the slab kernel's block integration (rodas3_slab.h) for
// one-cell grids, for CPU tests.  With a single cell every lane of the wave
// holds the same values (idle lanes shadow cell 1), so one lane emulates the
// wave exactly: cross-lane operations become the identity, LDS and slab
// regions plain arrays, buffer accesses plain loads and stores.
#include <cstdint>
#include <cstring>
#include <cstdio>

#include "hip/hip_runtime.h"

struct HostRsrc {
  char* base;
  long long n;
};
#define __amdgpu_buffer_rsrc_t HostRsrc
#define __builtin_amdgcn_make_buffer_rsrc(p, stride, n, flags) (HostRsrc{(char*)(p), (long long)(n)})
#define __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, aux) \
  (*(const bv2*)((r).base + (voff) + (soff)))
#define __builtin_amdgcn_raw_buffer_store_b64(v, r, voff, soff, aux) \
  (*(bv2*)((r).base + (voff) + (soff)) = (v))
#define __builtin_amdgcn_readfirstlane(x) (x)
#define __builtin_amdgcn_sched_barrier(x) ((void)0)
#define __shfl(x, l, w) (x)
#define __shfl_xor(x, o, w) (x)
#define __syncthreads() ((void)0)
#define __global__
#define __launch_bounds__(a, b)
#define __shared__ static
#define BRAMS_OPAQUE_SGPR(x) ((void)0)
#define atomicAdd(p, v) ((*(p) += (v)) - (v))
#define hipLaunchKernelGGL(...) ((void)0)
struct HostDim { int x = 0, y = 0, z = 0; };
static HostDim threadIdx, blockIdx;
using std::fmin;
using std::isfinite;
using std::sqrt;
using std::fmaxf;
static int min(int a, int b) { return a < b ? a : b; }

#include "rodas3_slab.h"

using namespace brams;

// one call of the block-1 form on a one-cell grid; arrays as the kernel gets them
extern "C" int slab_step_host(int m1, int m2, int m3, float dtlt, const float* pp, const float* pi0,
                              const float* theta, const float* rv, const double* jphoto, const float* weight,
                              const int* transported, float* sc_p, double* last_dt, const double* atol,
                              const double* rtol, int* counts /* att, sub, failed, thresh, order_thresh */,
                              unsigned* mask) {
  static double lu[MECH_NLU * 64], jac[MECH_NJAC * 64], rk[MECH_NR * 64], vec[2 * MECH_NS * 64];
  static double s_y[MECH_NS * 64], s_lu[32 * 64];
  StepParams p{};
  p.m1 = m1; p.m2 = m2; p.m3 = m3; p.ia = 2; p.iz = m2 - 1; p.ja = 2; p.jz = m3 - 1;
  p.nk = m1 - 2; p.ni = m2 - 2;
  p.dtlt = dtlt; p.ndyn = 1; p.split = 0;
  p.cp = 1004.0f; p.cpor = 1004.0f / 287.0f; p.p00 = 1.0e5f;
  p.pp = pp; p.pi0 = pi0; p.theta = theta; p.rv = rv; p.jphoto = jphoto;
  p.weight = weight; p.transported = transported; p.sc_p = sc_p;
  p.n3d = (long long)m1 * m2 * m3;
  static int block_start[2];
  block_start[0] = 0; block_start[1] = (m1 - 2) * (m2 - 2) * (m3 - 2);
  p.block_start = block_start;
  p.last_dt = last_dt;
  p.atol = atol; p.rtol = rtol;
  p.mask = mask;
  p.blk_attempts = counts; p.blk_substeps = counts + 1; p.blk_failed = counts + 2;
  p.blk_thresh = counts + 3; p.order_thresh = counts + 4;
  p.blk_forced = nullptr;
  const Lanes ys{s_y};
  const LuLanes<32> lua{s_lu, BufLanes{slab_rsrc(lu, MECH_NLU * 512LL), 0}};
  const BufLanes ja{slab_rsrc(jac, MECH_NJAC * 512LL), 0};
  const BufLanes rka{slab_rsrc(rk, MECH_NR * 512LL), 0};
  const BufLanes f0{slab_rsrc(vec, MECH_NS * 512LL), 0};
  const BufLanes p4{slab_rsrc(vec + MECH_NS * 64, MECH_NS * 512LL), 0};
  integrate_slab<32, false, true>(p, 0, ys, lua, ja, rka, f0, p4);
  return counts[2];
}
