// Photolysis rates on the device: the FAST-TUV output mapping of tuvDriver
// (ModTuvDriver2.7.f90:1302-1332) and the mechanism's J relations
// (:1333-1368; RELACS: J14 and J16, mech_j_relations of the generated code).
//
// tuvDriver gathers TUV's j-values into valJLoc(ia:iz, ja:jz, kj, kz) (double,
// :1263) and writes, for every TUV reaction k with xRef(k) /= 0 in ascending
// k, jphoto(xRef(k), m+1, i, j) = valJLoc(i, j, k, m) for m = 1..m1-1 and
// jphoto(xRef(k), 1, i, j) = jphoto(xRef(k), 2, i, j), over the interior
// columns.  RELACS then sets, at every point of the array,
//   J14 = 0.962055 J13 + 0.0106247 J9      J16 = 12 J13 + 208 J15
// with the constants as REAL literals, without FMA contraction (other
// mechanisms have their own relations, or none).
//
// The host resolves xRef into src[n]: the last k that writes jphoto index n
// (later reactions overwrite earlier ones), -1 where none does (that entry
// keeps its value).  One thread per point (level fastest, so a wave writes
// consecutive 136-byte rows of jphoto); its reads of valJLoc at one level are
// strided, and the neighbouring columns' waves use the rest of each line.
#pragma once
#include <hip/hip_runtime.h>

#include "mechanism.h"

namespace brams {

struct TuvMapParams {
  int m1, m2, m3, ia, iz, ja, jz;
  int kj, kz;
  const double* valj;            // (iz-ia+1, jz-ja+1, kj, kz), Fortran order
  double* jphoto;                // (MECH_NPHOT, m1, m2, m3)
  int src[MECH_NPHOT];           // TUV reaction (0-based) that writes index n, -1: none
};

// Workgroups are dealt to the eight XCDs in turn (workgroup w on XCD w % 8),
// each with its own L2.  Workgroup w takes tile (w % 8) * chunk + w / 8, so
// every XCD works through one contiguous run of columns, and the 16
// neighbouring columns that share a 128-byte line of valJLoc are read through
// one L2 instead of up to eight.
constexpr int kXcds = 8;
__device__ __forceinline__ long long xcd_tile(unsigned chunk) {
  return (long long)(blockIdx.x % kXcds) * chunk + blockIdx.x / kXcds;
}

// interior columns: the mapping, then the relations; launched with
// kXcds * chunk workgroups of 256 (chunk = ceil(tiles / kXcds))
__global__ __launch_bounds__(256) void tuv_map_interior(TuvMapParams p, long long npts, unsigned chunk) {
  const long long t = xcd_tile(chunk) * 256 + threadIdx.x;
  if (t >= npts) return;
  const int ni = p.iz - p.ia + 1, nj = p.jz - p.ja + 1;
  const int L = (int)(t % p.m1);
  const long long col = t / p.m1;
  const int ii = (int)(col % ni), jj = (int)(col / ni);
  const long long o = L + (long long)p.m1 * ((p.ia - 1 + ii) + (long long)p.m2 * (p.ja - 1 + jj));
  const int mz = L == 0 ? 0 : L - 1;                 // level 1 repeats level 2
  double v[MECH_NPHOT];
#pragma unroll
  for (int n = 0; n < MECH_NPHOT; ++n) {
    const int k = p.src[n];
    v[n] = k >= 0 ? p.valj[ii + (long long)ni * (jj + (long long)nj * (k + (long long)p.kj * mz))]
                  : p.jphoto[n + (long long)MECH_NPHOT * o];
  }
  mech_j_relations(v);
#pragma unroll
  for (int n = 0; n < MECH_NPHOT; ++n) p.jphoto[n + (long long)MECH_NPHOT * o] = v[n];
}

// points outside the interior columns: the relations alone (the entries they
// do not set are written back unchanged)
__global__ __launch_bounds__(256) void jphoto_relations_outside(TuvMapParams p, long long n3d) {
  const long long o = (long long)blockIdx.x * 256 + threadIdx.x;
  if (o >= n3d) return;
  const long long col = o / p.m1;
  const int i = (int)(col % p.m2) + 1, j = (int)(col / p.m2) + 1;
  if (i >= p.ia && i <= p.iz && j >= p.ja && j <= p.jz) return;
  double* jp = p.jphoto + (long long)MECH_NPHOT * o;
  double v[MECH_NPHOT];
#pragma unroll
  for (int n = 0; n < MECH_NPHOT; ++n) v[n] = jp[n];
  mech_j_relations(v);
#pragma unroll
  for (int n = 0; n < MECH_NPHOT; ++n) jp[n] = v[n];
}

}  // namespace brams
