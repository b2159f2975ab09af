// The QSSA solver of the reference (CHEMISTRY=1, chem_spack_qssa.f90:37-411)
// on CDNA4.
//
// Cells are independent: QSSA has no step-size control, so AllocIndex's
// blocks play no role and one lane takes one cell of the interior, in the
// j -> i -> k order of the blocks (k fastest, so a wave's loads of one
// species coalesce along the column).  Per cell:
//   gather          chem_spack_qssa.f90:161-195 (every species from ppbm)
//   kinetic         the generated mech_kinetic (chem_spack_kinetic.f90)
//   P and L         mech_prod_loss (rates, dratedc, fexprod, fexloss; :247-265)
//   itchim sub-steps of length dtlt/itchim with P and L held (solveur :343-348,
//                   qssa :368-411), negatives set to 0 (:352-354)
//   tendency        sc_t += (new - old) / dtlt for every species (:273-292)
// A cell where a species reaches 1e20 (solveur :357-361 stops the model) is
// counted and writes nothing.
#include <hip/hip_runtime.h>

#include "rodas3_kernel.h"
#include "rodas3_launch.h"

namespace brams {

// One wave per SIMD: the rate constants and the species' state, production
// and loss (about 270 doubles) need the whole register file, and a bound of
// two waves costs 300 spilled registers for no gain.  Species go through the
// sub-steps eight at a time, so each FMA chain has seven independent ones
// beside it.  1.6M cells (ms): groups of 1 / 2 / 4 / 8 / 12 / 16 / 48 species
// 3.13 / 2.60 / 2.27 / 2.12 / 2.13 / 2.17 / 2.61; 64, 128 or 256 lanes per
// workgroup alike.
constexpr int kQssaWG = 128;
constexpr int kQssaGroup = 8;                                   // species integrated together
constexpr int kQssaPad = (MECH_NS + kQssaGroup - 1) / kQssaGroup * kQssaGroup;

__global__ __launch_bounds__(kQssaWG) void qssa_cells(StepParams p, int itchim, long long ncell, int* nfatal) {
  const long long g = (long long)blockIdx.x * kQssaWG + threadIdx.x;
  if (g >= ncell) return;
  const int k = 1 + (int)(g % p.nk);
  const long long t = g / p.nk;
  const int i = p.ia - 1 + (int)(t % p.ni);
  const int j = p.ja - 1 + (int)(t / p.ni);
  const long long o = k + (long long)p.m1 * (i + (long long)p.m2 * j);

  // gather (:161-195): Exner in REAL, R and pmar as REAL literals, the REAL
  // product 6.02e23*1e-15; J for FAST-TUV as for FAST-JX
  const float exner = (p.pp[o] + p.pi0[o]) / p.cp;
  double press = (double)exner;
  const double temp = (double)p.theta[o] * press;
  press = pow(press, (double)p.cpor) * (double)p.p00;
  const double vapp = (double)p.rv[o];
  const double volmol = press / ((double)8.314f * temp);
  const double avo = (double)(6.02e23f * 1e-15f);
  const double pmar = (double)28.96f;
  double y[kQssaPad];
#pragma unroll
  for (int n = 0; n < MECH_NS; ++n)
    y[n] = (double)p.sc_p[n * p.n3d + o] * (avo * volmol) * pmar / (double)p.weight[n];
#pragma unroll
  for (int n = MECH_NS; n < kQssaPad; ++n) y[n] = 0.0;
  double J[MECH_NPHOT];
#pragma unroll
  for (int q = 0; q < MECH_NPHOT; ++q) J[q] = p.jphoto[q + (long long)MECH_NPHOT * o];
  double rk[MECH_NR];
  mech_kinetic(temp, press, vapp, J, rk);
  double P[kQssaPad], L[kQssaPad];
  mech_prod_loss(rk, y, P, L);
#pragma unroll
  for (int n = MECH_NS; n < kQssaPad; ++n) P[n] = L[n] = 0.0;

  // itchim sub-steps (qssa :392-406).  Per species: ybid = L if L > 0, else 1
  // (cvmgm_dp); the explicit step y + delt (P - L y) when L = 0, else the
  // steady state P/ybid when ybid delt >= 10, else P/ybid + (y - P/ybid)
  // exp(-ybid delt) (cvmgp_dp, cvmgz_dp; qssa:406 overrides the explicit
  // choice of :404).  The branch is fixed for the call, and all three are one
  // sub-step v <- (v - a) e + c with per-species coefficients:
  //   exponential  a = c = P/ybid, e = exp(-ybid delt)
  //   steady       a = c = P/ybid, e = 0 (v starts at P/ybid)
  //   explicit     a = 0, e = 1, c = delt P
  // so a wave takes one sub and one FMA per species and sub-step, whatever
  // branch each of its cells takes.
  const double delt = (double)(p.dtlt / (float)itchim);
  bool fatal = false;
#pragma unroll
  for (int s0 = 0; s0 < kQssaPad; s0 += kQssaGroup) {
    double v[kQssaGroup], a[kQssaGroup], e[kQssaGroup], c[kQssaGroup];
#pragma unroll
    for (int q = 0; q < kQssaGroup; ++q) {
      const double l = L[s0 + q];
      const double ybid = (-l < 0.0) ? l : 1.0;
      const double pr = P[s0 + q] / ybid;
      v[q] = y[s0 + q];
      if (l == 0.0) {
        a[q] = 0.0;
        e[q] = 1.0;
        c[q] = delt * (P[s0 + q] - l * v[q]);
      } else if (ybid * delt - 10.0 >= 0.0) {
        a[q] = c[q] = pr;
        e[q] = 0.0;
        if (itchim > 0) v[q] = pr;
      } else {
        a[q] = c[q] = pr;
        e[q] = exp(-ybid * delt);
      }
    }
    for (int it = 0; it < itchim; ++it) {
#pragma unroll
      for (int q = 0; q < kQssaGroup; ++q) v[q] = fma(v[q] - a[q], e[q], c[q]);
    }
#pragma unroll
    for (int q = 0; q < kQssaGroup; ++q) {
      double r = v[q];
      if (r < 0.0) r = 0.0;
      if (r >= 1.0e20) fatal = true;
      y[s0 + q] = r;
    }
  }
  if (fatal) {
    atomicAdd(nfatal, 1);
    return;
  }
  // tendency of every species (:273-292)
  const double dtlt_i = (double)(1.0f / p.dtlt);
#pragma unroll
  for (int n = 0; n < MECH_NS; ++n) {
    const double fxc = 1.0 / (avo * volmol * (pmar / (double)p.weight[n]));
    const long long a = n * p.n3d + o;
    p.sc_t[a] = (float)((double)p.sc_t[a] + (y[n] * fxc - (double)p.sc_p[a]) * dtlt_i);
  }
}

hipError_t launch_qssa(const StepParams& p, int itchim, long long ncell, int* nfatal, hipStream_t st) {
  if (ncell <= 0) return hipSuccess;
  const unsigned nwg = (unsigned)((ncell + kQssaWG - 1) / kQssaWG);
  hipLaunchKernelGGL(qssa_cells, dim3(nwg), dim3(kQssaWG), 0, st, p, itchim, ncell, nfatal);
  return hipGetLastError();
}

}  // namespace brams
