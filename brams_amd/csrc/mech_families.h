// Rate-law families of the mechanisms' kinetics (chem_spack_kinetic.f90 of
// RELACS_TUV, RELACS_MX, RACM_TUV, CB07_TUV), device side.  Each keeps the
// reference's evaluation order.
#pragma once
#include <hip/hip_runtime.h>

// exp(a - b/T), e.g. chem_spack_kinetic.f90:101
__device__ __forceinline__ double rl_arr(double T, double a, double b) {
  return exp(a - b / T);
}

// exp(a + n ln T - b/T), chem_spack_kinetic.f90:232-237
__device__ __forceinline__ double rl_arr_tn(double T, double lnT, double a, double n, double b) {
  return exp(a + n * lnT - b / T);
}

// Troe fall-off, chem_spack_kinetic.f90:126-133:
// Effko = k0 (T/300)^-n, Rapk = kinf (T/300)^-m,
// k = (Effko M / (1 + Effko M / Rapk)) * Fc^(1 / (1 + log10(Effko M / Rapk)^2))
__device__ __forceinline__ double rl_troe(double T, double M, double k0, double n,
                                          double kinf, double m, double fc) {
  const double effko = k0 * pow(T / 3.0e2, -(n));
  const double rapk = kinf * pow(T / 3.0e2, -(m));
  const double l = log10(effko * M / rapk);
  return (effko * M / (1.0 + effko * M / rapk)) * pow(fc, 1.0 / (1.0 + l * l));
}

// HNO3 + OH, chem_spack_kinetic.f90:181-184
__device__ __forceinline__ double rl_hno3_oh(double T, double M, double c1, double b1, double c2,
                                             double b2, double c3, double b3) {
  const double effko = c1 * exp(b1 / T);
  const double rapk = c2 * exp(b2 / T);
  const double fac = c3 * exp(b3 / T) * M;
  return effko + fac / (1.0 + fac / rapk);
}

// fall-off of the chemical-activation form, CB07_TUV/chem_spack_kinetic.f90:231-234:
// Effko = k0 (T/300)^n, Rapk = kinf (T/300)^m,
// k = Effko/(1 + Effko/(Rapk/M)) * Fc^(1 + (log10(Effko/(Rapk/M))^2)^-1), the
// exponent as the reference's parentheses group it (not the usual
// 1/(1 + log10^2)), with its real powers
__device__ __forceinline__ double rl_troe_ca(double T, double M, double k0, double n,
                                             double kinf, double m, double fc) {
  const double effko = k0 * pow(T / 300.0, n);
  const double rapk = kinf * pow(T / 300.0, m);
  const double x = log10(effko / (rapk / M));
  return (effko / (1.0 + effko / (rapk / M))) * pow(fc, 1.0 + pow(pow(x, 2.0), -1.0));
}
