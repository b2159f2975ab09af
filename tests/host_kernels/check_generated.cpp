// CPU check of the generated device functions of one mechanism (relacs_tuv
// by default; -DBRAMS_MECH_GEN/-DBRAMS_MECH_SLAB name another), compiled for
// the host: the accessor-templated kernels of <mech>_slab.h against the
// plain ones of <mech>_gen.h and against that mechanism's oracle, bit for bit.
//
//   mech_kinetic_t == mech_kinetic == ora_kinetic
//   mech_fun_t     == mech_fun     == ora_fexchem
//   mech_jac_t     == ora_jacobian (every entry)
//   mech_factor_rl (right-looking, lazy) == mech_factor (left-looking): the
//                     same LU array and threshold count, random masks too
//   mech_solve_t   == mech_solve   == ora_lu_solve_dense
//
// Usage: check_generated <n_states> <seed>; prints one line per failure
// and exits nonzero if any.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <random>

#include "mechanism.h"
#include BRAMS_MECH_SLAB

extern "C" {
void ora_kinetic(double temp, double press, double vapp, const double *jphoto, double *rk);
void ora_fexchem(const double *rk, const double *y, double *f);
void ora_jacobian(const double *rk, const double *y, double *jac);
int ora_lu_solve_dense(const double *a, const unsigned char *mask, const double *b, double *x);
}

static int nfail = 0;
static bool same(double a, double b) { return std::memcmp(&a, &b, sizeof a) == 0; }
#define CHECK(cond, ...)                    \
  do {                                      \
    if (!(cond)) {                          \
      if (nfail < 20) std::printf(__VA_ARGS__); \
      ++nfail;                              \
    }                                       \
  } while (0)

int main(int argc, char** argv) {
  const int n = argc > 1 ? std::atoi(argv[1]) : 200;
  std::mt19937_64 gen(argc > 2 ? std::atoll(argv[2]) : 1);
  std::uniform_real_distribution<double> U(0.0, 1.0);
  for (int t = 0; t < n; ++t) {
    const double T = 220.0 + 90.0 * U(gen), P = 2.0e4 + 8.5e4 * U(gen);
    const double rv = std::exp(std::log(1e-5) + U(gen) * (std::log(2e-2) - std::log(1e-5)));
    double J[MECH_NPHOT];
    const bool night = (t % 4) == 3;
    for (int k = 0; k < MECH_NPHOT; ++k) J[k] = night ? 0.0 : 1e-6 * std::exp(10.0 * U(gen));
    double rk0[MECH_NR], rk1[MECH_NR], rk2[MECH_NR];
    mech_kinetic(T, P, rv, J, rk0);
    mech_kinetic_t(T, P, rv, J, rk1);
    ora_kinetic(T, P, rv, J, rk2);
    for (int r = 0; r < MECH_NR; ++r)
      CHECK(same(rk0[r], rk1[r]) && same(rk0[r], rk2[r]), "state %d rk %d: %.17g %.17g %.17g\n", t, r,
            rk0[r], rk1[r], rk2[r]);
    double y[MECH_NS];
    for (int s = 0; s < MECH_NS; ++s)
      y[s] = (t % 7 == 5 && s % 5 == 0) ? 0.0 : std::exp(std::log(1e2) + U(gen) * std::log(1e12));
    double f0[MECH_NS], f1[MECH_NS], f2[MECH_NS];
    mech_fun(rk0, y, f0);
    mech_fun_t(rk0, y, f1);
    ora_fexchem(rk0, y, f2);
    for (int s = 0; s < MECH_NS; ++s)
      CHECK(same(f0[s], f1[s]) && same(f0[s], f2[s]), "state %d f %d\n", t, s);
    double jac[MECH_NJAC], dense[MECH_NS * MECH_NS];
    unsigned all[MECH_MASK_WORDS];
    for (int w = 0; w < MECH_MASK_WORDS; ++w) all[w] = 0xffffffffu;
    mech_jac_t(rk0, y, all, jac);
    ora_jacobian(rk0, y, dense);
    for (int e = 0; e < MECH_NJAC; ++e) {
      const int i = mech_jac_i_host[e], j = mech_jac_j_host[e];
      CHECK(same(jac[e], dense[i * MECH_NS + j]), "state %d jac (%d,%d)\n", t, i, j);
    }
    // the value-based mask from stored entries == mech_mask on the same cell
    {
      const double igdt = 1.0 / (0.5 * (1.0 + 100.0 * U(gen)));
      unsigned ref[MECH_MASK_WORDS], off[MECH_MASK_WORDS], dia[MECH_MASK_WORDS];
      mech_mask(rk0, y, igdt, ref);
      mech_mask_offdiag(jac, off);
      mech_mask_diag(jac, igdt, dia);
      for (int w = 0; w < MECH_MASK_WORDS; ++w)
        CHECK(ref[w] == (off[w] | dia[w]), "state %d mask word %d: %08x %08x\n", t, w, ref[w], off[w] | dia[w]);
    }
    // masks: every structural entry, then random subsets with the diagonal kept
    for (int mk = 0; mk < 3; ++mk) {
      unsigned mw[MECH_MASK_WORDS];
      unsigned char mask[MECH_NS * MECH_NS];
      std::memset(mask, 0, sizeof mask);
      for (int w = 0; w < MECH_MASK_WORDS; ++w) mw[w] = 0;
      for (int e = 0; e < MECH_NSTRUCT; ++e) {
        const int i = mech_struct_i_host[e], j = mech_struct_j_host[e];
        const bool keep = mk == 0 || i == j || U(gen) < 0.85;
        if (keep) {
          mw[e / 32] |= 1u << (e % 32);
          mask[i * MECH_NS + j] = 1;
        }
      }
      const double igdt = 1.0 / (0.5 * (0.5 + 119.5 * U(gen)));
      double lu0[MECH_NLU], lu1[MECH_NLU];
      int v0 = 0, v1 = 0;
      const int b0 = mech_factor<true>(rk0, y, igdt, mw, lu0, &v0);
      double jm[MECH_NJAC];
      mech_jac_t(rk0, y, mw, jm);               // masked per call, as the slab kernel does
      const int b1 = mech_factor_rl(jm, igdt, mw, lu1, &v1);
      CHECK(b0 == b1 && v0 == v1, "state %d mask %d: bad %d/%d viol %d/%d\n", t, mk, b0, b1, v0, v1);
      for (int k = 0; k < MECH_NLU; ++k)
        CHECK(same(lu0[k], lu1[k]), "state %d mask %d lu[%d]: %.17g %.17g\n", t, mk, k, lu0[k], lu1[k]);
      double x0[MECH_NS], x1[MECH_NS], x2[MECH_NS], a[MECH_NS * MECH_NS];
      for (int s = 0; s < MECH_NS; ++s) x0[s] = x1[s] = f0[s] + 1e3 * (U(gen) - 0.5);
      for (int i = 0; i < MECH_NS; ++i)
        for (int j = 0; j < MECH_NS; ++j)
          a[i * MECH_NS + j] = (i == j ? igdt : 0.0) - dense[i * MECH_NS + j];
      for (int i = 0; i < MECH_NS; ++i)     // the plain kernel's diagonal: igdt - J_ii
        a[i * MECH_NS + i] = igdt - dense[i * MECH_NS + i];
      const int rc = ora_lu_solve_dense(a, mask, x0, x2);
      mech_solve(lu0, x0);
      mech_solve_t(lu1, x1);
      for (int s = 0; s < MECH_NS; ++s) {
        CHECK(same(x0[s], x1[s]), "state %d mask %d x %d\n", t, mk, s);
        if (rc == 0) CHECK(same(x0[s], x2[s]), "state %d mask %d oracle x %d: %.17g %.17g\n", t, mk, s, x0[s], x2[s]);
      }
    }
  }
  std::printf("%s: %d states, %d mismatches\n", nfail ? "FAIL" : "OK", n, nfail);
  return nfail ? 1 : 0;
}
