/*
 * ORACLE -- test infrastructure only.
 *
 * CPU restatement of the reference's CCATT RODAS3 chemistry step
 * (src/brams/ccatt/chem_spack_rodas3_dyndt.f90) and of the RELACS_TUV
 * mechanism kernels it calls.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg load it, as the checker; the product path
 * (brams_amd/) never calls it.
 *
 * Pinning (see tests/):
 *   ora_kinetic / ora_fexchem / ora_jacobian  bit-for-bit against the
 *       reference's own compiled kernels (oracle/_ref) at random states;
 *   ora_lu_*        against the reference's compiled Sparse1.3 and against
 *       numpy.linalg.solve (residuals);
 *   ora_alloc_index against the reference's compiled AllocIndex;
 *   ora_rodas3_dyndt end to end against the same driver run on the
 *       reference's Sparse1.3 (ora_use_external_sparse), plus the closed-form
 *       checks of the Rodas3 tableau (order 3 on y'=Ay vs expm) and of the
 *       step-size controller.
 */
#ifndef BRAMS_ORACLE_H
#define BRAMS_ORACLE_H
#include <stdint.h>
#include "../include/brams_chem.h"

#ifdef __cplusplus
extern "C" {
#endif

/* One cell's rate coefficients rk[128] from T (K), p (Pa), rv (kg/kg), J[17]. */
void ora_kinetic(double temp, double press, double vapp, const double *jphoto, double *rk);
/* Net production f[47] = S w(rk, y). */
void ora_fexchem(const double *rk, const double *y, double *f);
/* Jacobian df/dy, dense row-major jac[i*47+j] = d f_i / d y_j. */
void ora_jacobian(const double *rk, const double *y, double *jac);

/* Fixed pivot order used by the oracle's LU: perm[k] = species pivoted at step k. */
int ora_lu_order(int *perm, int *nnz_lu);
/* Factor the masked matrix a (row-major 47x47, mask 0/1) and solve a x = b. */
int ora_lu_solve_dense(const double *a, const unsigned char *mask, const double *b, double *x);
/* Pivots of the fixed order that fail Sparse1.3's threshold rule when the
 * masked matrix a is factored (the count the device kernels report). */
int ora_lu_threshold_fails(const double *a, const unsigned char *mask);

/* Block partition, AllocIndex (chem_spack_utils.f90:39-183); returns nob. */
int ora_alloc_index(int block_size, int m1, int ia, int iz, int ja, int jz,
                    int32_t *block_end, int max_nob);

/* The whole step.  nthreads<=0: use all host cores.  mask_out (47*47, may be
 * NULL) receives the value-based sparsity mask blocks 2..nob used. */
int ora_rodas3_dyndt(const brams_chem_args *a, int nthreads, brams_chem_stats *st,
                     unsigned char *mask_out);

/* The same step on block 1 and the listed blocks only (0-based ids; others
 * are left untouched).  Blocks are independent once block 1 has fixed the
 * mask, so each listed block's result equals the full step's. */
int ora_rodas3_dyndt_blocks(const brams_chem_args *a, int nthreads, brams_chem_stats *st,
                            unsigned char *mask_out, const int32_t *blocks, int nblocks);

/* The RODAS3 stages, solution and error norm of the step applied to the
 * linear problem y' = A y (dense n x n, n <= 47) with nsteps fixed steps h;
 * err_out = error norm of the last step with atol = rtol = 1. */
int ora_rodas3_linear(int n, const double *A, const double *y0, double h, int nsteps,
                      double *y_out, double *err_out);
/* The step-size controller: dt * min(6, max(0.2, 0.9 err^-1/3)) in [dt_min, dt_max]. */
double ora_dt_new(double dt, double max_err, double dt_min, double dt_max);

/* The ROS2 solver of chem_spack_ros_dyndt.f90 (CHEMISTRY=3): the whole step,
 * and block 0 plus the listed blocks (blocks are independent: ROS2 has no
 * value mask).  mask_out receives the structural pattern it factors on. */
int ora_ros2_dyndt(const brams_chem_args *a, int nthreads, brams_chem_stats *st, unsigned char *mask_out);
int ora_ros2_dyndt_blocks(const brams_chem_args *a, int nthreads, brams_chem_stats *st,
                          unsigned char *mask_out, const int32_t *blocks, int nblocks);
/* ROS2 on the linear problem y' = A y, as ora_rodas3_linear. */
int ora_ros2_linear(int n, const double *A, const double *y0, double h, int nsteps,
                    double *y_out, double *err_out);
/* ROS2's controller: dt * min(6, max(0.2, 0.9 err^-1/2)) in [dt_min, dt_max]. */
double ora_ros2_dt_new(double dt, double max_err, double dt_min, double dt_max);

/* The QSSA solver of chem_spack_qssa.f90 (CHEMISTRY=1): the whole step with
 * itchim sub-steps (sc_t of every species += (new - old)/dtlt; sc_p is
 * unchanged).  Returns the number of cells where a species reached 1e20
 * (their tendencies are left unchanged), -1 on invalid arguments. */
long ora_qssa(const brams_chem_args *a, int itchim, int nthreads);
/* Production and loss of the P - L c split (rates, dratedc, fexprod,
 * fexloss) of one cell. */
void ora_prod_loss(const double *rk, const double *y, double *prod, double *loss);
/* One species through itchim QSSA sub-steps of length delt, P and L fixed. */
double ora_qssa_species(double y, double prod, double loss, double delt, int itchim);
/* One cell's QSSA gather (ppbm -> molec/cm^3 for every species). */
int ora_qssa_gather_cell(const brams_chem_args *a, long cell, double *y, double *temp, double *press,
                         double *volmol);

/* The FAST-TUV output mapping of tuvDriver and RELACS's J relations
 * (ModTuvDriver2.7.f90:1302-1332, 1363-1368): valj = valJLoc(ia:iz, ja:jz,
 * kj, kz) (Fortran order), xref[kj] (1-based jphoto index, 0 unused), jphoto
 * (np, m1, m2, m3) updated in place.  -1 on invalid arguments. */
int ora_tuv_jphoto(int m1, int m2, int m3, int ia, int iz, int ja, int jz, int kj, int kz,
                   const double *valj, const int32_t *xref, int np, double *jphoto);

/* Per-block counts of the last step: attempts, accepted sub-steps. */
int ora_block_counts(int64_t *attempts, int64_t *substeps, int max_nob);
/* Per-block largest error accepted at dt_min (0: none) of the last step. */
int ora_block_forced(float *forced, int max_nob);
/* Per-block count of pivots failing Sparse1.3's threshold rule, last step. */
int ora_block_thresholds(int32_t *counts, int max_nob);
/* per-block failure flags of the last step (a zero pivot or a non-finite
 * error sum: the point where the reference stops the model) */
int ora_block_failed(int32_t *failed, int max_nob);

/* Test hooks.  One cell's gather (Exner -> T, p; ppbm -> molec/cm^3) and
 * write-back alone; cell is the interior index in j -> i -> k order. */
int ora_gather_cell(const brams_chem_args *a, long cell, double *y, double *temp, double *press,
                    double *volmol);
int ora_scatter_cell(const brams_chem_args *a, long cell, const double *y);
/* Replace block 1's largest error of attempt i by errs[min(i, n-1)]
 * (NULL/0 switches it off); run single-threaded with a one-block grid. */
void ora_set_error_script(const double *errs, int n);
/* Block 1's attempts of the last step: dt tried, accepted (0 rejected, 1
 * accepted, 2 accepted with error > 1 at dt_min), time after the attempt.
 * Returns the number of attempts. */
int ora_controller_trace(double *dt, int *accepted, double *time_after, int max);
/* Jacobian and f(y0) evaluations (cells x sub-steps) of the last step. */
void ora_call_counts(int64_t *jacobian_calls, int64_t *f0_calls);

/* Route the LU through an external Sparse1.3 build (the reference's, via
 * oracle/ref_sparse_harness.c).  NULL pointers restore the oracle's own LU.
 * With an external LU the step runs single-threaded, exactly as the reference
 * drives Sparse (chem_spack_rodas3_dyndt.f90:470-603). */
typedef long (*ora_sp_create_fn)(int n, int nnz, const int *ipos, const int *jpos, long *elements);
typedef int (*ora_sp_load_factor_fn)(long h, int nnz, const long *elements, const double *vals);
typedef void (*ora_sp_solve_fn)(long h, double *rhs, double *x);
void ora_use_external_sparse(ora_sp_create_fn c, ora_sp_load_factor_fn lf, ora_sp_solve_fn s);

#ifdef __cplusplus
}
#endif
#endif
