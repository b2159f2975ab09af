/*
 * brams_surrogate.h -- learned surrogate of the CCATT chemistry step.
 *
 * No counterpart in the reference (which only has the RODAS3 solver,
 * src/brams/ccatt/chem_spack_rodas3_dyndt.f90).  A small MLP maps one cell's
 * state (log10 concentrations of the 47 species, T, ln p, log10 rv, the 17
 * photolysis rates, log10 dt) to the log10 increments of the 47 species over
 * one chemistry step.  Weights and activations are BF16 on the matrix cores,
 * accumulation and master weights FP32.  Cells whose prediction fails the
 * checks (non-finite or negative state, nitrogen/carbon budget, ensemble
 * spread) are integrated by the FP64 RODAS3 kernels instead.
 *
 * All array arguments are device pointers; calls are synchronous.
 * Status: 0 ok, <0 error (text in brams_chem_last_error()).
 */
#ifndef BRAMS_SURROGATE_H
#define BRAMS_SURROGATE_H

#include <stdint.h>
#include "brams_chem.h"

#ifdef __cplusplus
extern "C" {
#endif

#define BRAMS_SUR_NIN 68       /* features per cell */
#define BRAMS_SUR_NIN_PAD 128  /* padded feature row (bf16) */
#define BRAMS_SUR_NOUT 47      /* log10 increments */
#define BRAMS_SUR_NOUT_PAD 64  /* padded output row (f32) */

typedef struct brams_mlp brams_mlp;

/* One network: nlayers hidden layers of width nhidden (multiple of 64) with
 * GELU, input BRAMS_SUR_NIN_PAD, output BRAMS_SUR_NOUT_PAD; training batches
 * up to max_batch rows (multiple of 64).  Weights are initialised
 * U(-1/sqrt(fan_in), 1/sqrt(fan_in)) from seed. */
int brams_mlp_create(int nhidden, int nlayers, int max_batch, uint64_t seed, brams_mlp **out);
void brams_mlp_destroy(brams_mlp *m);
/* number of FP32 master parameters (layer by layer: W[out][in] then b[out]) */
int64_t brams_mlp_num_params(const brams_mlp *m);
/* copy master parameters device->host / host->device (n floats) */
int brams_mlp_get_params(const brams_mlp *m, float *host, int64_t n);
int brams_mlp_set_params(brams_mlp *m, const float *host, int64_t n);

/* Forward pass: X [M][BRAMS_SUR_NIN_PAD] bf16 -> out [M][BRAMS_SUR_NOUT_PAD]
 * f32, M a multiple of 64; synchronous.  Runs the fused inference kernel
 * (all layers in one launch, activations in registers). */
int brams_mlp_forward(brams_mlp *m, const void *X, int64_t M, float *out, void *stream);

/* Forward pass of n <= 4 networks of one shape over the same rows in one
 * launch: out[k] [BRAMS_SUR_NOUT_PAD][M] f32 (species-major, so that a
 * per-cell consumer reads it coalesced) for member k.  Stream-ordered
 * (returns before the kernel ends). */
int brams_mlp_forward_ensemble(brams_mlp *const *members, int n, const void *X, int64_t M,
                               float *const *out, void *stream);

/* Gradient of the mean squared error over the first BRAMS_SUR_NOUT outputs
 * of one batch (B <= max_batch): grads (host, num_params floats) receives
 * dL/dparams, loss (host) the loss.  Parameters are not changed. */
int brams_mlp_gradients(brams_mlp *m, const void *X, const float *Y, int B, float *grads,
                        float *loss);

/* Train on a resident dataset X [N][NIN_PAD] bf16, Y [N][NOUT_PAD] f32 for
 * `epochs` passes in shuffled batches of `batch` rows (the last partial batch
 * of an epoch is dropped), Adam (beta 0.9/0.999, eps 1e-8) with a cosine
 * learning-rate decay from lr0 to 0.  epoch_loss (host, epochs floats, may be
 * NULL) receives the mean training loss of each epoch. */
int brams_mlp_train(brams_mlp *m, const void *X, const float *Y, int64_t N, int batch, int epochs,
                    float lr0, uint64_t seed, float *epoch_loss);
/* Epochs [epoch0, epoch0 + nepochs) of a total_epochs run of brams_mlp_train
 * (same cosine schedule, same per-epoch permutations): a long run split
 * into calls, e.g. to report progress.  epoch_loss gets nepochs values. */
int brams_mlp_train_range(brams_mlp *m, const void *X, const float *Y, int64_t N, int batch,
                          int epoch0, int nepochs, int total_epochs, float lr0, uint64_t seed,
                          float *epoch_loss);

/* ------------------------------------------------------------------ step
 * The surrogate chemistry step on one grid, device-resident arrays as in
 * brams_chem_rodas3_dyndt_device (SYMMETRIC/SEQUENTIAL write-back). */
typedef struct brams_sur_model {
    brams_mlp *members[4];           /* ensemble */
    int32_t nmembers;
    const float *feat_mu, *feat_sd;  /* device [68] feature normalisation */
    const float *targ_mu, *targ_sd;  /* device [47]; targ_sd = 0 marks species the
                                        mechanism never changes (kept as they are) */
    const float *natoms, *catoms;    /* device [47] N and C atoms per molecule */
    double floor_c;                  /* floor (molec/cm^3) of the log10 features */
    double sig_c;                    /* species below this before and after skip the spread check */
    float spread_tol;                /* max relative ensemble spread of any species */
    float n_tol, c_tol;              /* max relative change of the N / C budgets */
} brams_sur_model;

typedef struct brams_sur_stats {
    int64_t cells;
    int64_t fallback_cells;          /* cells integrated by RODAS3 */
    int32_t reasons[4];              /* non-finite, N budget, C budget, spread */
    int64_t fallback_block_attempts; /* RODAS3 attempts summed over fallback blocks */
    int64_t fallback_cell_attempts;  /* attempts x cells, summed over fallback blocks */
    int64_t fallback_cell_substeps;  /* accepted sub-steps x cells, likewise */
    float features_ms, mlp_ms, apply_ms, fallback_ms, total_ms;
    int64_t fallback_failed_blocks;  /* fallback blocks ended by a zero pivot, a non-finite
                                        state or the attempt bound (also the return value) */
    int32_t fallback_wps;            /* waves per SIMD of this call's fallback kernel (0: no
                                        fallback): 2 from 8192 fallback blocks on, else 1; the
                                        forms give bit-identical results */
} brams_sur_stats;

/* Raw features [ncell][BRAMS_SUR_NIN] f32 of the current state (dataset). */
int brams_sur_features_raw(const brams_chem_dev_args *a, float *feat, double floor_c);
/* Raw targets [ncell][BRAMS_SUR_NOUT] f32: log10 increments from sc_p_before
 * (same layout as a->sc_p) to the current state a->sc_p (dataset). */
int brams_sur_targets_raw(const brams_chem_dev_args *a, const float *sc_p_before, float *targ,
                          double floor_c);
/* Column sums and sums of squares (FP64, device) of x [n][ncol] f32. */
int brams_sur_moments(const float *x, int64_t n, int ncol, double *sum, double *sumsq);
/* z-score columns of raw [n][ncol] into out [n][ldo] (bf16 if out_bf16, else f32),
 * padding columns set to 0. */
int brams_sur_normalize(const float *raw, int64_t n, int ncol, const float *mu, const float *sd,
                        void *out, int ldo, int out_bf16);
/* Surrogate + fallback step: features, ensemble inference, checks, write-back
 * of accepted cells, FP64 RODAS3 on the rejected ones.  SYMMETRIC and
 * SEQUENTIAL split methods only (PARALLEL is rejected with an error).
 * Returns 0, the number of fallback blocks with a numerical failure (>0), or
 * <0 on an error. */
int brams_sur_step(const brams_chem_dev_args *a, const brams_sur_model *m, brams_sur_stats *st);
/* Forget the fallback's per-cell memory of the last accepted step; the next
 * brams_sur_step starts it again from the partition blocks'
 * last_accepted_dt.  For a new run or a restored state. */
void brams_sur_reset(void);

#ifdef __cplusplus
}
#endif
#endif
