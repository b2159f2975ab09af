/*
 * brams_chem.h -- C-ABI of the MI355X CCATT gas-phase chemistry step.
 *
 * Drop-in boundary for BRAMS's RODAS3 chemistry solver (CHEMISTRY=4,
 * mechanism RELACS_TUV; libbrams_chem_<mech>.so gives the same entry points
 * for RELACS_MX, CB07_TUV and RACM_TUV, see brams_chem_mechanism), with the
 * ROS2 (CHEMISTRY=3) and QSSA (CHEMISTRY=1) solvers beside it.  The Fortran
 * module mod_chem_spack_rodas3_dyndt in brams_amd/fortran/ keeps the
 * reference interface of
 *   chem_rodas3_dyndt   src/brams/ccatt/chem_spack_rodas3_dyndt.f90:41-131
 * and calls brams_chem_rodas3_dyndt() below through ISO_C_BINDING.
 *
 * Conventions
 *   - Grid arrays are Fortran column-major (k,i,j) with extents (m1,m2,m3);
 *     jphoto is (nr_photo,m1,m2,m3).  Species follow the mechanism order
 *     (src/brams/ccatt/RELACS_TUV/chem1_list.f90:63-109, O3=1 ... XO2=47).
 *   - Transported species are in ppbm, non-transported ones in molec/cm^3
 *     (chem_spack_rodas3_dyndt.f90:261-317).
 *   - Indices passed in (ia..jz, transp_chem_index, ...) are 1-based like the
 *     Fortran caller's.
 *   - Every call is synchronous and returns a status: 0 ok, >0 count of
 *     blocks with a numerical failure (zero pivot or non-finite state, the
 *     reference would abort there: src/utils/lib/spSolve.c:158-164), <0 a
 *     fatal error whose text brams_chem_last_error() returns.
 *   - One call at a time per process (the reference keeps module state).
 */
#ifndef BRAMS_CHEM_H
#define BRAMS_CHEM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BRAMS_CHEM_NSPECIES 47
#define BRAMS_CHEM_NREACT 128
#define BRAMS_CHEM_NPHOTO 17

/* split_method (chem_spack_rodas3_dyndt.f90:62, chemistry.f90 SPLIT_METHOD) */
enum { BRAMS_SPLIT_SYMMETRIC = 0, BRAMS_SPLIT_SEQUENTIAL = 1, BRAMS_SPLIT_PARALLEL = 2 };
/* PhotojMethod (chem1_list.f90:784); LUT is rejected (not the RELACS_TUV path) */
enum { BRAMS_PHOTOJ_FAST_TUV = 0, BRAMS_PHOTOJ_FAST_JX = 1, BRAMS_PHOTOJ_LUT = 2 };

/*
 * Arguments of one chemistry step on one grid, host memory.  Field for field
 * the arguments of chem_rodas3_dyndt (chem_spack_rodas3_dyndt.f90:41-131);
 * arguments the RELACS_TUV path never reads (dn0, rcp, cosz, att, spack,
 * spack_2d, extra3d, get_non_zeros, indexk/i/j, kij_index) are not carried.
 */
typedef struct brams_chem_args {
    int32_t m1, m2, m3;          /* grid extents (k,i,j) */
    int32_t ia, iz, ja, jz;      /* interior columns, 1-based inclusive */
    float dtlt;                  /* dynamics step (s) */
    int32_t n_dyn_chem;          /* N_DYN_CHEM: chemistry step = dtlt*N */
    int32_t split_method;        /* BRAMS_SPLIT_* */
    int32_t photoj_method;       /* BRAMS_PHOTOJ_* */
    float cp, cpor, p00;         /* rconstants */
    const float *pp, *pi0, *theta, *rv;   /* (m1,m2,m3) */
    const double *jphoto;        /* (nr_photo,m1,m2,m3) */
    int32_t nspecies, nr, nr_photo;
    const float *weight;         /* (nspecies) molar weights */
    int32_t n_transported, n_not_transported;
    const int32_t *transp_chem_index;     /* (n_transported), 1-based species */
    const int32_t *no_transp_chem_index;  /* (n_not_transported) */
    float **sc_p;                /* nspecies pointers to (m1,m2,m3) state, in/out */
    float **sc_t;                /* nspecies pointers to (m1*m2*m3) tendencies (PARALLEL, N=1), else NULL */
    const float *const *sc_t_dyn;/* nspecies pointers (PARALLEL, N>1), else NULL */
    int32_t nob;                 /* number of blocks (AllocIndex) */
    int32_t maxblock_size;
    const int32_t *block_end;    /* (nob) cells per block */
    double *last_accepted_dt;    /* (nob) in/out, per-block step carried across calls */
    const double *atol, *rtol;   /* (nspecies) */
} brams_chem_args;

/*
 * The same step with every array already resident in device memory (HBM).
 * Species state is one contiguous buffer [nspecies][m1*m2*m3] (species-major);
 * sc_t / sc_t_dyn likewise.  The block partition is the AllocIndex one with
 * block size 64 (the reference's block_size_ros, chemistry.f90:275); the
 * library derives it from the grid.  last_accepted_dt lives on device.
 */
typedef struct brams_chem_dev_args {
    int32_t m1, m2, m3, ia, iz, ja, jz;
    float dtlt;
    int32_t n_dyn_chem, split_method, photoj_method;
    float cp, cpor, p00;
    const float *pp, *pi0, *theta, *rv;
    const double *jphoto;
    const float *weight;         /* device (nspecies) */
    const int32_t *transported;  /* device (nspecies) 1 = transported */
    float *sc_p;                 /* device [nspecies][m1*m2*m3] */
    float *sc_t;                 /* device or NULL */
    const float *sc_t_dyn;       /* device or NULL */
    int32_t nob;                 /* must equal brams_chem_alloc_index(64, ...) of this grid */
    double *last_accepted_dt;    /* device (nob) */
    const double *atol, *rtol;   /* device (nspecies) */
    void *stream;                /* hipStream_t, NULL = default stream */
} brams_chem_dev_args;

/* Per-call counters (no reference counterpart; the reference prints nothing). */
typedef struct brams_chem_stats {
    int64_t cells;               /* cells integrated */
    int64_t blocks;              /* blocks integrated */
    int64_t substeps;            /* accepted sub-steps, summed over blocks */
    int64_t attempts;            /* attempts (accepted + rejected), summed over blocks */
    int64_t failed_blocks;       /* blocks with a zero pivot or non-finite state */
    int32_t mask_nnz;            /* entries kept by the value-based sparsity mask */
    float kernel_ms;             /* device time of the last step, all kernels */
    int64_t cell_substeps;       /* sum over cells of their block's accepted sub-steps */
    int64_t cell_attempts;       /* sum over cells of their block's attempts */
    float block1_ms;             /* device time of the block-1 kernel (mask, serial prefix) */
    float main_ms;               /* device time of the wave64 kernel over the 64-cell blocks */
    int64_t main_cells;          /* cells that kernel integrated */
    int64_t forced_blocks;       /* blocks that accepted a step with error > 1 at dt_min */
    float max_forced_err;        /* largest such error (rodas3:671-676 accepts it) */
    int64_t threshold_pivots;    /* pivots, over the cells and attempts of block 1 and of the
                                    65..127-cell blocks, that Sparse1.3's threshold rule would
                                    reject: |pivot| <= RelThreshold (1e-3) x the largest other
                                    entry of its reduced column (spFactor.c:202-203,
                                    spAllocate.c:191).  Sparse applies the rule only where it
                                    orders, on block 1; the 64-cell blocks do not count */
    int64_t order_threshold_pivots; /* the same count on cell 1 of block 1, the matrix the
                                    reference orders on every block-1 attempt
                                    (chem_spack_rodas3_dyndt.f90:470-506): 0 means Sparse's
                                    threshold test accepts every pivot of the fixed order */
    int32_t kernel_form;         /* kernel form that integrated the 64-cell blocks (see
                                    brams_chem_set_form; -1: none, or several grids) */
} brams_chem_stats;

/* Select the HIP device and load the mechanism tables.  Idempotent. */
int brams_chem_init(int device);

/*
 * Drop-in for chem_rodas3_dyndt (chem_spack_rodas3_dyndt.f90:41): integrate
 * every interior cell of one grid over dtlt*N_DYN_CHEM with RODAS3, updating
 * sc_p (or sc_t), last_accepted_dt in place.  Host arrays; the library copies
 * them to and from the device.
 */
int brams_chem_rodas3_dyndt(const brams_chem_args *args);

/* Device-resident variant of brams_chem_rodas3_dyndt (same semantics). */
int brams_chem_rodas3_dyndt_device(const brams_chem_dev_args *args);

/*
 * Drop-in for chem_ros_dyndt (chem_spack_ros_dyndt.f90:39-45, CHEMISTRY=3):
 * the same step with the two-stage Rosenbrock method of that file (gamma =
 * 1 + 1/sqrt(2), error order 2), on the full structural Jacobian (no value
 * mask) and without clipping of the stages.  Arguments, units, status and
 * write-back as brams_chem_rodas3_dyndt; threshold counts are kept for the
 * 65..127-cell blocks only.  Host arrays.
 */
int brams_chem_ros_dyndt(const brams_chem_args *args);

/* Device-resident variant of brams_chem_ros_dyndt (same semantics). */
int brams_chem_ros_dyndt_device(const brams_chem_dev_args *args);

/*
 * Drop-in for chem_qssa (chem_spack_qssa.f90:37-43, CHEMISTRY=1; called at
 * chemistry.f90:399-407 with itchim = 75): every interior cell on its own,
 * P and L of the P - L c split (rates, dratedc, fexprod, fexloss) from the
 * state, then itchim sub-steps of length dtlt/itchim of Chang's formula with
 * P and L held (solveur/qssa, :324-411), negatives set to 0.  Every species
 * is read from sc_p in ppbm (non-transported ones too, :184-191) and
 * receives sc_t += (new - old)/dtlt (:273-292; sc_t must be given for every
 * species); sc_p is not changed.  split_method, N_DYN_CHEM, the block
 * partition and the tolerances are not used.  J is used for FAST-TUV as for
 * FAST-JX (the reference copies it for FAST-JX only).  Returns the number of
 * cells where a species reached 1e20 (where solveur stops the model, :357-361;
 * their tendencies are left unchanged), or a negative error code.  Host arrays.
 */
int brams_chem_qssa(const brams_chem_args *args, int32_t itchim);

/* Device-resident variant of brams_chem_qssa (same semantics; the species
 * are [47][m1*m2*m3] device arrays in sc_p and sc_t). */
int brams_chem_qssa_device(const brams_chem_dev_args *args, int32_t itchim);

/*
 * No counterpart in the reference.  brams_chem_rodas3_dyndt_device for
 * callers that cannot allocate device memory (code under stream capture,
 * tight serving loops, harnesses that forbid hipMalloc): the same step
 * through the same kernels, with the per-call bookkeeping (block starts,
 * mask words, per-block counters, the failure count) in a fixed device
 * array of the library (about 0.8 MB, one copy per device, held until the
 * process ends) instead of allocated buffers.  It calls no allocator.
 * At most 32768 blocks (2 097 152 cells); a larger grid is refused (-7)
 * before any device work.  Calls on one device share that array, so calls
 * that overlap on one device race.  Blocks run in partition order;
 * brams_chem_get_stats and brams_chem_block_status do not report these calls.
 * Synchronises the stream; returns the number of failed blocks (>= 0) or
 * an error code (< 0).
 */
int brams_chem_rodas3_dyndt_device_noalloc(const brams_chem_dev_args *args);

/*
 * The FAST-TUV output mapping of tuvDriver (ModTuvDriver2.7.f90:1302-1332) and
 * RELACS's J relations (:1363-1368) on device arrays, the photolysis input of
 * the chemistry step (called every radfrq, chemistry.f90:338-375).  For every
 * TUV reaction k (ascending) with xref[k] /= 0, over the interior columns:
 * jphoto(xref[k], m+1, i, j) = valJLoc(i, j, k, m) for m = 1..m1-1 and
 * jphoto(xref[k], 1, i, j) = jphoto(xref[k], 2, i, j); then at every point
 * J14 = 0.962055 J13 + 0.0106247 J9 and J16 = 12 J13 + 208 J15 (REAL-literal
 * constants, each product and sum rounded).  Entries no reaction maps keep
 * their values.  valj is valJLoc (:593, :672, :1263), the double copy of
 * TUV's j-values; xref is TUV's reaction cross-reference (read from
 * <mechanism>xreference.dat by ModTuv2.7.f90:17419-17426).  Stream-ordered;
 * returns when the kernels have run (their time is in brams_chem_get_stats).
 */
typedef struct brams_tuv_map_args {
    int32_t m1, m2, m3, ia, iz, ja, jz;
    int32_t kj, kz;             /* TUV reactions and levels of valj (ModTuv2.7.f90:79-86: 80, 125) */
    const double *valj;         /* valJLoc(ia:iz, ja:jz, kj, kz), Fortran order, device */
    const int32_t *xref;        /* [kj] host: jphoto index (1-based) of TUV reaction k, 0 = unused */
    int32_t nr_photo;           /* photolysis rates of the mechanism (17) */
    double *jphoto;             /* (nr_photo, m1, m2, m3), device, updated in place */
    void *stream;               /* hipStream_t, NULL for the default stream */
} brams_tuv_map_args;

int brams_chem_tuv_jphoto_device(const brams_tuv_map_args *args);

/*
 * chem_accum (chemistry.f90:528-604) on device arrays, for a model that keeps
 * the tracers in HBM.  With the PARALLEL split and N_DYN_CHEM > 1 the driver
 * accumulates the mean dynamic tendency of the transported species over a
 * chemistry step: ntask 1 (every dynamics step) adds REAL(1./N_DYN_CHEM) *
 * sc_t to sc_t_dyn, any other ntask (after the chemistry step) sets sc_t_dyn
 * to zero.  sc_t, sc_t_dyn: device [nspecies][m1*m2*m3] (sc_t may be NULL
 * when ntask != 1); transp_chem_index: host, 1-based, n_transported entries.
 * Stream-ordered; does not synchronise.
 */
int brams_chem_accum_device(int32_t m1, int32_t m2, int32_t m3, int32_t ntask, int32_t n_dyn_chem,
                            int32_t n_transported, const int32_t *transp_chem_index, float *sc_t_dyn,
                            const float *sc_t, void *stream);

/*
 * The nested grids of one model step in one pass.  chemistry_driver calls
 * chem_rodas3_dyndt once per grid (chemistry.f90:313-320); this integrates
 * grids[0..ngrids) together: block 1 of every grid first (each publishes its
 * own grid's mask), then every other block of every grid in shared
 * launches, the 64-cell blocks longest first across the grids from the
 * second call on.  Each grid keeps its own partition, mask,
 * last_accepted_dt and counters, so the results equal those of ngrids
 * brams_chem_rodas3_dyndt_device calls.  1 <= ngrids <= 16, grids of
 * distinct extents, one stream (all grids' stream fields equal).
 * brams_chem_get_stats then holds the sums over the grids (mask_nnz: the
 * first grid's) and brams_chem_block_status the last grid's blocks.
 * Synchronises the stream; returns the total number of failed blocks.
 */
int brams_chem_rodas3_dyndt_device_grids(const brams_chem_dev_args *grids, int32_t ngrids);

/*
 * Block partition of one grid, AllocIndex (chem_spack_utils.f90:39-183):
 * interior cells in j->i->k order cut into floor(N/block_size) blocks, the
 * remainder spread +1 over the first blocks.  Writes block_end[0..nob-1] when
 * block_end is non-NULL and max_nob is large enough; returns nob (<0 on error).
 */
int brams_chem_alloc_index(int32_t block_size, int32_t m1, int32_t ia, int32_t iz,
                           int32_t ja, int32_t jz, int32_t *block_end, int32_t max_nob);

/*
 * The mechanism the library is built for and its sizes (no reference
 * counterpart: a BRAMS build compiles one mechanism's chem1_list.f90, e.g.
 * src/brams/ccatt/RELACS_TUV/chem1_list.f90:11-14).  libbrams_chem.so holds
 * RELACS_TUV (the BRAMS_CHEM_N* values above); libbrams_chem_<mech>.so holds
 * RELACS_MX, CB07_TUV or RACM_TUV, whose brams_chem_args.nspecies, nr and
 * nr_photo must equal what this returns.  Any pointer may be NULL.
 */
const char *brams_chem_mechanism(int32_t *nspecies, int32_t *nr, int32_t *nr_photo);

/*
 * dt history checkpoint.  AllocIndex starts every block's step-size history
 * at dtlt*N_DYN_CHEM when a run starts (chem_spack_utils.f90:170-172), and its
 * comment there plans to keep that history in the history (restart) file so
 * that a restarted run reproduces the uninterrupted one; the reference never
 * writes it.  These calls do: pack writes one grid's last_accepted_dt with
 * the grid, block size and mechanism it belongs to and a checksum into a
 * byte record that the caller stores with its history file; unpack checks a
 * record against the grid it is restored into and writes last_accepted_dt
 * back.  A step's results depend on the species and this array only, so the
 * two restore a run exactly.  Host memory only (a device caller copies the
 * nob doubles across); no device work.  No reference counterpart.
 */
typedef struct brams_dt_history_grid {
  int32_t m1, m2, m3, ia, iz, ja, jz;  /* extents, 1-based interior bounds */
  int32_t block_size;                  /* AllocIndex's block size (64 here) */
} brams_dt_history_grid;
/* Bytes of the record of a grid of nob blocks (80 + 8*nob). */
int64_t brams_chem_dt_history_bytes(int32_t nob);
/* Writes the record into buf[0..cap); returns its size.  -8: nob is not the
 * grid's block count; -7: a dt is not finite and positive. */
int64_t brams_chem_dt_history_pack(const brams_dt_history_grid *grid, const double *last_accepted_dt,
                                   int32_t nob, void *buf, int64_t cap);
/* Checks buf[0..nbytes) and writes last_accepted_dt[0..nob); returns nob.
 * -7: not a record, another version, corrupt (checksum) or a bad dt; -8: a
 * record of another grid, block size or mechanism. */
int brams_chem_dt_history_unpack(const void *buf, int64_t nbytes, const brams_dt_history_grid *grid,
                                 double *last_accepted_dt, int32_t max_nob);

/* Counters of the last step. */
int brams_chem_get_stats(brams_chem_stats *out);

/* Per-block status of the last step (host arrays, any may be NULL): attempts,
 * accepted sub-steps, failure flag, largest forced-acceptance error.  Returns
 * nob. */
int brams_chem_block_status(int32_t *attempts, int32_t *substeps, int32_t *failed,
                            float *forced_err, int32_t max_nob);

/* Per-block count of threshold-rule pivot failures of the last step (see
 * brams_chem_stats.threshold_pivots; 0 for the 64-cell blocks, which do not
 * count); no reference counterpart.  Returns nob. */
int brams_chem_block_thresholds(int32_t *counts, int32_t max_nob);

/*
 * Kernel form of the 64-cell blocks (no reference counterpart).  The forms
 * give bit-identical results and differ in speed: 3, 4 and 5 are the
 * per-lane kernel with the register budget of 1, 2 and 4 waves per SIMD;
 * 0..2 (the slab layouts: per-cell working set in registers, LDS and a
 * per-slot scratch slab, persistent workgroups over a device work queue) are
 * switched off in this build and refused (-6).  -1 (default): the first call
 * of each grid times every available form on a scratch copy of its state and
 * keeps the fastest for the grid.  The ROS2 step has forms 3 and 5 and tunes
 * its own choice the same way; with form 4 forced its calls fail (-6).  A
 * lean library (libbrams_chem_<mech>.so) has forms 3 and 5 only and refuses
 * form 4 here (-6).
 */
int brams_chem_set_form(int32_t form);
/* The current grid's chosen form and the autotuner's timings (ms, one per
 * form, 0 if not timed) into ms[0..n); returns the number of forms. */
int brams_chem_get_tuning(int32_t *form, float *ms, int32_t n);

/* Text of the last error (thread-local, never NULL). */
const char *brams_chem_last_error(void);

/* Release device buffers. */
void brams_chem_finalize(void);

#ifdef __cplusplus
}
#endif
#endif /* BRAMS_CHEM_H */
