/*
 * ORACLE -- test infrastructure only (see brams_oracle.h).
 *
 * CPU restatement of src/brams/ccatt/chem_spack_rodas3_dyndt.f90 (and of the
 * ROS2 and QSSA drivers) and of a mechanism's kernels, written to follow the
 * reference's operation order so that the kernels agree bit for bit with the
 * compiled reference.  The mechanism is a build-time choice: ORA_TABLES names
 * its tables (relacs_tuv_tables.h by default; relacs_mx, cb07_tuv and
 * racm_tuv build liboracle_<mech>.so), each pinned against that mechanism's
 * own compiled kernels (oracle/_ref/libbrams_ref_<mech>.so).  The one
 * deliberate difference is the LU pivot order: the reference re-derives a
 * Markowitz order from cell 1 of block 1 on every block-1 attempt
 * (rodas3:470-506, spFactor.c:173-300); this oracle uses one fixed
 * diagonal Markowitz order of the structural pattern (the same rule the GPU
 * code generator uses).  Pivot order changes round-off only; the test suite
 * measures that against the reference's Sparse1.3 through
 * ora_use_external_sparse().
 */
#define _GNU_SOURCE
#include "brams_oracle.h"
#ifndef ORA_TABLES
#define ORA_TABLES "relacs_tuv_tables.h"
#endif
#include ORA_TABLES
#include <math.h>
#include <pthread.h>
#include <stdatomic.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#define NS ORA_NS
#define NR ORA_NR
#define NP ORA_NPHOT

/* ------------------------------------------------------------------ kinetic
 * chem_spack_kinetic.f90:7-435.  One function per rate-law family; each keeps
 * the reference's evaluation order. */

/* Troe fall-off, e.g. kinetic.f90:126-133 (rk 29):
 * Effko=k0 (T/300)^-n, Rapk=kinf (T/300)^-m,
 * k = (Effko M/(1+Effko M/Rapk)) * Fc^(1/(1+log10(Effko M/Rapk)^2)) */
static double ora_troe(double temp, double summ, const double *p)
{
    double effko = p[0] * pow(temp / 3.0e2, -(p[1]));
    double rapk = p[2] * pow(temp / 3.0e2, -(p[3]));
    double l = log10(effko * summ / rapk);
    return (effko * summ / (1.0 + effko * summ / rapk)) * pow(p[4], 1.0 / (1.0 + l * l));
}

void ora_kinetic(double temp, double press, double vapp, const double *jphoto, double *rk)
{
    /* third body and water, kinetic.f90:67,74 */
    const double summ = press * 7.243e16 / temp;
    const double h2o = 29.0 * summ * vapp / (18.0 + 11.0 * vapp);
    for (int r = 0; r < NR; r++) {
        const double *p = ora_params[r];
        double k = 0.0;
        switch (ora_family[r]) {
        case 0: /* never assigned (rk 97, kinetic.f90:356-358): stays 0 */
            k = 0.0;
            break;
        case 1: /* photolysis, kinetic.f90:82-98 */
            k = jphoto[(int)p[0] - 1];
            break;
        case 2: /* constant */
            k = p[0];
            break;
        case 3: /* Arrhenius exp(a - b/T), e.g. kinetic.f90:101 */
            k = exp(p[0] - p[1] / temp);
            break;
        case 4: /* exp(a + n ln T - b/T), kinetic.f90:232-237 */
            k = exp(p[0] + p[1] * log(temp) - p[2] / temp);
            break;
        case 5: { /* sum of c exp(n ln T - b/T) terms plus a constant, kinetic.f90:238-251 */
            int nt = (ora_nparams[r] - 1) / 3;
            double s = 0.0;
            for (int t = 0; t < nt; t++) {
                double c = p[1 + 3 * t], n = p[2 + 3 * t], b = p[3 + 3 * t];
                double x = (n != 0.0) ? n * log(temp) - b / temp : -(b / temp);
                double term = c * exp(x);
                s = (t == 0) ? term : s + term;
            }
            k = s + p[0];
            break;
        }
        case 6: /* (M c)(T/300)^n, kinetic.f90:99 */
            k = summ * p[0] * pow(temp / 3.0e2, p[1]);
            break;
        case 7: /* c1 e^(b1/T) + c2 M e^(b2/T), kinetic.f90:115-119 */
            k = p[0] * exp(p[1] / temp) + p[2] * summ * exp(p[3] / temp);
            break;
        case 8: /* Troe */
            k = ora_troe(temp, summ, p);
            break;
        case 9: { /* HNO3 + OH, kinetic.f90:181-184 */
            double effko = p[0] * exp(p[1] / temp);
            double rapk = p[2] * exp(p[3] / temp);
            double fac = p[4] * exp(p[5] / temp) * summ;
            k = effko + fac / (1.0 + fac / rapk);
            break;
        }
        case 10: /* c (1 + d M), kinetic.f90:229 */
            k = p[0] * (1.0 + p[1] * summ);
            break;
        case 11: /* c1 M e^(b1/T) (1 + c2 H2O e^(b2/T)), CB07_TUV/chem_spack_kinetic.f90:119-121 */
            k = p[0] * summ * exp(p[1] / temp) * (1.0 + p[2] * h2o * exp(p[3] / temp));
            break;
        case 12: /* c H2O H2O, CB07_TUV/chem_spack_kinetic.f90:220 */
            k = p[0] * h2o * h2o;
            break;
        case 13: { /* chemical-activation fall-off, CB07_TUV/chem_spack_kinetic.f90:231-234 */
            double effko = p[0] * pow(temp / 300.0, p[1]);
            double rapk = p[2] * pow(temp / 300.0, p[3]);
            double x = log10(effko / (rapk / summ));
            /* Fc^(1 + (x^2)^-1): the reference's parentheses, not 1/(1 + x^2) */
            k = (effko / (1.0 + effko / (rapk / summ))) * pow(p[4], 1.0 + pow(pow(x, 2.0), -1.0));
            break;
        }
        }
        switch (ora_post[r]) {
        case 1: k = k * summ * ora_post_params[r][0]; break;          /* e.g. kinetic.f90:100 */
        case 2: k = k * h2o; break;                                   /* kinetic.f90:109 */
        case 3: k = ora_troe(temp, summ, ora_post_params[r]) * k; break; /* kinetic.f90:178 */
        case 4: k = k * summ; break;                                  /* CB07_TUV kinetic.f90:273 */
        default: break;
        }
        rk[r] = k;
    }
}

/* ------------------------------------------------------------------ fexchem
 * rates: chem_spack_rates.f90:63-192, w_r = rk_r * y_a * y_b;
 * fexchem: chem_spack_fexchem.f90:73-676, f_s = sum_r c_sr w_r in r order.
 * Species whose row the reference never writes (13, 14, 16) stay 0. */
void ora_fexchem(const double *rk, const double *y, double *f)
{
    double w[NR];
    for (int r = 0; r < NR; r++) {
        double v = rk[r];
        if (ora_react[r][0] >= 0) v = v * y[ora_react[r][0]];
        if (ora_react[r][1] >= 0) v = v * y[ora_react[r][1]];
        w[r] = v;
    }
    for (int s = 0; s < NS; s++) {
        double acc = 0.0;
        for (int t = ora_prod_start[s]; t < ora_prod_start[s + 1]; t++) {
            double term = ora_prod_c[t] * w[ora_prod_r[t]];
            acc = (t == ora_prod_start[s]) ? term : acc + term;
        }
        f[s] = acc;
    }
}

/* ------------------------------------------------------------------ Jacobian
 * dratedc.f90:62-296 + jacdchemdc.f90:68-1191: J_ij = sum over reactions r
 * (ascending) and over each reactant slot of r equal to j of
 * c_ir * rk_r * y_other.  Entries the reference never assigns stay 0. */
typedef struct { double c; int r; int other; } jterm;
static int g_nje;                        /* number of Jacobian entries */
static int g_je_i[NS * NS], g_je_j[NS * NS], g_je_start[NS * NS + 1];
static jterm *g_jt;
static unsigned char g_struct[NS][NS];   /* structural pattern incl. diagonal */

/* LU symbolic data in internal (pivot) order */
static int g_perm[NS];                   /* internal -> species */
static unsigned char g_fill[NS][NS];     /* internal pattern after fill-in */
static int g_nnz_lu;
/* the filled pattern as ascending index lists, so that factor and solve do
 * the dense loops' work in the same order without scanning zeros */
static int g_ucol_n[NS], g_ucol[NS][NS];   /* column J: K < J with fill[K][J] */
static int g_lcol_n[NS], g_lcol[NS][NS];   /* column K: I > K with fill[I][K] */
static int g_urow_n[NS], g_urow[NS][NS];   /* row I: J > I with fill[I][J] */
static int g_nload, g_load_I[NS * NS], g_load_J[NS * NS], g_load_ij[NS * NS];

static pthread_once_t g_once = PTHREAD_ONCE_INIT;

static void ora_init_tables(void)
{
    /* stoichiometric coefficient lookup */
    static double S[NS][NR];
    static unsigned char has[NS][NR];
    memset(S, 0, sizeof S);
    memset(has, 0, sizeof has);
    for (int s = 0; s < NS; s++)
        for (int t = ora_prod_start[s]; t < ora_prod_start[s + 1]; t++) {
            S[s][ora_prod_r[t]] = ora_prod_c[t];
            has[s][ora_prod_r[t]] = 1;
        }
    int cap = 4096, n = 0;
    g_jt = (jterm *)malloc(sizeof(jterm) * cap);
    g_nje = 0;
    for (int i = 0; i < NS; i++)
        for (int j = 0; j < NS; j++) {
            int start = n;
            for (int r = 0; r < NR; r++) {
                if (!has[i][r]) continue;
                for (int slot = 0; slot < 2; slot++) {
                    if (ora_react[r][slot] != j) continue;
                    if (n == cap) { cap *= 2; g_jt = (jterm *)realloc(g_jt, sizeof(jterm) * cap); }
                    g_jt[n].c = S[i][r];
                    g_jt[n].r = r;
                    g_jt[n].other = ora_react[r][1 - slot];
                    n++;
                }
            }
            if (n > start) {
                g_je_i[g_nje] = i;
                g_je_j[g_nje] = j;
                g_je_start[g_nje] = start;
                g_nje++;
            }
        }
    g_je_start[g_nje] = n;

    memset(g_struct, 0, sizeof g_struct);
    for (int e = 0; e < g_nje; e++) g_struct[g_je_i[e]][g_je_j[e]] = 1;
    for (int i = 0; i < NS; i++) g_struct[i][i] = 1;

    /* Markowitz order with diagonal pivots on the structural pattern:
     * smallest (row count-1)*(col count-1) in the active submatrix, ties to
     * the lowest species index; fill-in added as each pivot is eliminated. */
    unsigned char P[NS][NS], active[NS];
    memcpy(P, g_struct, sizeof P);
    memset(active, 1, sizeof active);
    for (int step = 0; step < NS; step++) {
        int best = -1;
        long bestprod = 0;
        for (int c = 0; c < NS; c++) {
            if (!active[c]) continue;
            long rc = -1, cc = -1;
            for (int j = 0; j < NS; j++) if (active[j] && P[c][j]) rc++;
            for (int i = 0; i < NS; i++) if (active[i] && P[i][c]) cc++;
            long prod = rc * cc;
            if (best < 0 || prod < bestprod) { best = c; bestprod = prod; }
        }
        g_perm[step] = best;
        active[best] = 0;
        for (int i = 0; i < NS; i++) {
            if (!active[i] || !P[i][best]) continue;
            for (int j = 0; j < NS; j++)
                if (active[j] && P[best][j]) P[i][j] = 1;
        }
    }
    g_nnz_lu = 0;
    for (int I = 0; I < NS; I++)
        for (int J = 0; J < NS; J++) {
            g_fill[I][J] = P[g_perm[I]][g_perm[J]];
            g_nnz_lu += g_fill[I][J];
        }
    g_nload = 0;
    for (int J = 0; J < NS; J++) {
        g_ucol_n[J] = g_lcol_n[J] = g_urow_n[J] = 0;
        for (int K = 0; K < J; K++)
            if (g_fill[K][J]) g_ucol[J][g_ucol_n[J]++] = K;
        for (int I = J + 1; I < NS; I++)
            if (g_fill[I][J]) g_lcol[J][g_lcol_n[J]++] = I;
        for (int K = J + 1; K < NS; K++)
            if (g_fill[J][K]) g_urow[J][g_urow_n[J]++] = K;
        for (int I = 0; I < NS; I++)
            if (g_fill[I][J]) {
                g_load_I[g_nload] = I;
                g_load_J[g_nload] = J;
                g_load_ij[g_nload++] = g_perm[I] * NS + g_perm[J];
            }
    }
}

static void ora_init(void) { pthread_once(&g_once, ora_init_tables); }

void ora_jacobian(const double *rk, const double *y, double *jac)
{
    ora_init();
    memset(jac, 0, sizeof(double) * NS * NS);
    for (int e = 0; e < g_nje; e++) {
        double acc = 0.0;
        for (int t = g_je_start[e]; t < g_je_start[e + 1]; t++) {
            const jterm *q = &g_jt[t];
            double dw = rk[q->r];
            if (q->other >= 0) dw = dw * y[q->other];
            double term = q->c * dw;
            acc = (t == g_je_start[e]) ? term : acc + term;
        }
        jac[g_je_i[e] * NS + g_je_j[e]] = acc;
    }
}

int ora_lu_order(int *perm, int *nnz_lu)
{
    ora_init();
    for (int k = 0; k < NS; k++) perm[k] = g_perm[k];
    if (nnz_lu) *nnz_lu = g_nnz_lu;
    return NS;
}

/* ------------------------------------------------------------------ LU
 * Numeric refactor with Sparse1.3's arithmetic (spFactor.c:336-401): column by
 * column, U(k,J) = D_k * (1/L_kk) for k<J in row order, each followed by
 * D_i -= U(k,J) L(i,k) for i>k in row order; the diagonal keeps 1/pivot.
 * Solve as spSolve.c:183-218: forward with reciprocal pivots (skipping zero
 * entries), backward with unit-diagonal U.  Works on internal indices. */
typedef struct { double v[NS][NS]; } lumat;

/* Sparse1.3's threshold rule (spFactor.c:202-203 with FindBiggestInColExclude,
 * spFactor.c:2019-2050): a pivot is kept only if RelThreshold (1e-3,
 * spAllocate.c:191) times the largest other entry of its reduced column is
 * below |pivot|.  Counts the pivots of the fixed order that fail it. */
static int threshold_fails(const lumat *m, int J)
{
    if (g_lcol_n[J] == 0) return m->v[J][J] == 0.0;
    double cm = fabs(m->v[g_lcol[J][0]][J]);
    for (int t = 1; t < g_lcol_n[J]; t++) cm = fmax(cm, fabs(m->v[g_lcol[J][t]][J]));
    return cm * 1.0e-3 >= fabs(m->v[J][J]);
}

/* Left-looking in place: column J receives, for K < J in its pattern in
 * ascending order, U(K,J) = A(K,J) / pivot_K (as a product with the stored
 * reciprocal) and A(I,J) -= U(K,J) L(I,K) for I > K in ascending order;
 * then its pivot is replaced by the reciprocal.  Entries outside the filled
 * pattern are never read. */
static int ora_lu_factor(lumat *m, int *viol)
{
    for (int J = 0; J < NS; J++) {
        for (int a = 0; a < g_ucol_n[J]; a++) {
            const int K = g_ucol[J][a];
            const double u = m->v[K][J] * m->v[K][K];
            m->v[K][J] = u;
            for (int b = 0; b < g_lcol_n[K]; b++) {
                const int I = g_lcol[K][b];
                m->v[I][J] -= u * m->v[I][K];
            }
        }
        *viol += threshold_fails(m, J);
        if (m->v[J][J] == 0.0) return J + 1;
        m->v[J][J] = 1.0 / m->v[J][J];
    }
    return 0;
}

static void ora_lu_solve(const lumat *m, const double *rhs, double *x)
{
    double c[NS];
    for (int I = 0; I < NS; I++) c[I] = rhs[g_perm[I]];
    for (int I = 0; I < NS; I++) {
        double t = c[I];
        if (t != 0.0) {
            t *= m->v[I][I];
            c[I] = t;
            for (int b = 0; b < g_lcol_n[I]; b++) {
                const int R = g_lcol[I][b];
                c[R] -= t * m->v[R][I];
            }
        }
    }
    for (int I = NS - 1; I >= 0; I--) {
        double t = c[I];
        for (int b = 0; b < g_urow_n[I]; b++) {
            const int J = g_urow[I][b];
            t -= m->v[I][J] * c[J];
        }
        c[I] = t;
    }
    for (int I = 0; I < NS; I++) x[g_perm[I]] = c[I];
}

/* load a masked species-indexed matrix into internal order (filled pattern) */
static void ora_lu_load(lumat *m, const double *a, const unsigned char *mask)
{
    for (int t = 0; t < g_nload; t++) {
        const int ij = g_load_ij[t];
        m->v[g_load_I[t]][g_load_J[t]] = mask[ij] ? a[ij] : 0.0;
    }
}

/* the matrix 1/(gamma h) I - J of one cell, masked, straight into internal order */
static void ora_lu_load_jac(lumat *m, const double *jac, double igdt, const unsigned char *mask)
{
    for (int t = 0; t < g_nload; t++) {
        const int ij = g_load_ij[t];
        const int i = ij / NS, j = ij % NS;
        const double v = (i == j) ? igdt - jac[ij] : -jac[ij];
        m->v[g_load_I[t]][g_load_J[t]] = mask[ij] ? v : 0.0;
    }
}

int ora_lu_solve_dense(const double *a, const unsigned char *mask, const double *b, double *x)
{
    ora_init();
    lumat m;
    int viol = 0;
    ora_lu_load(&m, a, mask);
    int rc = ora_lu_factor(&m, &viol);
    if (rc) return rc;
    ora_lu_solve(&m, b, x);
    return 0;
}

int ora_lu_threshold_fails(const double *a, const unsigned char *mask)
{
    ora_init();
    lumat m;
    int viol = 0;
    ora_lu_load(&m, a, mask);
    ora_lu_factor(&m, &viol);
    return viol;
}

/* external Sparse1.3 hooks */
static ora_sp_create_fn g_sp_create;
static ora_sp_load_factor_fn g_sp_load_factor;
static ora_sp_solve_fn g_sp_solve;

void ora_use_external_sparse(ora_sp_create_fn c, ora_sp_load_factor_fn lf, ora_sp_solve_fn s)
{
    g_sp_create = c;
    g_sp_load_factor = lf;
    g_sp_solve = s;
}

/* ------------------------------------------------------------------ blocks
 * AllocIndex, chem_spack_utils.f90:93-125: nob = N/block_size (1 if N <=
 * block_size), the remainder added +1 to blocks 1,2,... in turn. */
int ora_alloc_index(int block_size, int m1, int ia, int iz, int ja, int jz,
                    int32_t *block_end, int max_nob)
{
    long ncell = (long)(m1 - 2) * (iz - ia + 1) * (jz - ja + 1);
    if (ncell <= 0 || block_size <= 0) return -1;
    int nob = (block_size >= ncell) ? 1 : (int)(ncell / block_size);
    if (!block_end) return nob;
    if (nob > max_nob) return -2;
    long rest = ncell;
    for (int b = 0; b < nob; b++) { block_end[b] = (block_size >= ncell) ? (int)ncell : block_size; }
    rest = ncell - (long)nob * ((block_size >= ncell) ? ncell : block_size);
    while (rest > 0)
        for (int b = 0; b < nob && rest > 0; b++) { block_end[b]++; rest--; }
    return nob;
}

/* ------------------------------------------------------------------ driver */
typedef struct {
    const brams_chem_args *a;
    int nob;
    long *bstart;                /* first global cell of each block */
    unsigned char mask[NS * NS]; /* value-based sparsity mask of block 1 */
    int64_t *attempts, *substeps;
    int *failed;
    float *forced;
    int *thresh;                 /* per block: threshold-rule pivot failures */
    int order_viol;              /* the same on cell 1 of block 1 */
    const int32_t *list;         /* blocks to run after block 1 (NULL: all) */
    int nlist;
    atomic_int next;
    int has_sparse;              /* external Sparse1.3 in use */
    long sp_handle;
    long *sp_elem;
    int *sp_ipos, *sp_jpos;
    int sp_nnz;
    int solver;                  /* ORA_RODAS3 or ORA_ROS2 */
} drv_t;

enum { ORA_RODAS3 = 0, ORA_ROS2 = 1 };

typedef struct {
    int n;
    long off[128];
    double press[128], temp[128], vapp[128], volmol[128], volmol_i[128];
    double y[128][NS], ynew[128][NS], rk[128][NR], jph[128][NP];
    double f0[128][NS], k1[128][NS], k2[128][NS], k3[128][NS], k4[128][NS], ftmp[NS];
    double err[128];
    double jac[128][NS * NS];
    lumat lu;
} blk_t;

static int64_t *g_last_attempts, *g_last_substeps;
static float *g_last_forced;
static int *g_last_thresh;
static int *g_last_failed;
static int g_last_nob;

static void cell_offset(const brams_chem_args *a, long g, long *off)
{
    long nk = a->m1 - 2, ni = a->iz - a->ia + 1;
    long k = 2 + g % nk;
    long i = a->ia + (g / nk) % ni;
    long j = a->ja + g / (nk * ni);
    *off = (k - 1) + (long)a->m1 * (i - 1) + (long)a->m1 * a->m2 * (j - 1);
}

/* gather and unit conversion, chem_spack_rodas3_dyndt.f90:226-354 */
static void gather(const brams_chem_args *a, blk_t *B)
{
    const float cp = a->cp, cpor = a->cpor, p00 = a->p00;
    for (int c = 0; c < B->n; c++) {
        long o = B->off[c];
        float exner = (a->pp[o] + a->pi0[o]) / cp;            /* REAL arithmetic */
        double press = (double)exner;
        B->temp[c] = (double)a->theta[o] * press;
        press = pow(press, (double)cpor) * (double)p00;
        B->press[c] = press;
        B->vapp[c] = (double)a->rv[o];
        B->volmol[c] = (6.02e23 * 1e-15 * 28.96) * press / (8.314 * B->temp[c]);
        B->volmol_i[c] = 1.0 / B->volmol[c];
        for (int t = 0; t < a->n_not_transported; t++) {
            int n = a->no_transp_chem_index[t] - 1;
            B->y[c][n] = (double)a->sc_p[n][o];
        }
        for (int t = 0; t < a->n_transported; t++) {
            int n = a->transp_chem_index[t] - 1;
            if (a->split_method == BRAMS_SPLIT_PARALLEL && a->n_dyn_chem > 1) {
                float start = a->sc_p[n][o] - a->sc_t_dyn[n][o] * (float)a->n_dyn_chem * a->dtlt;
                B->y[c][n] = (double)start * B->volmol[c] / (double)a->weight[n];
            } else {
                double v = (double)a->sc_p[n][o] * B->volmol[c] / (double)a->weight[n];
                B->y[c][n] = (v > 0.0) ? v : 0.0;
            }
        }
        for (int p = 0; p < NP; p++) B->jph[c][p] = a->jphoto[p + (long)NP * o];
    }
}

/* write-back, chem_spack_rodas3_dyndt.f90:713-798 */
static void scatter(const brams_chem_args *a, blk_t *B)
{
    for (int c = 0; c < B->n; c++) {
        long o = B->off[c];
        for (int t = 0; t < a->n_transported; t++) {
            int n = a->transp_chem_index[t] - 1;
            double conv = B->y[c][n] * (double)a->weight[n] * B->volmol_i[c];
            if (a->split_method == BRAMS_SPLIT_PARALLEL && a->n_dyn_chem > 1) {
                float base = a->sc_t_dyn[n][o] * (float)a->n_dyn_chem * a->dtlt;
                float v = (float)((double)base + conv);
                a->sc_p[n][o] = (v > 0.0f) ? v : 0.0f;
            } else if (a->split_method == BRAMS_SPLIT_PARALLEL) {
                double dti = 1.0 / (double)a->dtlt;
                a->sc_t[n][o] = (float)((double)a->sc_t[n][o] + (conv - (double)a->sc_p[n][o]) * dti);
            } else {
                float v = (float)conv;
                a->sc_p[n][o] = (v > 0.0f) ? v : 0.0f;
            }
        }
        for (int t = 0; t < a->n_not_transported; t++) {
            int n = a->no_transp_chem_index[t] - 1;
            float v = (float)B->y[c][n];
            a->sc_p[n][o] = (v > 0.0f) ? v : 0.0f;
        }
    }
}

static void set_mask_from_cell0(drv_t *D, const double *jac0, double igdt)
{
    /* mask = entries of DLmat(cell 1) that are nonzero, rodas3:470-482 */
    for (int i = 0; i < NS; i++)
        for (int j = 0; j < NS; j++) {
            double v = (i == j) ? igdt - jac0[i * NS + j] : -jac0[i * NS + j];
            D->mask[i * NS + j] = (v != 0.0);
        }
}

/* ------------------------------------------------------------------ RODAS3
 * The four stages of one attempt for one cell, rodas3:523-599 (the Rodas3
 * tableau of Sandu et al. 1997 with gamma = 1/2), with the reference's
 * non-negativity clipping of the stage states.  f and the linear solve with
 * the factored 1/(gamma h) I - J come in as callbacks, so the same code runs
 * the chemistry and the linear test problem of ora_rodas3_linear(). */
typedef void (*stage_fun_fn)(void *ctx, const double *y, double *f);
typedef void (*stage_solve_fn)(void *ctx, const double *rhs, double *x);

static void rodas3_stages(int n, void *ctx, stage_fun_fn fun, stage_solve_fn solve,
                          const double *y0, const double *f0, double dt, double *k1, double *k2,
                          double *k3, double *k4, double *yn)
{
    const double c83 = 8.0 / 3.0;
    const double dti = 1.0 / dt;
    double bb[NS], ft[NS];
    /* stage 1, rodas3:523-527 */
    solve(ctx, f0, k1);
    /* stage 2, rodas3:534-538 */
    for (int s = 0; s < n; s++) bb[s] = (4.0 * dti) * k1[s] + f0[s];
    solve(ctx, bb, k2);
    /* stage 3 with clipping, rodas3:542-566 */
    for (int s = 0; s < n; s++) {
        yn[s] = y0[s] + 2.0 * k1[s];
        if (yn[s] < 0.0) { yn[s] = 0.0; k1[s] = 0.5 * (yn[s] - y0[s]); }
    }
    fun(ctx, yn, ft);
    for (int s = 0; s < n; s++) bb[s] = ft[s] + dti * (k1[s] - k2[s]);
    solve(ctx, bb, k3);
    /* stage 4 with clipping, rodas3:573-599 */
    for (int s = 0; s < n; s++) {
        yn[s] = y0[s] + 2.0 * k1[s] + k3[s];
        if (yn[s] < 0.0) { yn[s] = 0.0; k3[s] = (yn[s] - y0[s]) - 2.0 * k1[s]; }
    }
    fun(ctx, yn, ft);
    for (int s = 0; s < n; s++) bb[s] = ft[s] + dti * (k1[s] - k2[s] - c83 * k3[s]);
    solve(ctx, bb, k4);
}

/* Solution y1 = max(y0 + 2 k1 + k3 + k4, 0) and error norm (k4 is the error
 * estimate), rodas3:609-648.  Returns the raw sum of squares. */
static double rodas3_solution(int n, const double *y0, const double *k1, const double *k3,
                              const double *k4, const double *atol, const double *rtol,
                              double *y1, double *err)
{
    double e = 0.0;
    for (int s = 0; s < n; s++) {
        double v = y0[s] + 2.0 * k1[s] + k3[s] + k4[s];
        v = fmax(v, 0.0);
        y1[s] = v;
        double tol = atol[s] + rtol[s] * fmax(fabs(y0[s]), fabs(v));
        double q = k4[s] / tol;
        e = e + q * q;
    }
    *err = fmax(1.0e-15, sqrt(e / n));
    return e;
}

/* New step from the block's largest error, rodas3:654-662:
 * dt * min(6, max(0.2, 0.9 err^-1/3)) clamped to [dt_min, dt_max]. */
static double rodas3_dt_new(double dt, double max_err, double dt_min, double dt_max)
{
    double fac = fmin(6.0, fmax(0.2, 0.9 / pow(max_err, 1.0 / 3.0)));
    double dt_new = dt * fac;
    return fmax(dt_min, fmin(dt_max, dt_new));
}

double ora_dt_new(double dt, double max_err, double dt_min, double dt_max)
{
    return rodas3_dt_new(dt, max_err, dt_min, dt_max);
}

/* ------------------------------------------------------------------ ROS2
 * chem_spack_ros_dyndt.f90 (CHEMISTRY=3): the two-stage Rosenbrock method
 * with gamma = 1 + 1/sqrt(2); the reference writes sqrt(2) as the double
 * value of its single-precision rounding (ros_dyndt:130). */
static const double ros2_gamma = 1.0 + 1.0 / 1.4142135381698608;

/* The two stages and the solution of one attempt for one cell,
 * ros_dyndt:404-477 (no clipping); returns the error sum of squares with
 * err_k = 0.5/gamma (k1 + k2), ros_dyndt:486-498. */
static double ros2_attempt(int n, void *ctx, stage_fun_fn fun, stage_solve_fn solve, const double *y0,
                           const double *f0, double dt, const double *atol, const double *rtol,
                           double *k1, double *k2, double *yn, double *err)
{
    const double gi = 1.0 / ros2_gamma;
    const double dti = 1.0 / dt;
    double ft[NS], bb[NS];
    solve(ctx, f0, k1);                                         /* ros_dyndt:404-410 */
    for (int s = 0; s < n; s++) yn[s] = y0[s] + gi * k1[s];     /* ros_dyndt:418-424 */
    fun(ctx, yn, ft);                                           /* ros_dyndt:433-437 */
    for (int s = 0; s < n; s++) bb[s] = ft[s] - 2.0 * gi * dti * k1[s];   /* ros_dyndt:440-445 */
    solve(ctx, bb, k2);                                         /* ros_dyndt:454-460 */
    double e = 0.0;
    for (int s = 0; s < n; s++) {
        double v = y0[s] + gi * (1.5 * k1[s] + 0.5 * k2[s]);    /* ros_dyndt:463-477 */
        yn[s] = v;
        double tol = atol[s] + rtol[s] * fmax(fabs(y0[s]), fabs(v));
        double q = (0.5 * gi * (k1[s] + k2[s])) / tol;
        e = e + q * q;
    }
    *err = fmax(1.0e-15, sqrt(e / n));
    return e;
}

/* New step, ros_dyndt:513-517: dt * min(6, max(0.2, 0.9 err^-1/2)) clamped */
static double ros2_dt_new(double dt, double max_err, double dt_min, double dt_max)
{
    double fac = fmin(6.0, fmax(0.2, 0.9 / pow(max_err, 1.0 / 2.0)));
    double dt_new = dt * fac;
    return fmax(dt_min, fmin(dt_max, dt_new));
}

double ora_ros2_dt_new(double dt, double max_err, double dt_min, double dt_max)
{
    return ros2_dt_new(dt, max_err, dt_min, dt_max);
}

typedef struct { drv_t *D; lumat *lu; const double *rk; } stage_ctx_t;

static void stage_fun(void *vctx, const double *y, double *f)
{
    stage_ctx_t *ctx = (stage_ctx_t *)vctx;
    ora_fexchem(ctx->rk, y, f);
}

static void stage_solve(void *vctx, const double *rhs, double *x)
{
    stage_ctx_t *ctx = (stage_ctx_t *)vctx;
    if (ctx->D->has_sparse) {
        double tmp[NS];
        memcpy(tmp, rhs, sizeof tmp);
        g_sp_solve(ctx->D->sp_handle, tmp, x);
    } else {
        ora_lu_solve(ctx->lu, rhs, x);
    }
}

/* Linear test problem y' = A y (n <= 47, dense A row-major) integrated over
 * nsteps fixed steps of size h with the same stage code; A is factored as
 * 1/(gamma h) I - A by dense elimination.  err_out receives the error norm of
 * the last step with atol = rtol = 1 (may be NULL). */
typedef struct { int n; const double *A; double lu[NS * NS]; int piv[NS]; } lin_ctx_t;

static void lin_fun(void *vctx, const double *y, double *f)
{
    lin_ctx_t *c = (lin_ctx_t *)vctx;
    for (int i = 0; i < c->n; i++) {
        double s = 0.0;
        for (int j = 0; j < c->n; j++) s += c->A[i * c->n + j] * y[j];
        f[i] = s;
    }
}

static void lin_solve(void *vctx, const double *rhs, double *x)
{
    lin_ctx_t *c = (lin_ctx_t *)vctx;
    int n = c->n;
    double b[NS];
    for (int i = 0; i < n; i++) b[i] = rhs[c->piv[i]];
    for (int i = 0; i < n; i++)
        for (int k = 0; k < i; k++) b[i] -= c->lu[i * n + k] * b[k];
    for (int i = n - 1; i >= 0; i--) {
        for (int k = i + 1; k < n; k++) b[i] -= c->lu[i * n + k] * b[k];
        b[i] /= c->lu[i * n + i];
    }
    for (int i = 0; i < n; i++) x[i] = b[i];
}

int ora_rodas3_linear(int n, const double *A, const double *y0, double h, int nsteps,
                      double *y_out, double *err_out)
{
    if (n < 1 || n > NS) return -1;
    lin_ctx_t c;
    c.n = n;
    c.A = A;
    /* dense LU with partial pivoting of 1/(gamma h) I - A */
    for (int i = 0; i < n; i++) {
        c.piv[i] = i;
        for (int j = 0; j < n; j++) c.lu[i * n + j] = (i == j ? 1.0 / (0.5 * h) : 0.0) - A[i * n + j];
    }
    for (int k = 0; k < n; k++) {
        int p = k;
        for (int i = k + 1; i < n; i++)
            if (fabs(c.lu[i * n + k]) > fabs(c.lu[p * n + k])) p = i;
        if (c.lu[p * n + k] == 0.0) return 1;
        if (p != k) {
            for (int j = 0; j < n; j++) {
                double t = c.lu[k * n + j]; c.lu[k * n + j] = c.lu[p * n + j]; c.lu[p * n + j] = t;
            }
            int t = c.piv[k]; c.piv[k] = c.piv[p]; c.piv[p] = t;
        }
        for (int i = k + 1; i < n; i++) {
            c.lu[i * n + k] /= c.lu[k * n + k];
            for (int j = k + 1; j < n; j++) c.lu[i * n + j] -= c.lu[i * n + k] * c.lu[k * n + j];
        }
    }
    double y[NS], f0[NS], k1[NS], k2[NS], k3[NS], k4[NS], yn[NS], ones[NS], err = 0.0;
    for (int i = 0; i < n; i++) { y[i] = y0[i]; ones[i] = 1.0; }
    for (int s = 0; s < nsteps; s++) {
        lin_fun(&c, y, f0);
        rodas3_stages(n, &c, lin_fun, lin_solve, y, f0, h, k1, k2, k3, k4, yn);
        rodas3_solution(n, y, k1, k3, k4, ones, ones, yn, &err);
        for (int i = 0; i < n; i++) y[i] = yn[i];
    }
    for (int i = 0; i < n; i++) y_out[i] = y[i];
    if (err_out) *err_out = err;
    return 0;
}

/* ROS2 on the linear test problem y' = A y, as ora_rodas3_linear */
int ora_ros2_linear(int n, const double *A, const double *y0, double h, int nsteps,
                    double *y_out, double *err_out)
{
    if (n < 1 || n > NS) return -1;
    lin_ctx_t c;
    c.n = n;
    c.A = A;
    for (int i = 0; i < n; i++) {
        c.piv[i] = i;
        for (int j = 0; j < n; j++) c.lu[i * n + j] = (i == j ? 1.0 / (ros2_gamma * h) : 0.0) - A[i * n + j];
    }
    for (int k = 0; k < n; k++) {
        int p = k;
        for (int i = k + 1; i < n; i++)
            if (fabs(c.lu[i * n + k]) > fabs(c.lu[p * n + k])) p = i;
        if (c.lu[p * n + k] == 0.0) return 1;
        if (p != k) {
            for (int j = 0; j < n; j++) {
                double t = c.lu[k * n + j]; c.lu[k * n + j] = c.lu[p * n + j]; c.lu[p * n + j] = t;
            }
            int t = c.piv[k]; c.piv[k] = c.piv[p]; c.piv[p] = t;
        }
        for (int i = k + 1; i < n; i++) {
            c.lu[i * n + k] /= c.lu[k * n + k];
            for (int j = k + 1; j < n; j++) c.lu[i * n + j] -= c.lu[i * n + k] * c.lu[k * n + j];
        }
    }
    double y[NS], f0[NS], k1[NS], k2[NS], yn[NS], ones[NS], err = 0.0;
    for (int i = 0; i < n; i++) { y[i] = y0[i]; ones[i] = 1.0; }
    for (int s = 0; s < nsteps; s++) {
        lin_fun(&c, y, f0);
        ros2_attempt(n, &c, lin_fun, lin_solve, y, f0, h, ones, ones, k1, k2, yn, &err);
        for (int i = 0; i < n; i++) y[i] = yn[i];
    }
    for (int i = 0; i < n; i++) y_out[i] = y[i];
    if (err_out) *err_out = err;
    return 0;
}

/* Test hooks of the controller (block 1 only, single-threaded use):
 * g_script replaces the block's largest error of attempt i by
 * g_script[min(i, n-1)] (every cell gets that error, so the accept test sees
 * it too); g_trace records each attempt's (dt, accepted, time after). */
static const double *g_script;
static int g_nscript;
#define ORA_TRACE_MAX 4096
static double g_trace_dt[ORA_TRACE_MAX], g_trace_t[ORA_TRACE_MAX];
static int g_trace_acc[ORA_TRACE_MAX], g_ntrace;
static atomic_long g_jac_calls, g_f0_calls;

void ora_set_error_script(const double *errs, int n)
{
    g_script = errs;
    g_nscript = (errs && n > 0) ? n : 0;
}

int ora_controller_trace(double *dt, int *accepted, double *time_after, int max)
{
    int n = g_ntrace < max ? g_ntrace : max;
    for (int i = 0; i < n; i++) {
        dt[i] = g_trace_dt[i];
        accepted[i] = g_trace_acc[i];
        time_after[i] = g_trace_t[i];
    }
    return g_ntrace;
}

void ora_call_counts(int64_t *jacobian_calls, int64_t *f0_calls)
{
    *jacobian_calls = atomic_load(&g_jac_calls);
    *f0_calls = atomic_load(&g_f0_calls);
}

/* RODAS3 on one block, chem_spack_rodas3_dyndt.f90:413-705 */
static int integrate_block(drv_t *D, blk_t *B, int b)
{
    const brams_chem_args *a = D->a;
    const float tn = a->dtlt * (float)a->n_dyn_chem;     /* REAL product */
    double dt = a->last_accepted_dt[b];
    const double dt_min = fmax(1.0, 1.0e-2 * (double)tn);
    const double dt_max = (double)tn;
    double time_c = 0.0;
    const double time_f = (double)tn;
    int failed = 0, viol = 0;
    int64_t natt = 0, nsub = 0;
    float forced = 0.f;
    if (b == 0) g_ntrace = 0;

    for (int c = 0; c < B->n; c++)
        ora_kinetic(B->temp[c], B->press[c], B->vapp[c], B->jph[c], B->rk[c]);

    while (time_c + 1.0e-8 < time_f) {
        for (int c = 0; c < B->n; c++) {
            ora_jacobian(B->rk[c], B->y[c], B->jac[c]);
            ora_fexchem(B->rk[c], B->y[c], B->f0[c]);
        }
        atomic_fetch_add(&g_jac_calls, B->n);
        atomic_fetch_add(&g_f0_calls, B->n);
        for (;;) {
            natt++;
            const double igdt = 1.0 / (0.5 * dt);
            if (b == 0) {
                set_mask_from_cell0(D, B->jac[0], igdt);
                if (D->has_sparse) {
                    /* new matrix + ordering on cell-1 values each attempt, rodas3:470-506 */
                    int nz = 0;
                    double vals[NS * NS];
                    for (int i = 0; i < NS; i++)
                        for (int j = 0; j < NS; j++)
                            if (D->mask[i * NS + j]) {
                                D->sp_ipos[nz] = i + 1;
                                D->sp_jpos[nz] = j + 1;
                                vals[nz] = (i == j) ? igdt - B->jac[0][i * NS + j] : -B->jac[0][i * NS + j];
                                nz++;
                            }
                    D->sp_nnz = nz;
                    D->sp_handle = g_sp_create(NS, nz, D->sp_ipos, D->sp_jpos, D->sp_elem);
                    g_sp_load_factor(D->sp_handle, nz, D->sp_elem, vals);
                }
            }
            for (int c = 0; c < B->n; c++) {
                if (D->has_sparse) {
                    double mat[NS * NS];
                    for (int i = 0; i < NS; i++)
                        for (int j = 0; j < NS; j++)
                            mat[i * NS + j] = (i == j) ? igdt - B->jac[c][i * NS + j] : -B->jac[c][i * NS + j];
                    double vals[NS * NS];
                    for (int z = 0; z < D->sp_nnz; z++)
                        vals[z] = mat[(D->sp_ipos[z] - 1) * NS + D->sp_jpos[z] - 1];
                    if (g_sp_load_factor(D->sp_handle, D->sp_nnz, D->sp_elem, vals) != 0) failed = 1;
                } else {
                    int v = 0;
                    ora_lu_load_jac(&B->lu, B->jac[c], igdt, D->mask);
                    if (ora_lu_factor(&B->lu, &v) != 0) failed = 1;
                    viol += v;
                    if (b == 0 && c == 0) D->order_viol += v;
                }
                stage_ctx_t ctx = {D, &B->lu, B->rk[c]};
                rodas3_stages(NS, &ctx, stage_fun, stage_solve, B->y[c], B->f0[c], dt,
                              B->k1[c], B->k2[c], B->k3[c], B->k4[c], B->ynew[c]);
            }
            /* solution and error, rodas3:609-648 */
            double max_err = 0.0;
            int all_accepted = 1;
            for (int c = 0; c < B->n; c++) {
                double e = rodas3_solution(NS, B->y[c], B->k1[c], B->k3[c], B->k4[c], a->atol, a->rtol,
                                           B->ynew[c], &B->err[c]);
                if (!isfinite(e)) failed = 1;
                if (b == 0 && g_nscript > 0)
                    B->err[c] = g_script[(natt - 1) < g_nscript ? (natt - 1) : g_nscript - 1];
                if (B->err[c] - 1.0e-8 > 1.0) all_accepted = 0;
                if (c == 0 || B->err[c] > max_err) max_err = B->err[c];
            }
            /* controller, rodas3:650-700 */
            double dt_new = rodas3_dt_new(dt, max_err, dt_min, dt_max);
            if (failed) { time_c = time_f; break; }
            if (!all_accepted && dt_new > dt_min) {
                if (b == 0 && g_ntrace < ORA_TRACE_MAX) {
                    g_trace_dt[g_ntrace] = dt; g_trace_acc[g_ntrace] = 0; g_trace_t[g_ntrace++] = time_c;
                }
                dt = dt_new;
                continue;
            }
            /* error above 1 accepted at dt_min, rodas3:671-676 */
            if (!all_accepted && max_err > forced) forced = (float)max_err;
            if (b == 0 && g_ntrace < ORA_TRACE_MAX) {
                g_trace_dt[g_ntrace] = dt; g_trace_acc[g_ntrace] = all_accepted ? 1 : 2;
                g_trace_t[g_ntrace++] = time_c + dt;
            }
            time_c = time_c + dt;
            dt = fmin(dt_new, time_f - time_c);
            if (time_c < time_f) a->last_accepted_dt[b] = dt_new;
            for (int c = 0; c < B->n; c++) memcpy(B->y[c], B->ynew[c], sizeof(double) * NS);
            nsub++;
            break;
        }
    }
    D->attempts[b] = natt;
    D->substeps[b] = nsub;
    D->failed[b] = failed;
    D->forced[b] = forced;
    D->thresh[b] = viol;
    return failed;
}

/* ROS2 on one block, chem_spack_ros_dyndt.f90:340-555.  The matrix is
 * 1/(gamma h) I - J on every structural entry (get_always_zeros,
 * chem_spack_ros.f90:578-627, keeps the nonzeros of J + I at a dummy state,
 * which are the structural ones); with the reference's Sparse1.3 the matrix
 * is created once and its first factorisation orders (chem_spack_solve_sparse.f90). */
static int integrate_block_ros2(drv_t *D, blk_t *B, int b)
{
    const brams_chem_args *a = D->a;
    const float tn = a->dtlt * (float)a->n_dyn_chem;     /* REAL product */
    double dt = a->last_accepted_dt[b];
    const double dt_min = fmax(1.0, 1.0e-2 * (double)tn);
    const double dt_max = (double)tn;
    double time_c = 0.0;
    const double time_f = (double)tn;
    int failed = 0, viol = 0;
    int64_t natt = 0, nsub = 0;
    float forced = 0.f;

    for (int c = 0; c < B->n; c++)
        ora_kinetic(B->temp[c], B->press[c], B->vapp[c], B->jph[c], B->rk[c]);

    while (time_c + 1.0e-8 < time_f) {                   /* ros_dyndt:348 */
        for (int c = 0; c < B->n; c++) {                 /* ros_dyndt:351-362, once per sub-step */
            ora_jacobian(B->rk[c], B->y[c], B->jac[c]);
            ora_fexchem(B->rk[c], B->y[c], B->f0[c]);
        }
        atomic_fetch_add(&g_jac_calls, B->n);
        atomic_fetch_add(&g_f0_calls, B->n);
        for (;;) {
            natt++;
            const double igdt = 1.0 / (ros2_gamma * dt);   /* ros_dyndt:387 */
            for (int c = 0; c < B->n; c++) {
                if (D->has_sparse) {
                    if (!D->sp_handle) {
                        int nz = 0;
                        for (int i = 0; i < NS; i++)
                            for (int j = 0; j < NS; j++)
                                if (D->mask[i * NS + j]) { D->sp_ipos[nz] = i + 1; D->sp_jpos[nz] = j + 1; nz++; }
                        D->sp_nnz = nz;
                        D->sp_handle = g_sp_create(NS, nz, D->sp_ipos, D->sp_jpos, D->sp_elem);
                    }
                    double vals[NS * NS];
                    for (int z = 0; z < D->sp_nnz; z++) {
                        int i = D->sp_ipos[z] - 1, j = D->sp_jpos[z] - 1;
                        vals[z] = (i == j) ? igdt - B->jac[c][i * NS + j] : -B->jac[c][i * NS + j];
                    }
                    if (g_sp_load_factor(D->sp_handle, D->sp_nnz, D->sp_elem, vals) != 0) failed = 1;
                } else {
                    int v = 0;
                    ora_lu_load_jac(&B->lu, B->jac[c], igdt, D->mask);
                    if (ora_lu_factor(&B->lu, &v) != 0) failed = 1;
                    viol += v;
                }
                stage_ctx_t ctx = {D, &B->lu, B->rk[c]};
                double e = ros2_attempt(NS, &ctx, stage_fun, stage_solve, B->y[c], B->f0[c], dt, a->atol,
                                        a->rtol, B->k1[c], B->k2[c], B->ynew[c], &B->err[c]);
                if (!isfinite(e)) failed = 1;
            }
            double max_err = 0.0;
            int all_accepted = 1;
            for (int c = 0; c < B->n; c++) {             /* ros_dyndt:501-510 */
                if (B->err[c] - 1.0e-8 > 1.0) all_accepted = 0;
                if (c == 0 || B->err[c] > max_err) max_err = B->err[c];
            }
            double dt_new = ros2_dt_new(dt, max_err, dt_min, dt_max);
            if (failed) { time_c = time_f; break; }
            if (!all_accepted && dt_new > dt_min) { dt = dt_new; continue; }   /* ros_dyndt:523-526 */
            if (!all_accepted && max_err > forced) forced = (float)max_err;
            time_c = time_c + dt;                        /* ros_dyndt:528-548 */
            dt = fmin(dt_new, time_f - time_c);
            if (time_c < time_f) a->last_accepted_dt[b] = dt_new;
            for (int c = 0; c < B->n; c++) memcpy(B->y[c], B->ynew[c], sizeof(double) * NS);
            nsub++;
            break;
        }
    }
    D->attempts[b] = natt;
    D->substeps[b] = nsub;
    D->failed[b] = failed;
    D->forced[b] = forced;
    D->thresh[b] = viol;
    return failed;
}

static void run_block(drv_t *D, blk_t *B, int b)
{
    B->n = D->a->block_end[b];
    for (int c = 0; c < B->n; c++) cell_offset(D->a, D->bstart[b] + c, &B->off[c]);
    gather(D->a, B);
    if (D->solver == ORA_ROS2)
        integrate_block_ros2(D, B, b);
    else
        integrate_block(D, B, b);
    scatter(D->a, B);
}

static void *worker(void *arg)
{
    drv_t *D = (drv_t *)arg;
    blk_t *B = (blk_t *)malloc(sizeof(blk_t));
    for (;;) {
        int k = atomic_fetch_add(&D->next, 1);
        int b;
        if (D->list) {
            if (k >= D->nlist) break;
            b = D->list[k];
            if (b <= 0 || b >= D->nob) continue;   /* block 1 ran first; ignore bad ids */
        } else {
            if (k >= D->nob) break;
            b = k;
        }
        run_block(D, B, b);
    }
    free(B);
    return NULL;
}

static int run_step(const brams_chem_args *a, int nthreads, brams_chem_stats *st,
                    unsigned char *mask_out, const int32_t *list, int nlist, int solver)
{
    ora_init();
    if (a->nspecies != NS || a->nr != NR || a->nr_photo != NP) return -1;
    if (a->photoj_method == BRAMS_PHOTOJ_LUT) return -2;
    if (a->maxblock_size > 128) return -3;
    drv_t *D = (drv_t *)calloc(1, sizeof(drv_t));
    D->a = a;
    D->nob = a->nob;
    D->list = list;
    D->nlist = nlist;
    D->bstart = (long *)malloc(sizeof(long) * (a->nob + 1));
    D->attempts = (int64_t *)calloc(a->nob, sizeof(int64_t));
    D->substeps = (int64_t *)calloc(a->nob, sizeof(int64_t));
    D->failed = (int *)calloc(a->nob, sizeof(int));
    D->forced = (float *)calloc(a->nob, sizeof(float));
    D->thresh = (int *)calloc(a->nob, sizeof(int));
    D->bstart[0] = 0;
    for (int b = 0; b < a->nob; b++) D->bstart[b + 1] = D->bstart[b] + a->block_end[b];
    D->has_sparse = (g_sp_create && g_sp_load_factor && g_sp_solve);
    D->sp_elem = (long *)malloc(sizeof(long) * NS * NS);
    D->sp_ipos = (int *)malloc(sizeof(int) * NS * NS);
    D->sp_jpos = (int *)malloc(sizeof(int) * NS * NS);
    D->solver = solver;
    if (solver == ORA_ROS2) memcpy(D->mask, g_struct, sizeof D->mask);   /* every structural entry */
    atomic_store(&g_jac_calls, 0);
    atomic_store(&g_f0_calls, 0);

    /* block 1 first: it fixes the mask (and, in the reference, the order) */
    blk_t *B = (blk_t *)malloc(sizeof(blk_t));
    run_block(D, B, 0);
    free(B);
    if (mask_out) memcpy(mask_out, D->mask, NS * NS);

    atomic_store(&D->next, list ? 0 : 1);
    if (nthreads <= 0) nthreads = (int)sysconf(_SC_NPROCESSORS_ONLN);
    if (D->has_sparse) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    pthread_t th[256];
    for (int t = 0; t < nthreads; t++) pthread_create(&th[t], NULL, worker, D);
    for (int t = 0; t < nthreads; t++) pthread_join(th[t], NULL);

    int nfail = 0;
    int64_t att = 0, sub = 0, cells = 0, viol = 0;
    char *ran = (char *)calloc(a->nob, 1);
    ran[0] = 1;
    if (list) {
        for (int k = 0; k < nlist; k++)
            if (list[k] > 0 && list[k] < a->nob) ran[list[k]] = 1;
    } else {
        memset(ran, 1, a->nob);
    }
    for (int b = 0; b < a->nob; b++) {
        if (!ran[b]) continue;
        nfail += D->failed[b];
        att += D->attempts[b];
        sub += D->substeps[b];
        cells += a->block_end[b];
        viol += D->thresh[b];
    }
    if (st) {
        memset(st, 0, sizeof *st);
        st->cells = cells;
        for (int b = 0; b < a->nob; b++) st->blocks += ran[b];
        st->substeps = sub;
        st->attempts = att;
        st->failed_blocks = nfail;
        for (int b = 0; b < a->nob; b++) {
            if (!ran[b]) continue;
            if (D->forced[b] > 1.0f) st->forced_blocks++;
            if (D->forced[b] > st->max_forced_err) st->max_forced_err = D->forced[b];
            st->cell_attempts += D->attempts[b] * a->block_end[b];
            st->cell_substeps += D->substeps[b] * a->block_end[b];
        }
        int nz = 0;
        for (int i = 0; i < NS * NS; i++) nz += D->mask[i];
        st->mask_nnz = nz;
        st->threshold_pivots = viol;
        st->order_threshold_pivots = D->order_viol;
    }
    free(ran);
    free(g_last_attempts);
    free(g_last_substeps);
    free(g_last_forced);
    free(g_last_thresh);
    free(g_last_failed);
    g_last_attempts = D->attempts;
    g_last_substeps = D->substeps;
    g_last_forced = D->forced;
    g_last_thresh = D->thresh;
    g_last_failed = D->failed;
    g_last_nob = a->nob;
    free(D->bstart);
    free(D->sp_elem);
    free(D->sp_ipos);
    free(D->sp_jpos);
    free(D);
    return nfail;
}

int ora_rodas3_dyndt(const brams_chem_args *a, int nthreads, brams_chem_stats *st,
                     unsigned char *mask_out)
{
    return run_step(a, nthreads, st, mask_out, NULL, 0, ORA_RODAS3);
}

int ora_rodas3_dyndt_blocks(const brams_chem_args *a, int nthreads, brams_chem_stats *st,
                            unsigned char *mask_out, const int32_t *blocks, int nblocks)
{
    static const int32_t none[1] = {0};
    return run_step(a, nthreads, st, mask_out, blocks ? blocks : none, blocks ? nblocks : 0, ORA_RODAS3);
}

int ora_ros2_dyndt(const brams_chem_args *a, int nthreads, brams_chem_stats *st, unsigned char *mask_out)
{
    return run_step(a, nthreads, st, mask_out, NULL, 0, ORA_ROS2);
}

int ora_ros2_dyndt_blocks(const brams_chem_args *a, int nthreads, brams_chem_stats *st,
                          unsigned char *mask_out, const int32_t *blocks, int nblocks)
{
    static const int32_t none[1] = {0};
    return run_step(a, nthreads, st, mask_out, blocks ? blocks : none, blocks ? nblocks : 0, ORA_ROS2);
}

int ora_block_counts(int64_t *attempts, int64_t *substeps, int max_nob)
{
    int n = g_last_nob < max_nob ? g_last_nob : max_nob;
    for (int b = 0; b < n; b++) {
        if (attempts) attempts[b] = g_last_attempts[b];
        if (substeps) substeps[b] = g_last_substeps[b];
    }
    return g_last_nob;
}

int ora_block_forced(float *forced, int max_nob)
{
    int n = g_last_nob < max_nob ? g_last_nob : max_nob;
    for (int b = 0; b < n; b++) forced[b] = g_last_forced[b];
    return g_last_nob;
}

int ora_block_thresholds(int32_t *counts, int max_nob)
{
    int n = g_last_nob < max_nob ? g_last_nob : max_nob;
    for (int b = 0; b < n; b++) counts[b] = g_last_thresh[b];
    return g_last_nob;
}

int ora_block_failed(int32_t *failed, int max_nob)
{
    int n = g_last_nob < max_nob ? g_last_nob : max_nob;
    for (int b = 0; b < n; b++) failed[b] = g_last_failed[b];
    return g_last_nob;
}

/* One cell's gather (rodas3:226-354) and write-back (rodas3:713-798) alone,
 * for unit tests of the conversions: cell is the global interior index in
 * j -> i -> k order. */
int ora_gather_cell(const brams_chem_args *a, long cell, double *y, double *temp, double *press,
                    double *volmol)
{
    blk_t *B = (blk_t *)malloc(sizeof(blk_t));
    if (!B) return -1;
    B->n = 1;
    cell_offset(a, cell, &B->off[0]);
    gather(a, B);
    memcpy(y, B->y[0], sizeof(double) * NS);
    *temp = B->temp[0];
    *press = B->press[0];
    *volmol = B->volmol[0];
    free(B);
    return 0;
}

int ora_scatter_cell(const brams_chem_args *a, long cell, const double *y)
{
    blk_t *B = (blk_t *)malloc(sizeof(blk_t));
    if (!B) return -1;
    B->n = 1;
    cell_offset(a, cell, &B->off[0]);
    gather(a, B);                 /* volmol of the cell */
    memcpy(B->y[0], y, sizeof(double) * NS);
    scatter(a, B);
    free(B);
    return 0;
}

/* ------------------------------------------------------------------ QSSA
 * The QSSA solver of chem_spack_qssa.f90 (CHEMISTRY=1).
 *
 * Production and loss of the P - L c split (chem_spack_qssa.f90:247-265):
 * rates.f90 w_r = rk_r y_a y_b; dratedc.f90 dw_rs = rk_r y_other, where the
 * other reactant of a self-reaction is s itself (dratedc assigns that value
 * twice) and a one-reactant rate has dw_rs = rk_r; fexprod.f90 P_s = the
 * positive stoichiometric terms c_sr w_r, fexloss.f90 L_s = -c_sr dw_rs over
 * the negative ones, both in reaction order.  Rows the reference never
 * assigns (no term) are 0. */
void ora_prod_loss(const double *rk, const double *y, double *prod, double *loss)
{
    double w[NR];
    for (int r = 0; r < NR; r++) {
        double v = rk[r];
        if (ora_react[r][0] >= 0) v = v * y[ora_react[r][0]];
        if (ora_react[r][1] >= 0) v = v * y[ora_react[r][1]];
        w[r] = v;
    }
    for (int s = 0; s < NS; s++) {
        double p = 0.0, l = 0.0;
        int np = 0, nl = 0;
        for (int t = ora_prod_start[s]; t < ora_prod_start[s + 1]; t++) {
            const int r = ora_prod_r[t];
            const double c = ora_prod_c[t];
            if (c > 0.0) {
                double term = c * w[r];
                p = np++ ? p + term : term;
            } else {
                /* dw_rs exists only where s is a reactant of r; a negative
                 * product coefficient (CB07_TUV) gives a fexloss term on a dw
                 * that dratedc never assigns, 0 here */
                double dw = 0.0;
                if (ora_react[r][0] == s || ora_react[r][1] == s) {
                    dw = rk[r];
                    if (ora_react[r][1] >= 0) dw = dw * y[ora_react[r][0] == s ? ora_react[r][1] : ora_react[r][0]];
                }
                double term = (-c) * dw;
                l = nl++ ? l + term : term;
            }
        }
        prod[s] = p;
        loss[s] = l;
    }
}

/* One species through itchim QSSA sub-steps of length delt with P and L held
 * for the whole call (solveur, chem_spack_qssa.f90:343-348, and qssa,
 * :368-411, Chang et al. 1987).  cvmgm_dp / cvmgp_dp / cvmgz_dp
 * (:421-477) are the selections written out below.  qssa:404 picks the
 * explicit step when 100 L delt <= 1, but qssa:406 assigns the species again
 * from cg1 and cg2 alone, so that choice never takes effect: the result is
 * the explicit step when L = 0 and the exponential (or, for L delt >= 10, the
 * steady-state) form otherwise. */
double ora_qssa_species(double y, double prod, double loss, double delt, int itchim)
{
    for (int it = 0; it < itchim; it++) {
        double ybid = (-loss < 0.0) ? loss : 1.0;
        double cg1 = y + delt * (prod - loss * y);
        double pr = prod / ybid;
        double pr1 = pr + (y - pr) * exp(-ybid * delt);
        double pr2 = ybid * delt - 10.0;
        double cg2 = (pr2 >= 0.0) ? pr : pr1;
        y = (loss == 0.0) ? cg1 : cg2;
    }
    return y;
}

/* gather of chem_spack_qssa.f90:161-195: the volmol of this driver (R as a
 * REAL literal), every species converted from ppbm with the REAL product
 * 6.02e23*1e-15 and pmar = 28.96 as a REAL literal (:100).  J is taken for
 * FAST-TUV as for FAST-JX (:197-210 copies it for FAST-JX only and leaves it
 * unset otherwise). */
static void qssa_gather(const brams_chem_args *a, long o, double *y, double *temp, double *press,
                        double *vapp, double *volmol, double *jph)
{
    const float cp = a->cp, cpor = a->cpor, p00 = a->p00;
    const double pmar = (double)28.96f;
    const double avo = (double)(6.02e23f * 1e-15f);
    float exner = (a->pp[o] + a->pi0[o]) / cp;
    double p = (double)exner;
    *temp = (double)a->theta[o] * p;
    p = pow(p, (double)cpor) * (double)p00;
    *press = p;
    *vapp = (double)a->rv[o];
    *volmol = p / ((double)8.314f * *temp);
    for (int n = 0; n < NS; n++)
        y[n] = (double)a->sc_p[n][o] * (avo * *volmol) * pmar / (double)a->weight[n];
    for (int q = 0; q < NP; q++) jph[q] = a->jphoto[q + (long)NP * o];
}

typedef struct {
    const brams_chem_args *a;
    int itchim;
    long ncell;
    atomic_long next;
    atomic_long nfatal;
} qssa_drv_t;

static void *qssa_worker(void *vd)
{
    qssa_drv_t *D = (qssa_drv_t *)vd;
    const brams_chem_args *a = D->a;
    const double delt = (double)(a->dtlt / (float)D->itchim);
    const double dtlt_i = (double)(1.0f / a->dtlt);
    const double pmar = (double)28.96f;
    const double avo = (double)(6.02e23f * 1e-15f);
    double y[NS], rk[NR], jph[NP], prod[NS], loss[NS];
    for (;;) {
        long c0 = atomic_fetch_add(&D->next, 64);
        if (c0 >= D->ncell) break;
        long c1 = c0 + 64 < D->ncell ? c0 + 64 : D->ncell;
        for (long g = c0; g < c1; g++) {
            long o;
            double temp, press, vapp, volmol;
            cell_offset(a, g, &o);
            qssa_gather(a, o, y, &temp, &press, &vapp, &volmol, jph);
            ora_kinetic(temp, press, vapp, jph, rk);
            ora_prod_loss(rk, y, prod, loss);
            int fatal = 0;
            for (int n = 0; n < NS; n++) {
                double v = ora_qssa_species(y[n], prod[n], loss[n], delt, D->itchim);
                if (v < 0.0) v = 0.0;               /* solveur:352-354 */
                if (v >= 1.0e20) fatal = 1;         /* solveur:357-361 stops the model */
                y[n] = v;
            }
            if (fatal) {
                atomic_fetch_add(&D->nfatal, 1);
                continue;
            }
            /* chem_spack_qssa.f90:273-292: tendency of every species */
            for (int n = 0; n < NS; n++) {
                double fxc = 1.0 / (avo * volmol * (pmar / (double)a->weight[n]));
                a->sc_t[n][o] = (float)((double)a->sc_t[n][o] + (y[n] * fxc - (double)a->sc_p[n][o]) * dtlt_i);
            }
        }
    }
    return NULL;
}

/* The whole QSSA step (chem_qssa, chem_spack_qssa.f90:37-321) with itchim
 * sub-steps (chemistry.f90:393 passes 75): sc_t of every species receives
 * (new - old) / dtlt; sc_p is not changed.  Cells are independent.  Returns
 * the number of cells where a species reached 1e20 (solveur stops the model
 * there); their tendencies are left unchanged. */
long ora_qssa(const brams_chem_args *a, int itchim, int nthreads)
{
    if (!a || itchim <= 0 || a->nspecies != NS || !a->sc_t) return -1;
    for (int n = 0; n < NS; n++)
        if (!a->sc_t[n]) return -1;
    qssa_drv_t D;
    D.a = a;
    D.itchim = itchim;
    D.ncell = (long)(a->m1 - 2) * (a->iz - a->ia + 1) * (a->jz - a->ja + 1);
    atomic_init(&D.next, 0);
    atomic_init(&D.nfatal, 0);
    if (nthreads <= 0) nthreads = (int)sysconf(_SC_NPROCESSORS_ONLN);
    if (nthreads > 256) nthreads = 256;
    pthread_t th[256];
    for (int t = 0; t < nthreads; t++) pthread_create(&th[t], NULL, qssa_worker, &D);
    for (int t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
    return atomic_load(&D.nfatal);
}

/* One cell's QSSA gather alone (test hook). */
int ora_qssa_gather_cell(const brams_chem_args *a, long cell, double *y, double *temp, double *press,
                         double *volmol)
{
    long o;
    double vapp, jph[NP];
    cell_offset(a, cell, &o);
    qssa_gather(a, o, y, temp, press, &vapp, volmol, jph);
    return 0;
}

/* ------------------------------------------------------------------ photolysis
 * The FAST-TUV output mapping of tuvDriver (ModTuvDriver2.7.f90:1302-1332)
 * in the reference's loop order: for k = 1..kj with xref(k) /= 0, over the
 * interior columns, jphoto(xref(k), m+1, i, j) = valJLoc(i, j, k, m) for
 * m = 1..m1-1 and jphoto(xref(k), 1, i, j) = jphoto(xref(k), 2, i, j); then
 * the mechanism's relations over the whole array (:1333-1368; RELACS: J14
 * and J16), the constants as REAL literals.  valj is valJLoc(ia:iz, ja:jz, kj, kz), Fortran order;
 * jphoto (np, m1, m2, m3).  Returns -1 on invalid arguments. */
int ora_tuv_jphoto(int m1, int m2, int m3, int ia, int iz, int ja, int jz, int kj, int kz,
                   const double *valj, const int32_t *xref, int np, double *jphoto)
{
    if (np != NP || kz < m1 - 1 || kj < 1) return -1;
    const long ni = iz - ia + 1, nj = jz - ja + 1;
#define JP(n, k, i, j) jphoto[(n) + (long)np * ((k) + (long)m1 * ((i) + (long)m2 * (j)))]
    for (int k = 0; k < kj; k++) {
        if (xref[k] == 0) continue;
        const int n = xref[k] - 1;
        if (n < 0 || n >= np) return -1;
        for (int j = ja - 1; j < jz; j++)
            for (int i = ia - 1; i < iz; i++) {
                for (int m = 1; m <= m1 - 1; m++)
                    JP(n, m, i, j) = valj[(i - (ia - 1)) + ni * ((j - (ja - 1)) + nj * (k + (long)kj * (m - 1)))];
                JP(n, 0, i, j) = JP(n, 1, i, j);
            }
    }
    /* the relations of the build's mechanism, each a whole-array statement
     * in the reference's order */
#define ALL for (int j = 0; j < m3; j++) for (int i = 0; i < m2; i++) for (int k = 0; k < m1; k++)
#if ORA_MECH_ID == 1   /* RELACS, :1363-1368 */
    ALL JP(13, k, i, j) = (double)0.962055f * JP(12, k, i, j) + (double)0.0106247f * JP(8, k, i, j);
    ALL JP(15, k, i, j) = (12.0 * JP(12, k, i, j)) + (208.0 * JP(14, k, i, j));
#elif ORA_MECH_ID == 2 /* RACM, :1349-1361 */
    ALL JP(14, k, i, j) = (double)0.28f * JP(8, k, i, j);
    ALL JP(19, k, i, j) = (double)0.2f * JP(0, k, i, j);
    ALL JP(13, k, i, j) = JP(12, k, i, j);
    ALL JP(15, k, i, j) = (JP(15, k, i, j) + (double)8.7e-07f) / 2.0;
    ALL JP(22, k, i, j) = JP(15, k, i, j);
#elif ORA_MECH_ID == 3 /* CB07, :1335-1339 */
    ALL JP(14, k, i, j) = JP(13, k, i, j);
#endif
#undef ALL
#undef JP
    return 0;
}
