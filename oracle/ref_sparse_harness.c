/*
 * C-callable harness around the reference's Sparse1.3 Fortran bridge,
 * compiled with the reference's own Sparse sources by oracle/Makefile into
 * oracle/_ref/libbrams_ref.so.  Test infrastructure only.
 *
 * It drives Sparse exactly as the RODAS3 driver does
 * (src/brams/ccatt/chem_spack_rodas3_dyndt.f90:470-517): create a matrix,
 * fetch one element handle per mask entry in row-major order, then per cell
 * clear + add values + factor, and solve.  The first factor of a fresh
 * matrix runs Sparse's Markowitz ordering on those values.
 */
#define spINSIDE_SPARSE
#include "spConfig.h"
#include "spmatrix.h"
#include "spDefs.h"

/* Fortran-callable entry points defined in src/utils/lib/spFortran.c:140-180 */
extern long sfcreate_solve_(int *Size, int *Complex, int *Error);
extern long sfgetelement_(long *Matrix, int *Row, int *Col);
extern void sfzero_(long *Matrix);
extern void sfadd1real_(long *Element, double *Real);
extern int sffactor_(long *Matrix);
extern void sfsolve_(long *Matrix, double *RHS, double *Solution);
extern void sfdestroy_(long *Matrix);

long ref_sp_create(int n, int nnz, const int *ipos, const int *jpos, long *elements)
{
    int complex_t = 0, err = 0, k;
    long h = sfcreate_solve_(&n, &complex_t, &err);
    if (err != 0 || h == 0) return 0;
    for (k = 0; k < nnz; k++) {
        int r = ipos[k], c = jpos[k];
        elements[k] = sfgetelement_(&h, &r, &c);
    }
    return h;
}

int ref_sp_load_factor(long h, int nnz, const long *elements, const double *vals)
{
    int k;
    sfzero_(&h);
    for (k = 0; k < nnz; k++) {
        long e = elements[k];
        double v = vals[k];
        sfadd1real_(&e, &v);
    }
    return sffactor_(&h);
}

void ref_sp_solve(long h, double *rhs, double *x)
{
    sfsolve_(&h, rhs, x);
}

/* Internal-to-external row and column maps chosen by the ordering (1-based). */
int ref_sp_order(long h, int *row_map, int *col_map, int *fillins)
{
    MatrixPtr m = (MatrixPtr)h;
    int i;
    for (i = 1; i <= m->Size; i++) {
        row_map[i - 1] = m->IntToExtRowMap[i];
        col_map[i - 1] = m->IntToExtColMap[i];
    }
    *fillins = m->Fillins;
    return m->Size;
}

void ref_sp_destroy(long h)
{
    sfdestroy_(&h);
}
