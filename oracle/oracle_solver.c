/*
 * oracle_solver.c -- CPU oracle linear algebra (FP64), TEST INFRASTRUCTURE ONLY.
 * Sparse kernels on the oracle's CSC matrix (Solver.jl:40-50 reads CSC column i
 * as row i because K is symmetric; here the generic CSC product is used).
 */
#include "pfem_oracle.h"

#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* y = K x for a CSC matrix (any symmetry). */
void or_csc_spmv(int64_t n, const int64_t *colptr, const int64_t *rowval,
                 const double *nzval, const double *x, double *y) {
    memset(y, 0, sizeof(double) * n);
    for (int64_t c = 0; c < n; ++c) {
        double xc = x[c];
        for (int64_t k = colptr[c]; k < colptr[c + 1]; ++k) y[rowval[k]] += nzval[k] * xc;
    }
}

/* Row-parallel product reading CSC column i as row i (SymThreadedK mul!,
 * Solver.jl:40-50): y = K^T x, which equals K x for a symmetric K. */
void or_csc_spmv_omp(int64_t n, const int64_t *colptr, const int64_t *rowval,
                     const double *nzval, const double *x, double *y, int nthreads) {
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        double s = 0.0;
        for (int64_t k = colptr[i]; k < colptr[i + 1]; ++k) s += nzval[k] * x[rowval[k]];
        y[i] = s;
    }
}
