/*
 * oracle_solver.c -- CPU oracle linear algebra (FP64), TEST INFRASTRUCTURE ONLY.
 *
 * PCG preconditioned by one V-cycle of smoothed-aggregation AMG, the algorithm of the
 * reference's linear solve (Solver.jl:235-290: Krylov.cg! with
 * AlgebraicMultigrid.smoothed_aggregation(K; B = rigid-body modes) and the Jacobi(2/3)
 * smoother option of Solver.jl:138), restated on scalar CSR matrices with OpenMP:
 *   aggregation    three-phase aggregation of mesh nodes (Vanek, Mandel & Brezina 1996),
 *                  all structural couplings strong
 *   tentative P    per-aggregate QR of the rigid-body modes (Solver.jl:117-129)
 *   smoothing      P = (I - w D^-1 A) Pt, w = 4/(3 rho(D^-1 A)), scalar diagonal D
 *   coarse ops     R A P with R = P^T (Gustavson row-wise SpGEMM)
 *   V-cycle        one damped-Jacobi (2/3) pre- and post-sweep, dense LU at the bottom
 *   PCG            x0 = 0, stop when sqrt(r'z) <= rtol sqrt(r0'z0)
 * Used by the CPU baseline and by the tests that pin it against direct solves.
 */
#include "pfem_oracle.h"

#include <math.h>
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

/* ------------------------------------------------------------------ CSR */
typedef struct csr {
    int64_t nr, nc;
    int64_t *ptr;
    int64_t *col;
    double *val;
} csr;

static void csr_free(csr *a) {
    free(a->ptr); free(a->col); free(a->val);
    memset(a, 0, sizeof *a);
}

static void csr_spmv(const csr *a, const double *x, double *y) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < a->nr; ++i) {
        double s = 0.0;
        for (int64_t k = a->ptr[i]; k < a->ptr[i + 1]; ++k) s += a->val[k] * x[a->col[k]];
        y[i] = s;
    }
}

/* transpose (counting sort); rows of the result come out column-sorted */
static csr csr_transpose(const csr *a) {
    csr t = {a->nc, a->nr, calloc(a->nc + 1, sizeof(int64_t)), NULL, NULL};
    int64_t nnz = a->ptr[a->nr];
    t.col = malloc(sizeof(int64_t) * (nnz ? nnz : 1));
    t.val = malloc(sizeof(double) * (nnz ? nnz : 1));
    for (int64_t k = 0; k < nnz; ++k) t.ptr[a->col[k] + 1]++;
    for (int64_t c = 0; c < a->nc; ++c) t.ptr[c + 1] += t.ptr[c];
    int64_t *f = malloc(sizeof(int64_t) * (a->nc + 1));
    memcpy(f, t.ptr, sizeof(int64_t) * (a->nc + 1));
    for (int64_t i = 0; i < a->nr; ++i)
        for (int64_t k = a->ptr[i]; k < a->ptr[i + 1]; ++k) {
            int64_t p = f[a->col[k]]++;
            t.col[p] = i;
            t.val[p] = a->val[k];
        }
    free(f);
    return t;
}

static int cmp64(const void *a, const void *b) {
    int64_t x = *(const int64_t *)a, y = *(const int64_t *)b;
    return (x > y) - (x < y);
}

/* C = A B, Gustavson row-by-row with a dense accumulator per thread (two passes). */
static csr csr_matmul(const csr *A, const csr *B) {
    csr C = {A->nr, B->nc, calloc(A->nr + 1, sizeof(int64_t)), NULL, NULL};
#pragma omp parallel
    {
        int64_t *mark = malloc(sizeof(int64_t) * (B->nc ? B->nc : 1));
        for (int64_t j = 0; j < B->nc; ++j) mark[j] = -1;
#pragma omp for schedule(dynamic, 256)
        for (int64_t i = 0; i < A->nr; ++i) {
            int64_t cnt = 0;
            for (int64_t ka = A->ptr[i]; ka < A->ptr[i + 1]; ++ka) {
                int64_t r = A->col[ka];
                for (int64_t kb = B->ptr[r]; kb < B->ptr[r + 1]; ++kb)
                    if (mark[B->col[kb]] != i) { mark[B->col[kb]] = i; ++cnt; }
            }
            C.ptr[i + 1] = cnt;
        }
        free(mark);
    }
    for (int64_t i = 0; i < A->nr; ++i) C.ptr[i + 1] += C.ptr[i];
    int64_t nnz = C.ptr[A->nr];
    C.col = malloc(sizeof(int64_t) * (nnz ? nnz : 1));
    C.val = malloc(sizeof(double) * (nnz ? nnz : 1));
#pragma omp parallel
    {
        double *acc = calloc(B->nc ? B->nc : 1, sizeof(double));
        int64_t *mark = malloc(sizeof(int64_t) * (B->nc ? B->nc : 1));
        for (int64_t j = 0; j < B->nc; ++j) mark[j] = -1;
#pragma omp for schedule(dynamic, 256)
        for (int64_t i = 0; i < A->nr; ++i) {
            int64_t p = C.ptr[i];
            for (int64_t ka = A->ptr[i]; ka < A->ptr[i + 1]; ++ka) {
                int64_t r = A->col[ka];
                double av = A->val[ka];
                for (int64_t kb = B->ptr[r]; kb < B->ptr[r + 1]; ++kb) {
                    int64_t j = B->col[kb];
                    if (mark[j] != i) { mark[j] = i; C.col[p++] = j; acc[j] = 0.0; }
                    acc[j] += av * B->val[kb];
                }
            }
            qsort(C.col + C.ptr[i], p - C.ptr[i], sizeof(int64_t), cmp64);
            for (int64_t q = C.ptr[i]; q < p; ++q) C.val[q] = acc[C.col[q]];
        }
        free(acc);
        free(mark);
    }
    return C;
}

/* ------------------------------------------------------------------ AMG */
typedef struct olevel {
    csr A, P, R;
    double *dinv;      /* 1 / diag(A), 0 where the diagonal vanishes */
    double *x, *b, *r; /* V-cycle work */
    double omega;      /* Jacobi damping of this level's smoother, 4 / (3 rho(D^-1 A)) */
} olevel;

typedef struct or_amg {
    int nlev;
    olevel L[16];
    int64_t nc;        /* coarsest size */
    double *lu;        /* dense LU of the coarsest operator (row-major), with pivots */
    int64_t *piv;
} or_amg;

static double *diag_inverse(const csr *A) {
    double *d = calloc(A->nr, sizeof(double));
    for (int64_t i = 0; i < A->nr; ++i)
        for (int64_t k = A->ptr[i]; k < A->ptr[i + 1]; ++k)
            if (A->col[k] == i && A->val[k] != 0.0) d[i] = 1.0 / A->val[k];
    return d;
}

static double spectral_radius_dinvA(const csr *A, const double *dinv, int iters) {
    int64_t n = A->nr;
    double *x = malloc(sizeof(double) * n), *y = malloc(sizeof(double) * n);
    uint64_t st = 0x9E3779B97F4A7C15ull;
    for (int64_t i = 0; i < n; ++i) {
        st = st * 6364136223846793005ull + 1442695040888963407ull;
        x[i] = (double)(st >> 11) * 0x1.0p-53 - 0.5;
    }
    double lam = 1.0;
    for (int it = 0; it < iters; ++it) {
        double nx = 0.0;
        for (int64_t i = 0; i < n; ++i) nx += x[i] * x[i];
        nx = sqrt(nx);
        if (nx == 0.0) break;
        for (int64_t i = 0; i < n; ++i) x[i] /= nx;
        csr_spmv(A, x, y);
        double ny = 0.0;
        for (int64_t i = 0; i < n; ++i) { y[i] *= dinv[i]; ny += y[i] * y[i]; }
        lam = sqrt(ny);
        double *t = x; x = y; y = t;
    }
    free(x); free(y);
    return lam;
}

/* three-phase aggregation on a node graph (CSR ptr/idx), inactive nodes excluded */
static int64_t aggregate_nodes(int64_t nb, const int64_t *gptr, const int64_t *gidx, const char *active,
                               int64_t *agg) {
    for (int64_t i = 0; i < nb; ++i) agg[i] = -1;
    int64_t na = 0;
    for (int64_t i = 0; i < nb; ++i) {
        if (!active[i] || agg[i] >= 0) continue;
        int ok = 1, nact = 0;
        for (int64_t k = gptr[i]; k < gptr[i + 1]; ++k) {
            int64_t j = gidx[k];
            if (j == i || !active[j]) continue;
            ++nact;
            if (agg[j] >= 0) { ok = 0; break; }
        }
        if (!ok || nact == 0) continue;
        for (int64_t k = gptr[i]; k < gptr[i + 1]; ++k)
            if (active[gidx[k]]) agg[gidx[k]] = na;
        agg[i] = na++;
    }
    int64_t *p1 = malloc(sizeof(int64_t) * nb);
    memcpy(p1, agg, sizeof(int64_t) * nb);
    for (int64_t i = 0; i < nb; ++i) {
        if (!active[i] || agg[i] >= 0) continue;
        for (int64_t k = gptr[i]; k < gptr[i + 1]; ++k)
            if (gidx[k] != i && p1[gidx[k]] >= 0) { agg[i] = p1[gidx[k]]; break; }
    }
    free(p1);
    for (int64_t i = 0; i < nb; ++i) {
        if (!active[i] || agg[i] >= 0) continue;
        for (int64_t k = gptr[i]; k < gptr[i + 1]; ++k)
            if (active[gidx[k]] && agg[gidx[k]] < 0) agg[gidx[k]] = na;
        agg[i] = na++;
    }
    return na;
}

/* tentative prolongator of block size bs (rows bs*nb, cols 6*na) from near-null
 * rows B (row-major ndof x 6); coarse near-null space Bc (6*na x 6) */
static csr tentative_P(int bs, int64_t nb, int64_t na, const int64_t *agg, const double *B, double *Bc) {
    int64_t *cnt = calloc(na + 1, sizeof(int64_t));
    for (int64_t i = 0; i < nb; ++i)
        if (agg[i] >= 0) cnt[agg[i] + 1]++;
    for (int64_t a = 0; a < na; ++a) cnt[a + 1] += cnt[a];
    int64_t *mem = malloc(sizeof(int64_t) * (cnt[na] ? cnt[na] : 1)), *f = malloc(sizeof(int64_t) * (na + 1));
    memcpy(f, cnt, sizeof(int64_t) * (na + 1));
    for (int64_t i = 0; i < nb; ++i)
        if (agg[i] >= 0) mem[f[agg[i]]++] = i;
    csr P = {bs * nb, 6 * na, calloc(bs * nb + 1, sizeof(int64_t)), NULL, NULL};
    for (int64_t i = 0; i < nb; ++i)
        for (int r = 0; r < bs; ++r) P.ptr[bs * i + r + 1] = agg[i] >= 0 ? 6 : 0;
    for (int64_t i = 0; i < bs * nb; ++i) P.ptr[i + 1] += P.ptr[i];
    P.col = malloc(sizeof(int64_t) * (P.ptr[bs * nb] ? P.ptr[bs * nb] : 1));
    P.val = malloc(sizeof(double) * (P.ptr[bs * nb] ? P.ptr[bs * nb] : 1));
    memset(Bc, 0, sizeof(double) * 36 * na);
    for (int64_t a = 0; a < na; ++a) {
        int64_t m = cnt[a + 1] - cnt[a], rows = bs * m;
        double *Q = calloc(rows * 6, sizeof(double)); /* column-major */
        for (int64_t t = 0; t < m; ++t)
            for (int r = 0; r < bs; ++r)
                for (int c = 0; c < 6; ++c) Q[c * rows + t * bs + r] = B[(bs * mem[cnt[a] + t] + r) * 6 + c];
        for (int c = 0; c < 6; ++c) { /* modified Gram-Schmidt, two passes */
            double *v = Q + c * rows, n0 = 0.0, nn = 0.0;
            for (int64_t r = 0; r < rows; ++r) n0 += v[r] * v[r];
            n0 = sqrt(n0);
            for (int pass = 0; pass < 2; ++pass)
                for (int j = 0; j < c; ++j) {
                    double d = 0.0, *q = Q + j * rows;
                    for (int64_t r = 0; r < rows; ++r) d += q[r] * v[r];
                    Bc[36 * a + 6 * j + c] += d;
                    for (int64_t r = 0; r < rows; ++r) v[r] -= d * q[r];
                }
            for (int64_t r = 0; r < rows; ++r) nn += v[r] * v[r];
            nn = sqrt(nn);
            if (n0 == 0.0 || nn <= 1e-10 * n0) {
                for (int64_t r = 0; r < rows; ++r) v[r] = 0.0;
                continue;
            }
            Bc[36 * a + 6 * c + c] = nn;
            for (int64_t r = 0; r < rows; ++r) v[r] /= nn;
        }
        for (int64_t t = 0; t < m; ++t)
            for (int r = 0; r < bs; ++r) {
                int64_t row = bs * mem[cnt[a] + t] + r, p = P.ptr[row];
                for (int c = 0; c < 6; ++c) { P.col[p + c] = 6 * a + c; P.val[p + c] = Q[c * rows + t * bs + r]; }
            }
        free(Q);
    }
    free(cnt); free(mem); free(f);
    return P;
}

/* dense LU with partial pivoting (row-major) */
static void dense_lu(int64_t n, double *a, int64_t *piv) {
    for (int64_t k = 0; k < n; ++k) {
        int64_t p = k;
        for (int64_t i = k + 1; i < n; ++i)
            if (fabs(a[i * n + k]) > fabs(a[p * n + k])) p = i;
        piv[k] = p;
        if (p != k)
            for (int64_t j = 0; j < n; ++j) { double t = a[k * n + j]; a[k * n + j] = a[p * n + j]; a[p * n + j] = t; }
        double d = a[k * n + k];
        if (d == 0.0) { a[k * n + k] = d = 1.0; }
        for (int64_t i = k + 1; i < n; ++i) {
            double f = (a[i * n + k] /= d);
            if (f != 0.0)
                for (int64_t j = k + 1; j < n; ++j) a[i * n + j] -= f * a[k * n + j];
        }
    }
}
static void dense_lu_solve(int64_t n, const double *a, const int64_t *piv, double *x) {
    for (int64_t k = 0; k < n; ++k) {
        int64_t p = piv[k];
        if (p != k) { double t = x[k]; x[k] = x[p]; x[p] = t; }
    }
    for (int64_t i = 0; i < n; ++i)
        for (int64_t j = 0; j < i; ++j) x[i] -= a[i * n + j] * x[j];
    for (int64_t i = n - 1; i >= 0; --i) {
        for (int64_t j = i + 1; j < n; ++j) x[i] -= a[i * n + j] * x[j];
        x[i] /= a[i * n + i];
    }
}

/* Build the hierarchy from a CSC operator (K^T is taken as K: the pattern is
 * symmetric; the Galerkin products use the CSR transpose of the given CSC).
 * isbc[d] != 0 marks constrained dofs (identity rows); nodes: 3 x nnodes. */
or_amg *or_amg_setup(int64_t n, const int64_t *colptr, const int64_t *rowval, const double *nzval,
                     const double *nodes, const uint8_t *isbc, int64_t coarse_max, int nthreads) {
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
    or_amg *amg = calloc(1, sizeof(or_amg));
    csr Kc = {n, n, (int64_t *)colptr, (int64_t *)rowval, (double *)nzval};
    csr A = csr_transpose(&Kc); /* CSR of K */
    int64_t nb = n / 3;
    int bs = 3;
    /* near-null space: rigid-body modes, zero on constrained dofs */
    double *B = calloc(n * 6, sizeof(double));
    char *active = malloc(nb);
    for (int64_t i = 0; i < nb; ++i) {
        double x = nodes[3 * i], y = nodes[3 * i + 1], z = nodes[3 * i + 2];
        double rb[3][6] = {{1, 0, 0, 0, z, -y}, {0, 1, 0, -z, 0, x}, {0, 0, 1, y, -x, 0}};
        int nfix = 0;
        for (int r = 0; r < 3; ++r) {
            int fx = isbc[3 * i + r] != 0;
            nfix += fx;
            for (int c = 0; c < 6; ++c) B[(3 * i + r) * 6 + c] = fx ? 0.0 : rb[r][c];
        }
        active[i] = nfix < 3;
    }
    int lev = 0;
    for (;;) {
        olevel *L = &amg->L[lev];
        L->A = A;
        L->dinv = diag_inverse(&A);
        L->x = calloc(A.nr, sizeof(double));
        L->b = calloc(A.nr, sizeof(double));
        L->r = calloc(A.nr, sizeof(double));
        if (A.nr <= coarse_max || lev == 15) break;
        /* node graph of this level: block pattern */
        int64_t *gptr = calloc(nb + 1, sizeof(int64_t));
        for (int64_t i = 0; i < nb; ++i) {
            int64_t last = -1, c = 0;
            for (int64_t k = A.ptr[bs * i]; k < A.ptr[bs * i + 1]; ++k) {
                int64_t j = A.col[k] / bs;
                if (j != last) { ++c; last = j; }
            }
            gptr[i + 1] = gptr[i] + c;
        }
        int64_t *gidx = malloc(sizeof(int64_t) * (gptr[nb] ? gptr[nb] : 1));
        for (int64_t i = 0; i < nb; ++i) {
            int64_t last = -1, p = gptr[i];
            for (int64_t k = A.ptr[bs * i]; k < A.ptr[bs * i + 1]; ++k) {
                int64_t j = A.col[k] / bs;
                if (j != last) { gidx[p++] = j; last = j; }
            }
        }
        int64_t *agg = malloc(sizeof(int64_t) * nb);
        int64_t na = aggregate_nodes(nb, gptr, gidx, active, agg);
        free(gptr); free(gidx);
        if (na == 0 || 6 * na >= A.nr) { free(agg); break; }
        double *Bc = malloc(sizeof(double) * 36 * na);
        csr Pt = tentative_P(bs, nb, na, agg, B, Bc);
        free(agg);
        double rho = spectral_radius_dinvA(&A, L->dinv, 10);
        double w = (4.0 / 3.0) / rho;
        /* the smoother takes the same damping: a fixed 2/3 exceeds 2 / rho(D^-1 A) for the
           elastic tangent (rho ~ 3-4 with a scalar diagonal), which makes the sweep divergent
           and the V-cycle indefinite (PCG broke down at its second iteration) */
        L->omega = w;
        csr AP = csr_matmul(&A, &Pt);
        /* P = Pt - w D^-1 A Pt: merge rows of Pt into the (superset) pattern of AP */
        for (int64_t i = 0; i < AP.nr; ++i) {
            for (int64_t k = AP.ptr[i]; k < AP.ptr[i + 1]; ++k) AP.val[k] *= -w * L->dinv[i];
            for (int64_t kt = Pt.ptr[i]; kt < Pt.ptr[i + 1]; ++kt) {
                for (int64_t k = AP.ptr[i]; k < AP.ptr[i + 1]; ++k)
                    if (AP.col[k] == Pt.col[kt]) { AP.val[k] += Pt.val[kt]; break; }
            }
        }
        csr_free(&Pt);
        L->P = AP;
        L->R = csr_transpose(&L->P);
        csr T = csr_matmul(&A, &L->P);
        csr Ac = csr_matmul(&L->R, &T);
        csr_free(&T);
        free(B);
        B = Bc;
        A = Ac;
        nb = na;
        bs = 6;
        active = realloc(active, nb);
        memset(active, 1, nb);
        ++lev;
    }
    amg->nlev = lev + 1;
    free(B);
    free(active);
    olevel *C = &amg->L[lev];
    int64_t nc = C->A.nr;
    amg->nc = nc;
    amg->lu = calloc(nc * nc, sizeof(double));
    amg->piv = malloc(sizeof(int64_t) * (nc ? nc : 1));
    for (int64_t i = 0; i < nc; ++i) {
        int any = 0;
        for (int64_t k = C->A.ptr[i]; k < C->A.ptr[i + 1]; ++k) {
            amg->lu[i * nc + C->A.col[k]] = C->A.val[k];
            any |= C->A.val[k] != 0.0;
        }
        if (!any) amg->lu[i * nc + i] = 1.0;
    }
    dense_lu(nc, amg->lu, amg->piv);
    return amg;
}

void or_amg_free(or_amg *amg) {
    if (!amg) return;
    for (int l = 0; l < amg->nlev; ++l) {
        olevel *L = &amg->L[l];
        csr_free(&L->A);
        csr_free(&L->P);
        csr_free(&L->R);
        free(L->dinv); free(L->x); free(L->b); free(L->r);
    }
    free(amg->lu);
    free(amg->piv);
    free(amg);
}

int or_amg_levels(const or_amg *amg, int64_t *sizes) {
    for (int l = 0; l < amg->nlev; ++l) sizes[l] = amg->L[l].A.nr;
    return amg->nlev;
}

static void vcycle(const or_amg *amg, int l, const double *b, double *x) {
    const olevel *L = &amg->L[l];
    int64_t n = L->A.nr;
    if (l == amg->nlev - 1) {
        memcpy(x, b, sizeof(double) * n);
        dense_lu_solve(n, amg->lu, amg->piv, x);
        return;
    }
    const double w = L->omega;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) x[i] = w * L->dinv[i] * b[i]; /* pre-sweep from x = 0 */
    csr_spmv(&L->A, x, L->r);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) L->r[i] = b[i] - L->r[i];
    const olevel *N = &amg->L[l + 1];
    csr_spmv(&L->R, L->r, N->b);
    vcycle(amg, l + 1, N->b, N->x);
    csr_spmv(&L->P, N->x, L->r);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) x[i] += L->r[i];
    csr_spmv(&L->A, x, L->r); /* post-sweep */
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) x[i] += w * L->dinv[i] * (b[i] - L->r[i]);
}

void or_amg_vcycle(const or_amg *amg, const double *b, double *x) { vcycle(amg, 0, b, x); }

static double dot(int64_t n, const double *a, const double *b) {
    double s = 0.0;
#pragma omp parallel for reduction(+ : s) schedule(static)
    for (int64_t i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
}

/* PCG with the V-cycle preconditioner on the CSR operator of the hierarchy's level 0.
 * Returns iterations (negative on breakdown, -itmax-1 if not converged). */
int or_amg_pcg(const or_amg *amg, const double *b, double *x, double rtol, int itmax) {
    const csr *A = &amg->L[0].A;
    int64_t n = A->nr;
    double *r = malloc(sizeof(double) * n), *z = malloc(sizeof(double) * n), *p = malloc(sizeof(double) * n),
           *q = malloc(sizeof(double) * n);
    memset(x, 0, sizeof(double) * n);
    memcpy(r, b, sizeof(double) * n);
    vcycle(amg, 0, r, z);
    double rz = dot(n, r, z);
    int ret = -itmax - 1;
    double eps = rtol * sqrt(fabs(rz));
    if (rz <= 0.0) { ret = rz == 0.0 ? 0 : -1; goto done; }
    memcpy(p, z, sizeof(double) * n);
    for (int it = 1; it <= itmax; ++it) {
        csr_spmv(A, p, q);
        double pq = dot(n, p, q);
        if (!(pq > 0)) { ret = -it; goto done; }
        double a = rz / pq;
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; ++i) { x[i] += a * p[i]; r[i] -= a * q[i]; }
        vcycle(amg, 0, r, z);
        double rzn = dot(n, r, z);
        if (!(rzn >= 0)) { ret = -it; goto done; }
        if (sqrt(rzn) <= eps) { ret = it; goto done; }
        double be = rzn / rz;
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; ++i) p[i] = z[i] + be * p[i];
        rz = rzn;
    }
done:
    free(r); free(z); free(p); free(q);
    return ret;
}

/* y = A x with the hierarchy's level-0 CSR operator (the operator PCG multiplies by). */
void or_amg_apply_A(const or_amg *amg, const double *x, double *y) { csr_spmv(&amg->L[0].A, x, y); }
