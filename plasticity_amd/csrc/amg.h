// amg.h -- smoothed-aggregation AMG with rigid-body near-null space (internal).
//
// Replaces AlgebraicMultigrid.jl's smoothed_aggregation(K; B) + aspreconditioner
// (Solver.jl:136-154) with a node-block variant: aggregates of mesh nodes (not
// scalar DOFs), tentative prolongator from a per-aggregate QR of the six rigid-body
// modes (Solver.jl:117-129), Jacobi-smoothed prolongator P = (I - w D^-1 A) Pt with
// block-diagonal D and w = 4/(3 rho(D^-1 A)), Galerkin coarse operators R A P with
// R = P^T.  The symbolic part (aggregates, QR, sparsity of P, A P and R A P) depends
// only on the mesh graph, coordinates and constraints and is built once on the host;
// the numeric part runs on the device at every rebuild (once per load step).
#pragma once
#include "common.h"

struct pfem_ctx;

namespace pfem {

// most blocks per chunk of the streamed prolongation (bsr_spmv_stream)
constexpr int P_STREAM_CAP = 256;

struct AmgLevel {
    int bs = 3;                  // dofs per block row on this level (3 fine, 6 coarse)
    int64_t nb = 0;              // block rows
    int64_t n = 0;               // dofs = bs * nb
    // operator (level 0 uses the context's SELL K)
    int64_t nnzb = 0;
    dbuf<int32_t> rowptr, colind, diagpos;
    dbuf<double> A;
    dbuf<float> A32;
    std::vector<int32_t> h_rowptr, h_colind;
    // block-diagonal inverse and spectral radius of D^-1 A
    dbuf<double> dinv;
    dbuf<float> dinv32;
    double rho = 2.0;
    // prolongator to this level from the next: P (nb x nagg, bs x 6 blocks)
    int64_t nagg = 0;
    int64_t nP = 0, nQ = 0, nR = 0;  // true block counts (buffers hold at least one entry)
    dbuf<int32_t> agg;           // node -> aggregate (-1 none)
    dbuf<double> Pt;             // tentative block per node (bs x 6), zero if agg < 0
    dbuf<int32_t> P_rowptr, P_colind, P_row;
    dbuf<int32_t> P_chunk;       // row boundaries of the streamed prolongation's chunks
    int64_t nPchunk = 0;         // 0: thread-per-row prolongation
    dbuf<double> P;
    dbuf<float> P32;
    // restriction R = P^T (nagg x nb, 6 x bs blocks)
    dbuf<int32_t> R_rowptr, R_colind, R_src;
    dbuf<double> R;
    dbuf<float> R32;
    // Q = A P (nb x nagg), and the row of every Q block / coarse block
    dbuf<int32_t> Q_rowptr, Q_colind, Q_row;
    dbuf<double> Q;
    dbuf<int32_t> C_row;         // row of every block of the next level's operator
    // V-cycle vectors
    dbuf<double> x, b, r, d, t;
    dbuf<double> pw;             // power-iteration vector, kept between set-ups (warm start)
    bool pw_warm = false;
    dbuf<float> x32, b32, r32, d32, t32;
    dbuf<double> wr, we;         // W-cycle: coarse residual and correction of the second visit
    dbuf<float> wr32, we32;
};

// Numeric data of the levels below the finest (operators, block-diagonal inverses, rho)
// and the coarsest inverse, kept aside after the set-up of one Newton iteration index
// (rebuild policy 7) for the same iteration of later load steps.  Valid while the
// prolongators it was built with are kept.
struct AmgStash {
    struct Level {
        dbuf<double> A, dinv;
        dbuf<float> A32, dinv32;
        double rho = 2.0;
    };
    std::vector<Level> lv;       // levels 1 .. nlev-1
    dbuf<double> coarse_inv;
    bool ok = false, have32 = false;
    int iters = 0;               // Krylov iterations of the solve right after it was built
    bool stale = true;           // a later solve on it needed clearly more: rebuild
};

struct Amg {
    std::vector<AmgLevel> L;
    // V-cycle replayed as a HIP graph ([0] FP64, [1] FP32), captured at the first
    // application after a (re)build for one (input, output) pair
    hipGraphExec_t gexec[2] = {nullptr, nullptr};
    const void *g_in[2] = {nullptr, nullptr}, *g_out[2] = {nullptr, nullptr};
    // coarsest level: dense inverse (FP64)
    int64_t nc = 0;
    dbuf<double> coarse_inv;
    dbuf<double> gj_work, gj_prow, gj_fcol, gj_dw;  // Gauss-Jordan scratch
    dbuf<int> gj_iw;
    // concurrent set-up: the rho power iterations run on their own stream (with their own
    // reduction scratch) beside the Galerkin products when the prolongators are kept
    hipStream_t s_rho = nullptr;
    std::vector<hipEvent_t> ev_level;  // level l's operator and block-diagonal inverse ready
    hipEvent_t ev_rho = nullptr;
    dbuf<double> rho_partial, rho_scal;
    int smoother = PFEM_SMOOTH_CHEBYSHEV;
    int degree = 2;         // fine level
    int coarse_degree = 2;  // levels below the finest
    int degree_at(int l) const { return l == 0 ? degree : coarse_degree; }
    bool wcycle = false;
    bool have32 = false;
    double op_complexity = 1.0;
    std::vector<AmgStash> stash;  // policy 7: one per Newton iteration index (slot)
};

Amg *amg_create(pfem_ctx *c);                 // symbolic set-up (host)
void amg_destroy(Amg *a);
// numeric set-up (device); keep_P: the prolongator and restriction of the last set-up
// are kept, everything that depends on the operator (smoothers, A P, coarse operators,
// coarse inverse) is recomputed from the current one
void amg_numeric_setup(pfem_ctx *c, Amg *a, bool keep_P = false);
void amg_refresh_fine(pfem_ctx *c, Amg *a);  // level-0 smoother from the current K only
// Copy the numeric data of the levels below the finest (operators, block-diagonal
// inverses, rho) and the coarsest inverse into / out of stash slot `slot`.
// amg_restore_coarse then refreshes the level-0 smoother from the current K; it returns
// false (nothing changed) when the slot holds no copy made with the current prolongators
// and precision.
AmgStash &amg_stash_slot(Amg *a, int slot);
void amg_stash_coarse(pfem_ctx *c, Amg *a, int slot);
bool amg_restore_coarse(pfem_ctx *c, Amg *a, int slot);
void amg_apply64(pfem_ctx *c, Amg *a, const double *r, double *z);  // z = V-cycle(r)
void amg_apply32(pfem_ctx *c, Amg *a, const float *r, float *z);
// Times `reps` cycles (prec 0 FP64, 1 FP32) and `reps` fused fine-level Chebyshev steps
// on the current hierarchy; returns seconds per cycle / step and their algorithmic bytes.
void amg_bench(pfem_ctx *c, Amg *a, int prec, int reps, double *sec_cycle, double *bytes_cycle, double *sec_step,
               double *bytes_step);

// amg.hip launchers
template <int BS>
void amg_block_inverse(int64_t nb, const int32_t *diagpos, const double *A, double *dinv, hipStream_t s);
template <int BS, bool SELL>
void amg_smooth_P(int64_t nP, const int32_t *P_row, const int32_t *P_colind, const int64_t *slice_ptr,
                  const int32_t *rowlen, const int32_t *rowptr, const int32_t *colind, const double *A,
                  const int32_t *agg, const double *Pt, const double *dinv, double omega, double *P, hipStream_t s);
template <int BS, bool SELL>
void amg_AP(int64_t nQ, const int32_t *Q_row, const int32_t *Q_colind, const int64_t *slice_ptr,
            const int32_t *rowlen, const int32_t *rowptr, const int32_t *colind, const double *A,
            const int32_t *P_rowptr, const int32_t *P_colind, const double *P, double *Q, hipStream_t s);
template <int BS>
void amg_RAP(int64_t nC, const int32_t *C_row, const int32_t *C_colind, const int32_t *R_rowptr,
             const int32_t *R_colind, const double *R, const int32_t *Q_rowptr, const int32_t *Q_colind,
             const double *Q, double *C, hipStream_t s);
template <int BS>
void amg_transpose_P(int64_t nR, const int32_t *R_src, const double *P, double *R, hipStream_t s);
template <int BR, int BC, class VT, class XT>
void bsr_spmv(int64_t nb, const int32_t *rowptr, const int32_t *colind, const VT *vals, const XT *x, XT *y,
              double beta, hipStream_t s);  // y = A x + beta*y
// y = A x + beta*y over chunks of consecutive block rows (chunk_row: nchunk + 1 row
// boundaries, at most P_STREAM_CAP blocks per chunk)
template <int BR, int BC, class VT, class XT>
void bsr_spmv_stream(int64_t nchunk, const int32_t *chunk_row, const int32_t *rowptr, const int32_t *colind,
                     const VT *vals, const XT *x, XT *y, double beta, hipStream_t s);
template <int BR, int BC, class VT, class XT>
void bsr_spmv_wave(int64_t nb, const int32_t *rowptr, const int32_t *colind, const VT *vals, const XT *x, XT *y,
                   hipStream_t s);  // y = A x, one wave per block row (long rows)
template <int BS, class VT, class XT>
void bsr_residual(int64_t nb, const int32_t *rowptr, const int32_t *colind, const VT *vals, const XT *x,
                  const XT *b, XT *r, hipStream_t s);
template <int BS, class VT, class XT>
void sell_residual(const FineMatrix &K, const VT *vals, const XT *x, const XT *b, XT *r, hipStream_t s);
template <int BS, class VT, class XT>
void sell_spmv(const FineMatrix &K, const VT *vals, const XT *x, XT *y, hipStream_t s);
// fine-level Chebyshev step with its product: t = K d; r -= t; d' = c1 d + c2 Dinv r;
// x += d' (d' into d_out; d_out == null: the last step, r and d' not stored)
template <class VT, class XT>
void sell_cheb(const FineMatrix &K, const VT *vals, const XT *d, const VT *dinv, XT *r, XT *d_out, XT *x, double c1,
               double c2, hipStream_t s);
template <int BS, class DT, class XT>
void block_diag_apply(int64_t nb, const DT *dinv, const XT *r, XT *z, double scale, hipStream_t s);
// d = scale Dinv v; zero_init: r = v, x = d; else x += d
template <int BS, class DT, class XT>
void diag_first(int64_t nb, const DT *dinv, const XT *v, XT *r, XT *d, XT *x, double scale, bool zero_init,
                hipStream_t s);
template <int BS, class DT, class XT>
void cheb_step(int64_t nb, const DT *dinv, const XT *Ad, XT *r, XT *d, XT *x, double c1, double c2,
               hipStream_t s);  // r -= Ad; d = c1 d + c2 Dinv r; x += d
// coarse-level Chebyshev step with its product: t = A d; r -= t; d' = c1 d + c2 Dinv r;
// x += d' (d' into d_out; d_out == null: the last step, r and d' not stored)
template <int BS, class DT, class XT>
void bsr_cheb(int64_t nb, const int32_t *rowptr, const int32_t *colind, const DT *vals, const XT *d, const DT *dinv,
              XT *r, XT *d_out, XT *x, double c1, double c2, hipStream_t s);
template <class XT>
void dense_gemv(int64_t n, const double *Ainv, const XT *b, XT *x, hipStream_t s);
void scale_by_rnorm(int64_t n, const double *nrm2, double *x, hipStream_t s);
// dense inverse of a block-sparse operator (SELL fine level or CSR of blocks) by
// Gauss-Jordan with partial pivoting; W is 2 n^2, prow 2n, fcol n scratch
void dense_coarse_inverse(int64_t nb, int bs, bool sell, const int64_t *slice_ptr, const int32_t *rowlen,
                          const int32_t *rowptr, const int32_t *colind, const double *A, double *W, double *prow,
                          double *fcol, int *iw, double *dw, double *inv, hipStream_t s);
// inv = M^-1 for the dense n x n M stored in the left half of W (n x 2n, row-major);
// zero rows of M are decoupled with a unit diagonal.  W, prow (2n), fcol (n), iw
// (gj_scratch_ints(n)) and dw (gj_scratch_doubles(n)) are scratch.
size_t gj_scratch_ints(int64_t n);
size_t gj_scratch_doubles(int64_t n);
void gauss_jordan_inverse(int64_t n, double *W, double *prow, double *fcol, int *iw, double *dw, double *inv,
                          hipStream_t s);

}  // namespace pfem
