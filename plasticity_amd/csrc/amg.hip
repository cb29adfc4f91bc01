// amg.hip -- device kernels of the smoothed-aggregation AMG (numeric set-up and
// V-cycle operators).  Every kernel is deterministic: each output block is owned by
// exactly one thread, so repeated set-ups give bit-identical hierarchies.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "amg.h"
#include "sell.h"

namespace pfem {
namespace {

inline unsigned grid_for(int64_t n, int bs) {
    int64_t g = (n + bs - 1) / bs;
    return (unsigned)(g < 1 ? 1 : g);
}

// A-row accessor: SELL-64 (fine level) or CSR of blocks (coarse levels).
template <bool SELL>
struct RowIt {
    int64_t start, stride;
    int len;
    __device__ RowIt(int64_t i, const int64_t *slice_ptr, const int32_t *rowlen, const int32_t *rowptr) {
        if (SELL) {
            start = slice_ptr[i >> 6] + (i & 63);
            stride = SELL_C;
            len = rowlen[i];
        } else {
            start = rowptr[i];
            stride = 1;
            len = rowptr[i + 1] - rowptr[i];
        }
    }
    __device__ int64_t pos(int k) const { return start + (int64_t)k * stride; }
};

__device__ __forceinline__ int find_sorted(const int32_t *a, int lo, int hi, int32_t key) {
    // index in [lo, hi) with a[idx] == key, or -1
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        const int32_t v = a[mid];
        if (v == key) return mid;
        if (v < key) lo = mid + 1; else hi = mid;
    }
    return -1;
}

// G independent searches of sorted rows advanced in lockstep, so that their dependent
// loads overlap: m[u] = index in [lo[u], hi[u]) with a[index] == key, or -1.  The
// Galerkin kernels take SEARCH_G terms of a row at a time (AP at 100^3: 4.74 -> 3.84 ms;
// 8 at a time cost occupancy: 4.18 ms).
constexpr int SEARCH_G = 4;
template <int G>
__device__ __forceinline__ void find_sorted_lockstep(const int32_t *__restrict__ a, int (&lo)[G], int (&hi)[G],
                                                     int32_t key, int (&m)[G]) {
#pragma unroll
    for (int u = 0; u < G; ++u) m[u] = -1;
    bool more = true;
    while (more) {
        int mid[G];
        int32_t v[G];
#pragma unroll
        for (int u = 0; u < G; ++u) {
            mid[u] = (lo[u] + hi[u]) >> 1;
            v[u] = lo[u] < hi[u] ? a[mid[u]] : 0;
        }
        more = false;
#pragma unroll
        for (int u = 0; u < G; ++u) {
            if (lo[u] < hi[u]) {
                if (v[u] == key) {
                    m[u] = mid[u];
                    hi[u] = lo[u];
                } else if (v[u] < key) {
                    lo[u] = mid[u] + 1;
                } else {
                    hi[u] = mid[u];
                }
            }
            more = more || lo[u] < hi[u];
        }
    }
}

// Gauss-Jordan inverse with partial pivoting of a BS x BS diagonal block.  A zero
// row (a coarse dof whose rigid-body column was dropped by the QR) is treated as
// decoupled: its diagonal is taken as 1.
template <int BS>
__global__ void block_inverse_kernel(int64_t nb, const int32_t *__restrict__ diagpos, const double *__restrict__ A,
                                     double *__restrict__ dinv) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nb) return;
    double a[BS][BS], v[BS][BS];
    const double *m = A + (int64_t)BS * BS * diagpos[i];
    for (int r = 0; r < BS; ++r) {
        bool zero = true;
        for (int c = 0; c < BS; ++c) {
            a[r][c] = m[BS * r + c];
            v[r][c] = (r == c) ? 1.0 : 0.0;
            zero = zero && a[r][c] == 0.0;
        }
        if (zero) a[r][r] = 1.0;
    }
    for (int col = 0; col < BS; ++col) {
        int piv = col;
        double best = fabs(a[col][col]);
        for (int r = col + 1; r < BS; ++r)
            if (fabs(a[r][col]) > best) { best = fabs(a[r][col]); piv = r; }
        if (piv != col)
            for (int c = 0; c < BS; ++c) {
                double t = a[col][c]; a[col][c] = a[piv][c]; a[piv][c] = t;
                t = v[col][c]; v[col][c] = v[piv][c]; v[piv][c] = t;
            }
        const double d = a[col][col];
        const double id = d != 0.0 ? 1.0 / d : 0.0;
        for (int c = 0; c < BS; ++c) { a[col][c] *= id; v[col][c] *= id; }
        for (int r = 0; r < BS; ++r) {
            if (r == col) continue;
            const double f = a[r][col];
            if (f == 0.0) continue;
            for (int c = 0; c < BS; ++c) { a[r][c] -= f * a[col][c]; v[r][c] -= f * v[col][c]; }
        }
    }
    double *o = dinv + (int64_t)BS * BS * i;
    for (int r = 0; r < BS; ++r)
        for (int c = 0; c < BS; ++c) o[BS * r + c] = v[r][c];
}

// P block (i, J) = Pt_i [if J = agg(i)] - omega * Dinv_i * sum_{k: agg(col k) = J} A_ik Pt_{col k};
// one thread per P block, accumulated in registers.
template <int BS, bool SELL>
__global__ __launch_bounds__(128) void smooth_P_kernel(int64_t nP, const int32_t *__restrict__ P_row, const int32_t *__restrict__ P_colind,
                                const int64_t *__restrict__ slice_ptr, const int32_t *__restrict__ rowlen,
                                const int32_t *__restrict__ rowptr, const int32_t *__restrict__ colind,
                                const double *__restrict__ A, const int32_t *__restrict__ agg,
                                const double *__restrict__ Pt, const double *__restrict__ dinv, double omega,
                                double *__restrict__ P) {
    const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= nP) return;
    const int64_t i = P_row[m];
    const int32_t J = P_colind[m];
    constexpr int B6 = BS * 6;
    double acc[B6];
#pragma unroll
    for (int t = 0; t < B6; ++t) acc[t] = 0.0;
    RowIt<SELL> row(i, slice_ptr, rowlen, rowptr);
    for (int k = 0; k < row.len; ++k) {
        const int64_t pos = row.pos(k);
        const int32_t j = colind[pos];
        if (agg[j] != J) continue;
        const double *a = A + (int64_t)BS * BS * pos;
        const double *pt = Pt + (int64_t)B6 * j;
#pragma unroll
        for (int r = 0; r < BS; ++r)
#pragma unroll
            for (int c = 0; c < 6; ++c) {
                double s = 0.0;
#pragma unroll
                for (int q = 0; q < BS; ++q) s += a[BS * r + q] * pt[6 * q + c];
                acc[6 * r + c] += s;
            }
    }
    const double *di = dinv + (int64_t)BS * BS * i;
    const bool own = (agg[i] == J);
    double *o = P + (int64_t)B6 * m;
#pragma unroll
    for (int r = 0; r < BS; ++r)
#pragma unroll
        for (int c = 0; c < 6; ++c) {
            double s = 0.0;
#pragma unroll
            for (int q = 0; q < BS; ++q) s += di[BS * r + q] * acc[6 * q + c];
            o[6 * r + c] = (own ? Pt[(int64_t)B6 * i + 6 * r + c] : 0.0) - omega * s;
        }
}

// Q block (i, J) = sum_k A_ik P_{col k, J}
template <int BS, bool SELL>
__global__ __launch_bounds__(128) void AP_kernel(int64_t nQ, const int32_t *__restrict__ Q_row, const int32_t *__restrict__ Q_colind,
                          const int64_t *__restrict__ slice_ptr, const int32_t *__restrict__ rowlen,
                          const int32_t *__restrict__ rowptr, const int32_t *__restrict__ colind,
                          const double *__restrict__ A, const int32_t *__restrict__ P_rowptr,
                          const int32_t *__restrict__ P_colind, const double *__restrict__ P,
                          double *__restrict__ Q) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nQ) return;
    const int64_t i = Q_row[q];
    const int32_t J = Q_colind[q];
    constexpr int B6 = BS * 6;
    double acc[B6];
    for (int t = 0; t < B6; ++t) acc[t] = 0.0;
    RowIt<SELL> row(i, slice_ptr, rowlen, rowptr);
    // G terms of the row at a time: their searches run in lockstep, then the hits are
    // added in row order
    for (int k0 = 0; k0 < row.len; k0 += SEARCH_G) {
        int lo[SEARCH_G], hi[SEARCH_G], m[SEARCH_G];
        int32_t j[SEARCH_G];
#pragma unroll
        for (int u = 0; u < SEARCH_G; ++u) j[u] = k0 + u < row.len ? colind[row.pos(k0 + u)] : -1;
#pragma unroll
        for (int u = 0; u < SEARCH_G; ++u) {
            lo[u] = j[u] >= 0 ? P_rowptr[j[u]] : 0;
            hi[u] = j[u] >= 0 ? P_rowptr[j[u] + 1] : 0;
        }
        find_sorted_lockstep(P_colind, lo, hi, J, m);
#pragma unroll
        for (int u = 0; u < SEARCH_G; ++u) {
            if (m[u] < 0) continue;
            const double *a = A + (int64_t)BS * BS * row.pos(k0 + u);
            const double *p = P + (int64_t)B6 * m[u];
            for (int r = 0; r < BS; ++r)
                for (int c = 0; c < 6; ++c) {
                    double s = 0.0;
                    for (int t = 0; t < BS; ++t) s += a[BS * r + t] * p[6 * t + c];
                    acc[6 * r + c] += s;
                }
        }
    }
    double *o = Q + (int64_t)B6 * q;
    for (int t = 0; t < B6; ++t) o[t] = acc[t];
}

// Coarse block (I, J) = sum_{m in R row I} R_m Q_{R_col m, J}
template <int BS>
__global__ __launch_bounds__(128) void RAP_kernel(int64_t nC, const int32_t *__restrict__ C_row, const int32_t *__restrict__ C_colind,
                           const int32_t *__restrict__ R_rowptr, const int32_t *__restrict__ R_colind,
                           const double *__restrict__ R, const int32_t *__restrict__ Q_rowptr,
                           const int32_t *__restrict__ Q_colind, const double *__restrict__ Q,
                           double *__restrict__ C) {
    const int64_t cb = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (cb >= nC) return;
    const int32_t I = C_row[cb], J = C_colind[cb];
    double acc[36];
    for (int t = 0; t < 36; ++t) acc[t] = 0.0;
    const int m1 = R_rowptr[I + 1];
    for (int m0 = R_rowptr[I]; m0 < m1; m0 += SEARCH_G) {  // lockstep searches, hits in order
        int lo[SEARCH_G], hi[SEARCH_G], q[SEARCH_G];
        int32_t i[SEARCH_G];
#pragma unroll
        for (int u = 0; u < SEARCH_G; ++u) i[u] = m0 + u < m1 ? R_colind[m0 + u] : -1;
#pragma unroll
        for (int u = 0; u < SEARCH_G; ++u) {
            lo[u] = i[u] >= 0 ? Q_rowptr[i[u]] : 0;
            hi[u] = i[u] >= 0 ? Q_rowptr[i[u] + 1] : 0;
        }
        find_sorted_lockstep(Q_colind, lo, hi, J, q);
#pragma unroll
        for (int u = 0; u < SEARCH_G; ++u) {
            if (q[u] < 0) continue;
            const double *rb = R + (int64_t)6 * BS * (m0 + u);
            const double *qb = Q + (int64_t)BS * 6 * q[u];
            for (int r = 0; r < 6; ++r)
                for (int c = 0; c < 6; ++c) {
                    double s = 0.0;
                    for (int t = 0; t < BS; ++t) s += rb[BS * r + t] * qb[6 * t + c];
                    acc[6 * r + c] += s;
                }
        }
    }
    double *o = C + 36 * cb;
    for (int t = 0; t < 36; ++t) o[t] = acc[t];
}

template <int BS>
__global__ void transpose_P_kernel(int64_t nR, const int32_t *__restrict__ R_src, const double *__restrict__ P,
                                   double *__restrict__ R) {
    const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= nR) return;
    const double *p = P + (int64_t)BS * 6 * R_src[m];
    double *o = R + (int64_t)6 * BS * m;
    for (int r = 0; r < 6; ++r)
        for (int c = 0; c < BS; ++c) o[BS * r + c] = p[6 * c + r];
}

// Row r of the blocks k0..k1-1 of a block row times x at their columns, summed in
// storage order; ROW_U blocks per trip, all loaded before any is added, so each thread
// keeps ROW_U blocks' loads in flight.
constexpr int ROW_U = 4;
template <int BR, int BC, class VT, class XT>
__device__ __forceinline__ VT block_row_dot(int k0, int k1, int r, const int32_t *__restrict__ colind,
                                            const VT *__restrict__ vals, const XT *__restrict__ x) {
    VT s = 0;
    int k = k0;
    for (; k + ROW_U <= k1; k += ROW_U) {
        VT a[ROW_U][BC], p[ROW_U][BC];
#pragma unroll
        for (int u = 0; u < ROW_U; ++u) {
            const VT *v = vals + (int64_t)BR * BC * (k + u) + BC * r;
            const XT *xx = x + (int64_t)BC * colind[k + u];
#pragma unroll
            for (int c = 0; c < BC; ++c) {
                a[u][c] = v[c];
                p[u][c] = (VT)xx[c];
            }
        }
#pragma unroll
        for (int u = 0; u < ROW_U; ++u)
#pragma unroll
            for (int c = 0; c < BC; ++c) s += a[u][c] * p[u][c];
    }
    for (; k < k1; ++k) {
        const VT *v = vals + (int64_t)BR * BC * k + BC * r;
        const XT *xx = x + (int64_t)BC * colind[k];
#pragma unroll
        for (int c = 0; c < BC; ++c) s += v[c] * (VT)xx[c];
    }
    return s;
}

// y = A x + beta*y; one thread per (block row, row-in-block)
template <int BR, int BC, class VT, class XT>
__global__ void bsr_spmv_kernel(int64_t nb, const int32_t *__restrict__ rowptr, const int32_t *__restrict__ colind,
                                const VT *__restrict__ vals, const XT *__restrict__ x, XT *__restrict__ y,
                                double beta) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t i = t / BR;
    const int r = (int)(t % BR);
    if (i >= nb) return;
    const VT s = block_row_dot<BR, BC, VT, XT>(rowptr[i], rowptr[i + 1], r, colind, vals, x);
    const int64_t o = (int64_t)BR * i + r;
    y[o] = (XT)(beta == 0.0 ? (double)s : (double)s + beta * (double)y[o]);
}

// One 64-lane wave per block row, for long rows (the restriction R = P^T holds ~10^2
// blocks per coarse row): lanes stride over the row's blocks, so consecutive lanes read
// consecutive blocks, and the BR partial sums are reduced across the wave.  y = A x.
template <int BR, int BC, class VT, class XT>
__global__ __launch_bounds__(256) void bsr_spmv_wave_kernel(int64_t nb, const int32_t *__restrict__ rowptr,
                                                            const int32_t *__restrict__ colind,
                                                            const VT *__restrict__ vals, const XT *__restrict__ x,
                                                            XT *__restrict__ y) {
    const int64_t i = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (i >= nb) return;
    VT acc[BR];
#pragma unroll
    for (int r = 0; r < BR; ++r) acc[r] = 0;
    // two of the lane's blocks per trip, both loaded before either is used (a row has
    // ~10^2 blocks: two trips); the lane still adds its blocks in storage order
    const int k1 = rowptr[i + 1];
    for (int k = rowptr[i] + lane; k < k1; k += 128) {
        const bool two = k + 64 < k1;
        VT va[BR * BC], vb[BR * BC], xa[BC], xb[BC];
        {
            const VT *v = vals + (int64_t)BR * BC * k;
            const XT *xx = x + (int64_t)BC * colind[k];
#pragma unroll
            for (int t = 0; t < BR * BC; ++t) va[t] = v[t];
#pragma unroll
            for (int c = 0; c < BC; ++c) xa[c] = (VT)xx[c];
        }
        if (two) {
            const VT *v = vals + (int64_t)BR * BC * (k + 64);
            const XT *xx = x + (int64_t)BC * colind[k + 64];
#pragma unroll
            for (int t = 0; t < BR * BC; ++t) vb[t] = v[t];
#pragma unroll
            for (int c = 0; c < BC; ++c) xb[c] = (VT)xx[c];
        }
#pragma unroll
        for (int r = 0; r < BR; ++r)
#pragma unroll
            for (int c = 0; c < BC; ++c) acc[r] += va[BC * r + c] * xa[c];
        if (two)
#pragma unroll
            for (int r = 0; r < BR; ++r)
#pragma unroll
                for (int c = 0; c < BC; ++c) acc[r] += vb[BC * r + c] * xb[c];
    }
#pragma unroll
    for (int r = 0; r < BR; ++r)
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) acc[r] += __shfl_xor(acc[r], off, 64);
    if (lane == 0)
#pragma unroll
        for (int r = 0; r < BR; ++r) y[(int64_t)BR * i + r] = (XT)acc[r];
}

// y = A x + beta*y over runs of consecutive block rows ("chunks" of at most
// P_STREAM_CAP blocks, row boundaries listed on the host), for short rows of varying
// length (the fine prolongator: 1-8 blocks per row).  One 64-lane wave per chunk:
// pass 1 gives lane t row r = t % BR of block b = t / BR, so the wave reads the chunk's
// blocks as one contiguous stream and keeps each block row's product with x in LDS;
// pass 2 gives each lane one output row, which sums its blocks' products in order.
template <int BR, int BC, class VT, class XT>
__global__ __launch_bounds__(64) void bsr_stream_kernel(const int32_t *__restrict__ chunk_row,
                                                        const int32_t *__restrict__ rowptr,
                                                        const int32_t *__restrict__ colind,
                                                        const VT *__restrict__ vals, const XT *__restrict__ x,
                                                        XT *__restrict__ y, double beta) {
    __shared__ VT part[P_STREAM_CAP * BR];
    const int lane = threadIdx.x;
    const int32_t r0 = chunk_row[blockIdx.x], r1 = chunk_row[blockIdx.x + 1];
    const int32_t b0 = rowptr[r0];
    const int nt = (rowptr[r1] - b0) * BR;
    for (int t0 = lane; t0 < nt; t0 += 64 * ROW_U) {  // ROW_U products per trip, loads first
        VT a[ROW_U][BC], p[ROW_U][BC];
#pragma unroll
        for (int u = 0; u < ROW_U; ++u) {
            const int t = t0 + 64 * u;
            if (t < nt) {
                const int b = t / BR, r = t - BR * b;
                const int64_t k = (int64_t)b0 + b;
                const VT *v = vals + (int64_t)BR * BC * k + BC * r;
                const XT *xx = x + (int64_t)BC * colind[k];
#pragma unroll
                for (int c = 0; c < BC; ++c) {
                    a[u][c] = v[c];
                    p[u][c] = (VT)xx[c];
                }
            }
        }
#pragma unroll
        for (int u = 0; u < ROW_U; ++u) {
            const int t = t0 + 64 * u;
            if (t < nt) {
                VT sum = 0;
#pragma unroll
                for (int c = 0; c < BC; ++c) sum += a[u][c] * p[u][c];
                part[t] = sum;
            }
        }
    }
    __syncthreads();
    const int no = (r1 - r0) * BR;
    for (int o = lane; o < no; o += 64) {
        const int32_t i = r0 + o / BR;
        const int r = o - BR * (o / BR);
        VT s = 0;
        const int32_t k1 = rowptr[i + 1];
        for (int32_t k = rowptr[i]; k < k1; ++k) s += part[(k - b0) * BR + r];
        const int64_t oi = (int64_t)BR * i + r;
        y[oi] = (XT)(beta == 0.0 ? (double)s : (double)s + beta * (double)y[oi]);
    }
}

template <int BS, class VT, class XT>
__global__ void bsr_residual_kernel(int64_t nb, const int32_t *__restrict__ rowptr, const int32_t *__restrict__ colind,
                                    const VT *__restrict__ vals, const XT *__restrict__ x, const XT *__restrict__ b,
                                    XT *__restrict__ r) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t i = t / BS;
    const int rr = (int)(t % BS);
    if (i >= nb) return;
    const VT s = block_row_dot<BS, BS, VT, XT>(rowptr[i], rowptr[i + 1], rr, colind, vals, x);
    const int64_t o = (int64_t)BS * i + rr;
    r[o] = (XT)((VT)b[o] - s);
}

// One Chebyshev step of a coarse-level smoother with its product in the same pass: the
// thread of row rr of block row i forms (A d)_i,rr (as bsr_spmv_kernel), then the block
// row's BS threads exchange r - A d through LDS and each forms its component of
// d' = c1 d + c2 Dinv (r - A d) and x += d' -- cheb_step_kernel's arithmetic in the same
// order, so results are bitwise equal.  d is read by other block rows' products, so d'
// goes to d_out; d_out == null (the last step of a smoothing) skips the r and d' stores.
// Workgroups of 42 block rows (252 threads) keep every block row's threads together.
constexpr int CHEB_ROWS = 42;
template <int BS, class DT, class XT>
__global__ __launch_bounds__(CHEB_ROWS * BS) void bsr_cheb_kernel(int64_t nb, const int32_t *__restrict__ rowptr,
                                                                  const int32_t *__restrict__ colind,
                                                                  const DT *__restrict__ vals,
                                                                  const XT *__restrict__ d,
                                                                  const DT *__restrict__ dinv, XT *__restrict__ r,
                                                                  XT *__restrict__ d_out, XT *__restrict__ x,
                                                                  double c1, double c2) {
    __shared__ DT srv[CHEB_ROWS * BS];
    const int64_t t = (int64_t)blockIdx.x * (CHEB_ROWS * BS) + threadIdx.x;
    const int64_t i = t / BS;
    const int rr = (int)(t % BS);
    const bool act = i < nb;
    const int64_t o = (int64_t)BS * i + rr;
    DT rv = 0;
    if (act) {
        const DT s = block_row_dot<BS, BS, DT, XT>(rowptr[i], rowptr[i + 1], rr, colind, vals, d);
        rv = (DT)r[o] - (DT)(XT)(double)s;  // bsr_spmv_kernel stores (XT)(double)s
        if (d_out) r[o] = (XT)rv;
    }
    srv[threadIdx.x] = rv;
    __syncthreads();
    if (!act) return;
    const DT *m = dinv + (int64_t)BS * BS * i;
    const DT *rw = srv + (threadIdx.x - rr);
    DT s2 = 0;
#pragma unroll
    for (int c = 0; c < BS; ++c) s2 = fma(m[BS * rr + c], rw[c], s2);
    const double dn = fma(c1, (double)d[o], c2 * (double)s2);
    if (d_out) d_out[o] = (XT)dn;
    x[o] = (XT)((double)x[o] + dn);
}

template <int BS, class DT, class XT>
__global__ void block_diag_apply_kernel(int64_t nb, const DT *__restrict__ dinv, const XT *__restrict__ r,
                                        XT *__restrict__ z, double scale) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t i = t / BS;
    const int rr = (int)(t % BS);
    if (i >= nb) return;
    const DT *m = dinv + (int64_t)BS * BS * i + BS * rr;
    const XT *rv = r + (int64_t)BS * i;
    DT s = 0;
#pragma unroll
    for (int c = 0; c < BS; ++c) s += m[c] * (DT)rv[c];
    z[(int64_t)BS * i + rr] = (XT)(scale * (double)s);
}

// First step of a polynomial smoother, d = scale Dinv v, fused with what surrounds it:
// zero initial guess (v = b): also r = b and x = d; otherwise (v = r): x += d.  Same
// arithmetic as block_diag_apply followed by the copies / the axpby.
template <int BS, class DT, class XT, bool ZERO>
__global__ void diag_first_kernel(int64_t nb, const DT *__restrict__ dinv, const XT *__restrict__ v,
                                  XT *__restrict__ r, XT *__restrict__ d, XT *__restrict__ x, double scale) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t i = t / BS;
    const int rr = (int)(t % BS);
    if (i >= nb) return;
    const DT *m = dinv + (int64_t)BS * BS * i + BS * rr;
    const XT *rv = v + (int64_t)BS * i;
    DT s = 0;
#pragma unroll
    for (int c = 0; c < BS; ++c) s += m[c] * (DT)rv[c];
    const int64_t o = (int64_t)BS * i + rr;
    const XT dv = (XT)(scale * (double)s);
    d[o] = dv;
    if (ZERO) {
        r[o] = v[o];
        x[o] = dv;
    } else {
        // x += d with d already rounded (no fused multiply-add of scale * s into x, which
        // the compiler would otherwise form in FP64)
#pragma clang fp contract(off)
        x[o] = (XT)((double)dv + (double)x[o]);
    }
}

// r -= Ad; d = c1 d + c2 Dinv r; x += d
template <int BS, class DT, class XT>
__global__ void cheb_step_kernel(int64_t nb, const DT *__restrict__ dinv, const XT *__restrict__ Ad,
                                 XT *__restrict__ r, XT *__restrict__ d, XT *__restrict__ x, double c1, double c2) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nb) return;
    DT rv[BS];
#pragma unroll
    for (int c = 0; c < BS; ++c) {
        const int64_t o = (int64_t)BS * i + c;
        rv[c] = (DT)r[o] - (DT)Ad[o];
        r[o] = (XT)rv[c];
    }
    const DT *m = dinv + (int64_t)BS * BS * i;
#pragma unroll
    for (int rr = 0; rr < BS; ++rr) {
        DT s = 0;
#pragma unroll
        for (int c = 0; c < BS; ++c) s = fma(m[BS * rr + c], rv[c], s);
        const int64_t o = (int64_t)BS * i + rr;
        // explicit fused multiply-adds: the fine level's fused product + step
        // (sell_cheb_kernel) rounds the same way
        const double dn = fma(c1, (double)d[o], c2 * (double)s);
        d[o] = (XT)dn;
        x[o] = (XT)((double)x[o] + dn);
    }
}

// x *= 1/sqrt(*nrm2) (power-iteration normalisation without a host round trip)
__global__ void scale_by_rnorm_kernel(int64_t n, const double *__restrict__ nrm2, double *__restrict__ x) {
    const double s = *nrm2 > 0.0 ? 1.0 / sqrt(*nrm2) : 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        x[i] *= s;
}

template <class XT>
__global__ void dense_gemv_kernel(int64_t n, const double *__restrict__ Ainv, const XT *__restrict__ b,
                                  XT *__restrict__ x) {
    // one 64-lane wave per row
    const int64_t row = (int64_t)blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (row >= n) return;
    double s = 0.0;
    for (int64_t c = lane; c < n; c += 64) s += Ainv[row * n + c] * (double)b[c];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    if (lane == 0) x[row] = (XT)s;
}


// ---- coarsest level: dense [M | I] Gauss-Jordan with partial pivoting on the device
template <bool SELL>
__global__ void dense_from_blocks_kernel(int64_t nb, int bs, const int64_t *__restrict__ slice_ptr,
                                         const int32_t *__restrict__ rowlen, const int32_t *__restrict__ rowptr,
                                         const int32_t *__restrict__ colind, const double *__restrict__ A,
                                         int64_t n, double *__restrict__ W) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nb) return;
    RowIt<SELL> row(i, slice_ptr, rowlen, rowptr);
    for (int k = 0; k < row.len; ++k) {
        const int64_t pos = row.pos(k);
        const int64_t j = colind[pos];
        const double *a = A + (int64_t)bs * bs * pos;
        for (int r = 0; r < bs; ++r)
            for (int q = 0; q < bs; ++q) W[(bs * i + r) * 2 * n + bs * j + q] = a[bs * r + q];
    }
}
__global__ __launch_bounds__(64) void dense_prepare_kernel(int64_t n, double *__restrict__ W) {
    // one wave per row: unit right half; a zero row (dropped coarse dof) is decoupled
    // with a unit diagonal
    const int64_t r = blockIdx.x;
    const int lane = threadIdx.x;
    double *row = W + r * 2 * n;
    int nz = 0;
    for (int64_t c = lane; c < n; c += 64) nz |= row[c] != 0.0;
    for (int64_t c = lane; c < n; c += 64) row[n + c] = (c == r) ? 1.0 : 0.0;
    nz = __any(nz);
    if (lane == 0 && !nz) row[r] = 1.0;
}
__global__ __launch_bounds__(1024) void gj_pivot_kernel(int64_t n, int64_t k, double *__restrict__ W,
                                                        double *__restrict__ prow, double *__restrict__ fcol) {
    __shared__ double bv[1024];
    __shared__ int64_t bi[1024];
    const int t = threadIdx.x;
    double best = -1.0;
    int64_t arg = k;
    for (int64_t r = k + t; r < n; r += 1024) {
        const double v = fabs(W[r * 2 * n + k]);
        if (v > best) { best = v; arg = r; }
    }
    bv[t] = best;
    bi[t] = arg;
    __syncthreads();
    for (int off = 512; off > 0; off >>= 1) {
        if (t < off && (bv[t + off] > bv[t] || (bv[t + off] == bv[t] && bi[t + off] < bi[t]))) {
            bv[t] = bv[t + off];
            bi[t] = bi[t + off];
        }
        __syncthreads();
    }
    const int64_t p = bi[0];
    if (p != k)
        for (int64_t c = t; c < 2 * n; c += 1024) {
            const double a = W[k * 2 * n + c];
            W[k * 2 * n + c] = W[p * 2 * n + c];
            W[p * 2 * n + c] = a;
        }
    __syncthreads();
    const double piv = W[k * 2 * n + k];
    const double ip = piv != 0.0 ? 1.0 / piv : 0.0;
    for (int64_t c = t; c < 2 * n; c += 1024) prow[c] = W[k * 2 * n + c] * ip;
    for (int64_t r = t; r < n; r += 1024) fcol[r] = W[r * 2 * n + k];
}
__global__ void gj_update_kernel(int64_t n, int64_t k, double *__restrict__ W, const double *__restrict__ prow,
                                 const double *__restrict__ fcol) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t r = t / (2 * n), c = t % (2 * n);
    if (r >= n) return;
    if (r == k) W[r * 2 * n + c] = prow[c];
    else W[r * 2 * n + c] -= fcol[r] * prow[c];
}
__global__ void gj_extract_kernel(int64_t n, const double *__restrict__ W, double *__restrict__ inv) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * n) return;
    const int64_t r = t / n, c = t % n;
    inv[t] = W[r * 2 * n + n + c];
}

// ---- blocked Gauss-Jordan: b pivots per panel, 3 launches per panel instead of 2 per pivot.
// Panel [k0, k0+b): one workgroup selects the pivot rows S by partial pivoting on the
// panel (kept in LDS) and inverts M = W[S, panel]; then Y = M^-1 W[S, :] replaces the
// pivot rows and every other row r takes W[r, :] -= W[r, panel] Y.  Rows are never
// swapped: the inverse's row c is the right half of the pivot row chosen for column c.
// M^-1 comes out of the panel elimination itself: every panel row carries b tracking
// columns E (zero), the row chosen as the j-th pivot gets E_j += 1 before it is scaled,
// and at the end E[S_j] is row j of M^-1 (Gauss-Jordan on [M | I] with the rows of M in
// pivot order; an unchosen row's E only ever involves columns of earlier pivots).
constexpr int GJ_MAXB = 16;
constexpr int GJ_THREADS = 1024;

template <int B>
__global__ __launch_bounds__(GJ_THREADS) void gj_panel_kernel(int64_t n, int64_t k0, int bb, int panel,
                                                              const double *__restrict__ W, int *__restrict__ used,
                                                              int *__restrict__ ptag, int *__restrict__ pivrow,
                                                              double *__restrict__ L, double *__restrict__ Minv,
                                                              int *__restrict__ S) {
    extern __shared__ double gsm[];
    // the last panel may be narrower (bb < B): its extra columns are zero and never pivoted
    const int b = B;
    const int ld = 2 * B + 1;              // padded row stride (LDS banks)
    double *P = gsm;                       // n x (2b) [panel | E], stride ld
    double *prow = P + n * ld;             // scaled pivot row (2b)
    int *usedL = reinterpret_cast<int *>(prow + 2 * GJ_MAXB);
    __shared__ double wv[GJ_THREADS / 64];
    __shared__ int wi[GJ_THREADS / 64];
    __shared__ int Ss[GJ_MAXB];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int64_t n2 = 2 * n;
    const int b2 = 2 * b;
    for (int64_t idx = t; idx < n * b; idx += GJ_THREADS) {
        const int64_t r = idx / b, j = idx % b;
        const double v = j < bb ? W[r * n2 + k0 + j] : 0.0;
        P[r * ld + j] = v;
        P[r * ld + b + j] = 0.0;
        if (j < bb) L[r * bb + j] = v;
    }
    for (int64_t r = t; r < n; r += GJ_THREADS) usedL[r] = used[r];
    __syncthreads();
    for (int j = 0; j < bb; ++j) {
        // pivot: largest |P[r, j]| over unused rows (any unused row beats none, so a NaN
        // column still yields a valid row); every thread reduces the per-wave winners
        double best = -1.0;
        int arg = (int)n;
        for (int64_t r = t; r < n; r += GJ_THREADS)
            if (!usedL[r]) {
                const double v = fabs(P[r * ld + j]);
                if (arg == (int)n || v > best) { best = v; arg = (int)r; }
            }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            const double ob = __shfl_xor(best, off, 64);
            const int oa = __shfl_xor(arg, off, 64);
            if (oa < (int)n && (arg == (int)n || ob > best || (ob == best && oa < arg))) { best = ob; arg = oa; }
        }
        if (lane == 0) { wv[w] = best; wi[w] = arg; }
        __syncthreads();
        int p = (int)n;
        double pb = -1.0;
        for (int q = 0; q < GJ_THREADS / 64; ++q)
            if (wi[q] < (int)n && (p == (int)n || wv[q] > pb || (wv[q] == pb && wi[q] < p))) { pb = wv[q]; p = wi[q]; }
        const double pv = P[(int64_t)p * ld + j];
        const double ip = pv != 0.0 ? 1.0 / pv : 0.0;
        if (t < b2) prow[t] = (P[(int64_t)p * ld + t] + (t == b + j ? 1.0 : 0.0)) * ip;
        if (t == 0) { usedL[p] = 1; Ss[j] = p; }
        __syncthreads();
        for (int64_t r = t; r < n; r += GJ_THREADS) {  // one row per thread
            double *Pr = P + r * ld;
            if (r == p) {
#pragma unroll
                for (int c = 0; c < 2 * B; ++c) Pr[c] = prow[c];
            } else {
                const double f = Pr[j];
                double v[2 * B];
#pragma unroll
                for (int c = 0; c < 2 * B; ++c) v[c] = Pr[c];
#pragma unroll
                for (int c = 0; c < 2 * B; ++c) Pr[c] = v[c] - f * prow[c];
            }
        }
        __syncthreads();
    }
    for (int idx = t; idx < bb * bb; idx += GJ_THREADS) Minv[idx] = P[(int64_t)Ss[idx / bb] * ld + b + idx % bb];
    if (t < bb) {
        S[t] = Ss[t];
        pivrow[k0 + t] = Ss[t];
        ptag[Ss[t]] = panel * 64 + t + 1;
    }
    for (int64_t r = t; r < n; r += GJ_THREADS) used[r] = usedL[r];
}

// Y[jj, c] = sum_i Minv[jj, i] W[S_i, c] for c in [k0, 2n)
__global__ void gj_rows_kernel(int64_t n, int64_t k0, int b, const int *__restrict__ S,
                               const double *__restrict__ Minv, const double *__restrict__ W,
                               double *__restrict__ Y) {
    const int64_t w = 2 * n - k0;
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= b * w) return;
    const int64_t jj = idx / w, c = k0 + idx % w;
    double s = 0.0;
    for (int i = 0; i < b; ++i) s += Minv[jj * b + i] * W[(int64_t)S[i] * 2 * n + c];
    Y[jj * 2 * n + c] = s;
}

// W[r, c] -= L[r, :] Y[:, c] (pivot rows of this panel: W[S_jj, c] = Y[jj, c]), c in [k0, 2n);
// the columns left of k0 are already eliminated and Y is zero there.
__global__ void gj_apply_kernel(int64_t n, int64_t k0, int b, int panel, const int *__restrict__ ptag,
                                const double *__restrict__ L, const double *__restrict__ Y, double *__restrict__ W) {
    const int64_t w = 2 * n - k0;
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n * w) return;
    const int64_t r = idx / w, c = k0 + idx % w;
    const int tg = ptag[r];
    if (tg > 0 && (tg - 1) / 64 == panel) {
        W[r * 2 * n + c] = Y[(int64_t)((tg - 1) % 64) * 2 * n + c];
        return;
    }
    double s = 0.0;
    for (int jj = 0; jj < b; ++jj) s += L[r * b + jj] * Y[(int64_t)jj * 2 * n + c];
    W[r * 2 * n + c] -= s;
}

__global__ void gj_extract_perm_kernel(int64_t n, const int *__restrict__ pivrow, const double *__restrict__ W,
                                       double *__restrict__ inv) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * n) return;
    const int64_t c = t / n, q = t % n;
    inv[t] = W[(int64_t)pivrow[c] * 2 * n + n + q];
}

size_t gj_panel_lds(int64_t n, int b) {
    return (size_t)(n * (2 * b + 1) + 2 * GJ_MAXB) * sizeof(double) + n * sizeof(int);
}
// Widest panel whose LDS footprint fits (0: none, use the per-pivot path).
int gj_panel_width(int64_t n) {
    for (int b = GJ_MAXB; b >= 4; b /= 2)
        if (gj_panel_lds(n, b) <= 150 * 1024) return b;
    return 0;
}

}  // namespace

size_t gj_scratch_ints(int64_t n) { return (size_t)3 * n + GJ_MAXB; }
size_t gj_scratch_doubles(int64_t n) { return (size_t)n * GJ_MAXB + GJ_MAXB * GJ_MAXB + (size_t)GJ_MAXB * 2 * n; }

void dense_coarse_inverse(int64_t nb, int bs, bool sell, const int64_t *slice_ptr, const int32_t *rowlen,
                          const int32_t *rowptr, const int32_t *colind, const double *A, double *W, double *prow,
                          double *fcol, int *iw, double *dw, double *inv, hipStream_t s) {
    const int64_t n = nb * bs;
    PF_HIP(hipMemsetAsync(W, 0, sizeof(double) * 2 * n * n, s));
    if (sell)
        hipLaunchKernelGGL((dense_from_blocks_kernel<true>), dim3(grid_for(nb, 64)), dim3(64), 0, s, nb, bs,
                           slice_ptr, rowlen, rowptr, colind, A, n, W);
    else
        hipLaunchKernelGGL((dense_from_blocks_kernel<false>), dim3(grid_for(nb, 64)), dim3(64), 0, s, nb, bs,
                           slice_ptr, rowlen, rowptr, colind, A, n, W);
    gauss_jordan_inverse(n, W, prow, fcol, iw, dw, inv, s);
}

void gauss_jordan_inverse(int64_t n, double *W, double *prow, double *fcol, int *iw, double *dw, double *inv,
                          hipStream_t s) {
    hipLaunchKernelGGL(dense_prepare_kernel, dim3((unsigned)n), dim3(64), 0, s, n, W);
    const int b = gj_panel_width(n);
    if (b > 0) {
        // scratch: used, ptag, pivrow (n ints each), S (b ints), L (n x b), Minv (b x b), Y (b x 2n)
        int *used = iw, *ptag = iw + n, *pivrow = iw + 2 * n, *S = iw + 3 * n;
        double *L = dw, *Minv = dw + n * GJ_MAXB, *Y = Minv + GJ_MAXB * GJ_MAXB;
        PF_HIP(hipMemsetAsync(iw, 0, sizeof(int) * 2 * n, s));
        const size_t lds = gj_panel_lds(n, b);
        const void *pk = b == 16 ? (const void *)gj_panel_kernel<16>
                                 : b == 8 ? (const void *)gj_panel_kernel<8> : (const void *)gj_panel_kernel<4>;
        PF_HIP(hipFuncSetAttribute(pk, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        int panel = 0;
        for (int64_t k0 = 0; k0 < n; k0 += b, ++panel) {
            const int bb = (int)std::min<int64_t>(b, n - k0);
            if (b == 16)
                hipLaunchKernelGGL(gj_panel_kernel<16>, dim3(1), dim3(GJ_THREADS), lds, s, n, k0, bb, panel, W, used,
                                   ptag, pivrow, L, Minv, S);
            else if (b == 8)
                hipLaunchKernelGGL(gj_panel_kernel<8>, dim3(1), dim3(GJ_THREADS), lds, s, n, k0, bb, panel, W, used,
                                   ptag, pivrow, L, Minv, S);
            else
                hipLaunchKernelGGL(gj_panel_kernel<4>, dim3(1), dim3(GJ_THREADS), lds, s, n, k0, bb, panel, W, used,
                                   ptag, pivrow, L, Minv, S);
            const int64_t w = 2 * n - k0;
            hipLaunchKernelGGL(gj_rows_kernel, dim3(grid_for(bb * w, 256)), dim3(256), 0, s, n, k0, bb, S, Minv, W, Y);
            hipLaunchKernelGGL(gj_apply_kernel, dim3(grid_for(n * w, 256)), dim3(256), 0, s, n, k0, bb, panel, ptag,
                               L, Y, W);
        }
        hipLaunchKernelGGL(gj_extract_perm_kernel, dim3(grid_for(n * n, 256)), dim3(256), 0, s, n, pivrow, W, inv);
        PF_HIP(hipGetLastError());
        return;
    }
    for (int64_t k = 0; k < n; ++k) {
        hipLaunchKernelGGL(gj_pivot_kernel, dim3(1), dim3(1024), 0, s, n, k, W, prow, fcol);
        hipLaunchKernelGGL(gj_update_kernel, dim3(grid_for(2 * n * n, 256)), dim3(256), 0, s, n, k, W, prow, fcol);
    }
    hipLaunchKernelGGL(gj_extract_kernel, dim3(grid_for(n * n, 256)), dim3(256), 0, s, n, W, inv);
    PF_HIP(hipGetLastError());
}

namespace {
}  // namespace

template <int BS>
void amg_block_inverse(int64_t nb, const int32_t *diagpos, const double *A, double *dinv, hipStream_t s) {
    hipLaunchKernelGGL((block_inverse_kernel<BS>), dim3(grid_for(nb, 128)), dim3(128), 0, s, nb, diagpos, A, dinv);
    PF_HIP(hipGetLastError());
}
template <int BS, bool SELL>
void amg_smooth_P(int64_t nP, const int32_t *P_row, const int32_t *P_colind, const int64_t *slice_ptr,
                  const int32_t *rowlen, const int32_t *rowptr, const int32_t *colind, const double *A,
                  const int32_t *agg, const double *Pt, const double *dinv, double omega, double *P, hipStream_t s) {
    hipLaunchKernelGGL((smooth_P_kernel<BS, SELL>), dim3(grid_for(nP, 128)), dim3(128), 0, s, nP, P_row, P_colind,
                       slice_ptr, rowlen, rowptr, colind, A, agg, Pt, dinv, omega, P);
    PF_HIP(hipGetLastError());
}
template <int BS, bool SELL>
void amg_AP(int64_t nQ, const int32_t *Q_row, const int32_t *Q_colind, const int64_t *slice_ptr, const int32_t *rowlen,
            const int32_t *rowptr, const int32_t *colind, const double *A, const int32_t *P_rowptr,
            const int32_t *P_colind, const double *P, double *Q, hipStream_t s) {
    hipLaunchKernelGGL((AP_kernel<BS, SELL>), dim3(grid_for(nQ, 128)), dim3(128), 0, s, nQ, Q_row, Q_colind,
                       slice_ptr, rowlen, rowptr, colind, A, P_rowptr, P_colind, P, Q);
    PF_HIP(hipGetLastError());
}
template <int BS>
void amg_RAP(int64_t nC, const int32_t *C_row, const int32_t *C_colind, const int32_t *R_rowptr,
             const int32_t *R_colind, const double *R, const int32_t *Q_rowptr, const int32_t *Q_colind,
             const double *Q, double *C, hipStream_t s) {
    hipLaunchKernelGGL((RAP_kernel<BS>), dim3(grid_for(nC, 128)), dim3(128), 0, s, nC, C_row, C_colind, R_rowptr,
                       R_colind, R, Q_rowptr, Q_colind, Q, C);
    PF_HIP(hipGetLastError());
}
template <int BS>
void amg_transpose_P(int64_t nR, const int32_t *R_src, const double *P, double *R, hipStream_t s) {
    hipLaunchKernelGGL((transpose_P_kernel<BS>), dim3(grid_for(nR, 256)), dim3(256), 0, s, nR, R_src, P, R);
    PF_HIP(hipGetLastError());
}
template <int BR, int BC, class VT, class XT>
void bsr_spmv(int64_t nb, const int32_t *rowptr, const int32_t *colind, const VT *vals, const XT *x, XT *y,
              double beta, hipStream_t s) {
    if (nb <= 0) return;
    hipLaunchKernelGGL((bsr_spmv_kernel<BR, BC, VT, XT>), dim3(grid_for(nb * BR, 256)), dim3(256), 0, s, nb, rowptr,
                       colind, vals, x, y, beta);
    PF_HIP(hipGetLastError());
}
template <int BR, int BC, class VT, class XT>
void bsr_spmv_stream(int64_t nchunk, const int32_t *chunk_row, const int32_t *rowptr, const int32_t *colind,
                     const VT *vals, const XT *x, XT *y, double beta, hipStream_t s) {
    if (nchunk <= 0) return;
    hipLaunchKernelGGL((bsr_stream_kernel<BR, BC, VT, XT>), dim3((unsigned)nchunk), dim3(64), 0, s, chunk_row, rowptr,
                       colind, vals, x, y, beta);
    PF_HIP(hipGetLastError());
}
template <int BR, int BC, class VT, class XT>
void bsr_spmv_wave(int64_t nb, const int32_t *rowptr, const int32_t *colind, const VT *vals, const XT *x, XT *y,
                   hipStream_t s) {
    if (nb <= 0) return;
    hipLaunchKernelGGL((bsr_spmv_wave_kernel<BR, BC, VT, XT>), dim3((unsigned)((nb + 3) / 4)), dim3(256), 0, s, nb,
                       rowptr, colind, vals, x, y);
    PF_HIP(hipGetLastError());
}
template <int BS, class VT, class XT>
void bsr_residual(int64_t nb, const int32_t *rowptr, const int32_t *colind, const VT *vals, const XT *x,
                  const XT *b, XT *r, hipStream_t s) {
    hipLaunchKernelGGL((bsr_residual_kernel<BS, VT, XT>), dim3(grid_for(nb * BS, 256)), dim3(256), 0, s, nb, rowptr,
                       colind, vals, x, b, r);
    PF_HIP(hipGetLastError());
}
template <int BS, class VT, class XT>
void sell_residual(const FineMatrix &K, const VT *vals, const XT *x, const XT *b, XT *r, hipStream_t s) {
    static_assert(BS == 3, "fine level is 3x3");
    hipLaunchKernelGGL((sell_lds_kernel<VT, XT, SELL_RES>), dim3((unsigned)K.nslices), dim3(SELL_C), 0, s, K.nb,
                       K.slice_ptr.p, K.colind.p, vals, x, b, r, nullptr);
    PF_HIP(hipGetLastError());
}
template <int BS, class VT, class XT>
void sell_spmv(const FineMatrix &K, const VT *vals, const XT *x, XT *y, hipStream_t s) {
    static_assert(BS == 3, "fine level is 3x3");
    hipLaunchKernelGGL((sell_lds_kernel<VT, XT, SELL_Y>), dim3((unsigned)K.nslices), dim3(SELL_C), 0, s, K.nb,
                       K.slice_ptr.p, K.colind.p, vals, x, (const XT *)nullptr, y, nullptr);
    PF_HIP(hipGetLastError());
}
template <class VT, class XT>
void sell_cheb(const FineMatrix &K, const VT *vals, const XT *d, const VT *dinv, XT *r, XT *d_out, XT *x, double c1,
               double c2, hipStream_t s) {
    hipLaunchKernelGGL((sell_cheb_kernel<VT, XT>), dim3((unsigned)K.nslices), dim3(SELL_C), 0, s, K.nb,
                       K.slice_ptr.p, K.colind.p, vals, d, dinv, r, d_out, x, c1, c2);
    PF_HIP(hipGetLastError());
}
template <int BS, class DT, class XT>
void block_diag_apply(int64_t nb, const DT *dinv, const XT *r, XT *z, double scale, hipStream_t s) {
    hipLaunchKernelGGL((block_diag_apply_kernel<BS, DT, XT>), dim3(grid_for(nb * BS, 256)), dim3(256), 0, s, nb,
                       dinv, r, z, scale);
    PF_HIP(hipGetLastError());
}
template <int BS, class DT, class XT>
void bsr_cheb(int64_t nb, const int32_t *rowptr, const int32_t *colind, const DT *vals, const XT *d, const DT *dinv,
              XT *r, XT *d_out, XT *x, double c1, double c2, hipStream_t s) {
    const int64_t threads = nb * BS;
    hipLaunchKernelGGL((bsr_cheb_kernel<BS, DT, XT>), dim3((unsigned)((threads + CHEB_ROWS * BS - 1) / (CHEB_ROWS * BS))),
                       dim3(CHEB_ROWS * BS), 0, s, nb, rowptr, colind, vals, d, dinv, r, d_out, x, c1, c2);
    PF_HIP(hipGetLastError());
}
template <int BS, class DT, class XT>
void cheb_step(int64_t nb, const DT *dinv, const XT *Ad, XT *r, XT *d, XT *x, double c1, double c2, hipStream_t s) {
    hipLaunchKernelGGL((cheb_step_kernel<BS, DT, XT>), dim3(grid_for(nb, 256)), dim3(256), 0, s, nb, dinv, Ad, r, d,
                       x, c1, c2);
    PF_HIP(hipGetLastError());
}
void scale_by_rnorm(int64_t n, const double *nrm2, double *x, hipStream_t s) {
    int64_t g = (n + 255) / 256;
    if (g > 4096) g = 4096;
    hipLaunchKernelGGL(scale_by_rnorm_kernel, dim3((unsigned)(g < 1 ? 1 : g)), dim3(256), 0, s, n, nrm2, x);
    PF_HIP(hipGetLastError());
}

template <class XT>
void dense_gemv(int64_t n, const double *Ainv, const XT *b, XT *x, hipStream_t s) {
    hipLaunchKernelGGL((dense_gemv_kernel<XT>), dim3(grid_for(n, 4)), dim3(256), 0, s, n, Ainv, b, x);
    PF_HIP(hipGetLastError());
}

template <int BS, class DT, class XT>
void diag_first(int64_t nb, const DT *dinv, const XT *v, XT *r, XT *d, XT *x, double scale, bool zero_init,
                hipStream_t s) {
    if (zero_init)
        hipLaunchKernelGGL((diag_first_kernel<BS, DT, XT, true>), dim3(grid_for(nb * BS, 256)), dim3(256), 0, s, nb,
                           dinv, v, r, d, x, scale);
    else
        hipLaunchKernelGGL((diag_first_kernel<BS, DT, XT, false>), dim3(grid_for(nb * BS, 256)), dim3(256), 0, s, nb,
                           dinv, v, r, d, x, scale);
    PF_HIP(hipGetLastError());
}
template void diag_first<3, double, double>(int64_t, const double *, const double *, double *, double *, double *,
                                            double, bool, hipStream_t);
template void diag_first<6, double, double>(int64_t, const double *, const double *, double *, double *, double *,
                                            double, bool, hipStream_t);
template void diag_first<3, float, float>(int64_t, const float *, const float *, float *, float *, float *, double,
                                          bool, hipStream_t);
template void diag_first<6, float, float>(int64_t, const float *, const float *, float *, float *, float *, double,
                                          bool, hipStream_t);

// explicit instantiations
template void amg_block_inverse<3>(int64_t, const int32_t *, const double *, double *, hipStream_t);
template void amg_block_inverse<6>(int64_t, const int32_t *, const double *, double *, hipStream_t);
#define PF_SM(BS, SELL)                                                                                            \
    template void amg_smooth_P<BS, SELL>(int64_t, const int32_t *, const int32_t *, const int64_t *,                \
                                         const int32_t *, const int32_t *, const int32_t *, const double *,         \
                                         const int32_t *, const double *, const double *, double, double *,         \
                                         hipStream_t);                                                              \
    template void amg_AP<BS, SELL>(int64_t, const int32_t *, const int32_t *, const int64_t *, const int32_t *,     \
                                   const int32_t *, const int32_t *, const double *, const int32_t *,               \
                                   const int32_t *, const double *, double *, hipStream_t);
PF_SM(3, true)
PF_SM(6, false)
#undef PF_SM
template void amg_RAP<3>(int64_t, const int32_t *, const int32_t *, const int32_t *, const int32_t *, const double *,
                         const int32_t *, const int32_t *, const double *, double *, hipStream_t);
template void amg_RAP<6>(int64_t, const int32_t *, const int32_t *, const int32_t *, const int32_t *, const double *,
                         const int32_t *, const int32_t *, const double *, double *, hipStream_t);
template void amg_transpose_P<3>(int64_t, const int32_t *, const double *, double *, hipStream_t);
template void amg_transpose_P<6>(int64_t, const int32_t *, const double *, double *, hipStream_t);
#define PF_OPS(VT, XT)                                                                                             \
    template void bsr_spmv<6, 6, VT, XT>(int64_t, const int32_t *, const int32_t *, const VT *, const XT *, XT *,  \
                                         double, hipStream_t);                                                     \
    template void bsr_spmv<3, 6, VT, XT>(int64_t, const int32_t *, const int32_t *, const VT *, const XT *, XT *,  \
                                         double, hipStream_t);                                                     \
    template void bsr_spmv<6, 3, VT, XT>(int64_t, const int32_t *, const int32_t *, const VT *, const XT *, XT *,  \
                                         double, hipStream_t);                                                     \
    template void bsr_spmv_stream<3, 6, VT, XT>(int64_t, const int32_t *, const int32_t *, const int32_t *,        \
                                                const VT *, const XT *, XT *, double, hipStream_t);                \
    template void bsr_spmv_wave<6, 3, VT, XT>(int64_t, const int32_t *, const int32_t *, const VT *, const XT *,    \
                                              XT *, hipStream_t);                                                  \
    template void bsr_spmv_wave<6, 6, VT, XT>(int64_t, const int32_t *, const int32_t *, const VT *, const XT *,    \
                                              XT *, hipStream_t);                                                  \
    template void bsr_residual<6, VT, XT>(int64_t, const int32_t *, const int32_t *, const VT *, const XT *,       \
                                          const XT *, XT *, hipStream_t);                                          \
    template void sell_residual<3, VT, XT>(const FineMatrix &, const VT *, const XT *, const XT *, XT *,           \
                                           hipStream_t);                                                           \
    template void sell_spmv<3, VT, XT>(const FineMatrix &, const VT *, const XT *, XT *, hipStream_t);             \
    template void sell_cheb<VT, XT>(const FineMatrix &, const VT *, const XT *, const VT *, XT *, XT *, XT *,      \
                                    double, double, hipStream_t);                                                  \
    template void block_diag_apply<3, VT, XT>(int64_t, const VT *, const XT *, XT *, double, hipStream_t);         \
    template void block_diag_apply<6, VT, XT>(int64_t, const VT *, const XT *, XT *, double, hipStream_t);         \
    template void cheb_step<3, VT, XT>(int64_t, const VT *, const XT *, XT *, XT *, XT *, double, double,          \
                                       hipStream_t);                                                               \
    template void cheb_step<6, VT, XT>(int64_t, const VT *, const XT *, XT *, XT *, XT *, double, double,          \
                                       hipStream_t);                                                               \
    template void bsr_cheb<6, VT, XT>(int64_t, const int32_t *, const int32_t *, const VT *, const XT *,          \
                                      const VT *, XT *, XT *, XT *, double, double, hipStream_t);
PF_OPS(double, double)
PF_OPS(float, float)
#undef PF_OPS
template void dense_gemv<double>(int64_t, const double *, const double *, double *, hipStream_t);
template void dense_gemv<float>(int64_t, const double *, const float *, float *, hipStream_t);

}  // namespace pfem
