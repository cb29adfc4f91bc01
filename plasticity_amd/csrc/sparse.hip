// sparse.hip -- fine-level block-sparse kernels for gfx950 (HBM-bound work).
//
//  * SELL-64 BSR(3x3) SpMV, FP64 and FP32: one thread per block row, the 64 rows of
//    a slice read their k-th blocks as one contiguous 64 x 72 B (FP64) stream, so
//    every wavefront load is fully coalesced (replaces the CSC mul! of
//    Solver.jl:40-50 and Krylov's operator products).
//  * Deterministic reductions: fixed grid, per-block partials, one finishing block.
//  * Dirichlet imposition on the device K/R (BoundaryConditions.jl:62-109) over the
//    precomputed list of block rows the constraints touch.
//  * Block-diagonal (3x3) inverse and block-Jacobi application.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "common.h"
#include "kernels.h"
#include "sell.h"

namespace pfem {
namespace {

template <class T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

// Block-wide sum of a double (blockDim.x multiple of 64, <= 1024); result in thread 0.
__device__ __forceinline__ double block_sum(double v) {
    __shared__ double red[16];
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) red[wid] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0) {
        const int nw = blockDim.x >> 6;
        for (int i = 0; i < nw; ++i) s += red[i];
    }
    __syncthreads();
    return s;
}

// ---------------------------------------------------------------- reductions
constexpr int RED_BLOCKS = 1024;
constexpr int MV_MAX = 16;  // vectors per fused Gram-Schmidt pass

template <class A, class B>
__global__ __launch_bounds__(256) void dot_partial_kernel(int64_t n, const A *__restrict__ a,
                                                          const B *__restrict__ b, double *__restrict__ partial) {
    double s = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
        s += (double)a[i] * (double)b[i];
    s = block_sum(s);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

__global__ __launch_bounds__(1024) void finish_sum_kernel(int n, const double *__restrict__ partial,
                                                          double *__restrict__ out) {
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += 1024) s += partial[i];
    s = block_sum(s);
    if (threadIdx.x == 0) *out = s;
}

// Up to MV_MAX dot products a . v_j in one pass over a (the Gram-Schmidt step of the
// flexible GCR); per-thread order, grid and finishing sum as dot_partial_kernel /
// finish_sum_kernel, so each result equals the separate dot product bit for bit.
struct MultiVec {
    const double *v[MV_MAX];
    int nv;
};
__global__ __launch_bounds__(256) void multi_dot_partial_kernel(int64_t n, const double *__restrict__ a, MultiVec V,
                                                                double *__restrict__ partial) {
    double s[MV_MAX];
#pragma unroll
    for (int j = 0; j < MV_MAX; ++j) s[j] = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
        const double ai = a[i];
#pragma unroll
        for (int j = 0; j < MV_MAX; ++j)
            if (j < V.nv) s[j] += ai * V.v[j][i];
    }
#pragma unroll
    for (int j = 0; j < MV_MAX; ++j)
        if (j < V.nv) {
            const double t = block_sum(s[j]);
            if (threadIdx.x == 0) partial[(int64_t)j * gridDim.x + blockIdx.x] = t;
        }
}
// block j sums row j of the partials
__global__ __launch_bounds__(1024) void multi_finish_kernel(int nblk, const double *__restrict__ partial,
                                                            double *__restrict__ out) {
    const double *row = partial + (int64_t)blockIdx.x * nblk;
    double s = 0.0;
    for (int i = threadIdx.x; i < nblk; i += 1024) s += row[i];
    s = block_sum(s);
    if (threadIdx.x == 0) out[blockIdx.x] = s;
}
// p -= sum_j c_j P_j and q -= sum_j c_j Q_j in one pass (the j-th step is the axpby
// y = -c_j x + y, in the same order)
struct MultiAxpy {
    const double *P[MV_MAX];
    const double *Q[MV_MAX];
    double c[MV_MAX];
    int nv;
};
__global__ __launch_bounds__(256) void multi_axpy2_kernel(int64_t n, MultiAxpy V, double *__restrict__ p,
                                                          double *__restrict__ q) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        double pa = p[i], qa = q[i];
#pragma unroll
        for (int j = 0; j < MV_MAX; ++j)
            if (j < V.nv) {
                pa = -V.c[j] * V.P[j][i] + pa;
                qa = -V.c[j] * V.Q[j][i] + qa;
            }
        p[i] = pa;
        q[i] = qa;
    }
}
// x += a p; r -= a q; per-block partials of r . r (grid and order of dot_partial_kernel)
__global__ __launch_bounds__(256) void update_dot_kernel(int64_t n, double a, const double *__restrict__ p,
                                                         const double *__restrict__ q, double *__restrict__ x,
                                                         double *__restrict__ r, double *__restrict__ partial) {
    double s = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
        x[i] = a * p[i] + x[i];
        const double ri = -a * q[i] + r[i];
        r[i] = ri;
        s += ri * ri;
    }
    s = block_sum(s);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// ---------------------------------------------------------------- vector ops
template <class T>
__global__ void axpy_ratio_kernel(int64_t n, const double *num, const double *den, double sign,
                                  const T *__restrict__ x, T *__restrict__ y) {
    // y += sign * (num/den) * x
    const double al = sign * (*num / *den);
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        y[i] = (T)((double)y[i] + al * (double)x[i]);
}
// y += sign * a * x with a = num/den, or a = 0 when den <= 0 (a search direction of
// non-positive curvature contributes nothing)
__global__ void axpy_safe_ratio_kernel(int64_t n, const double *num, const double *den, double sign,
                                       const double *__restrict__ x, double *__restrict__ y) {
    const double a = *den > 0 ? sign * (*num / *den) : 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        y[i] += a * x[i];
}
template <class T>
__global__ void xpby_ratio_kernel(int64_t n, const double *num, const double *den, const T *__restrict__ z,
                                  T *__restrict__ p) {
    // p = z + (num/den) p
    const double be = *num / *den;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = (T)((double)z[i] + be * (double)p[i]);
}
// x += a p ; r -= a q  (a = num/den), fused
template <class T>
__global__ void cg_update_kernel(int64_t n, const double *num, const double *den, const T *__restrict__ p,
                                 const T *__restrict__ q, T *__restrict__ x, T *__restrict__ r) {
    const double al = *num / *den;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        x[i] = (T)((double)x[i] + al * (double)p[i]);
        r[i] = (T)((double)r[i] - al * (double)q[i]);
    }
}
// x64 += a p32 ; r32 -= a q32  (a = num/den): FP32 Krylov vectors, FP64 solution
__global__ void cg_update_mixed_kernel(int64_t n, const double *num, const double *den, const float *__restrict__ p,
                                       const float *__restrict__ q, double *__restrict__ x, float *__restrict__ r) {
    const double al = *num / *den;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        x[i] += al * (double)p[i];
        r[i] = (float)((double)r[i] - al * (double)q[i]);
    }
}
// cg_update_kernel<double> that also writes r32 = (float) r: the FP32 copy of the
// residual the FP32 V-cycle takes, without a separate conversion pass
__global__ void cg_update_r32_kernel(int64_t n, const double *num, const double *den, const double *__restrict__ p,
                                     const double *__restrict__ q, double *__restrict__ x, double *__restrict__ r,
                                     float *__restrict__ r32) {
    const double al = *num / *den;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        x[i] = x[i] + al * p[i];
        const double ri = r[i] - al * q[i];
        r[i] = ri;
        r32[i] = (float)ri;
    }
}
// The three dot products of a flexible-CG step in one pass, r'z, r'r and z'q, with z
// read as ZT (FP32: the V-cycle's output, written out as FP64 z); per-thread order and
// grid as dot_partial_kernel, so each equals the separate dot product bit for bit.
// partial: 3 rows of gridDim.x.
template <class ZT>
__global__ __launch_bounds__(256) void fcg_dots_kernel(int64_t n, const double *__restrict__ r,
                                                       const ZT *__restrict__ zin, double *__restrict__ zout,
                                                       const double *__restrict__ q, double *__restrict__ partial) {
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
        const double ri = r[i], zi = (double)zin[i];
        if (zout) zout[i] = zi;
        s0 += ri * zi;
        s1 += ri * ri;
        s2 += zi * q[i];
    }
    s0 = block_sum(s0);
    s1 = block_sum(s1);
    s2 = block_sum(s2);
    if (threadIdx.x == 0) {
        partial[blockIdx.x] = s0;
        partial[gridDim.x + blockIdx.x] = s1;
        partial[2 * gridDim.x + blockIdx.x] = s2;
    }
}
template <class T>
__global__ void axpby_kernel(int64_t n, double a, const T *__restrict__ x, double b, T *__restrict__ y) {
    // b = 0 overwrites y without reading it: y may hold anything (a fresh buffer), and
    // 0 * NaN would be NaN
    if (b == 0.0) {
        for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
            y[i] = (T)(a * (double)x[i]);
        return;
    }
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        y[i] = (T)(a * (double)x[i] + b * (double)y[i]);
}
// axpby with b = -(num/den) read on the device: the same expression (and the same b = 0
// overwrite) as axpby_kernel, so a flexible-CG direction update formed here equals the
// one the host forms from the same scalars
template <class T>
__global__ void axpby_negratio_kernel(int64_t n, double a, const T *__restrict__ x, const double *num,
                                      const double *den, T *__restrict__ y) {
    const double b = -(*num / *den);
    if (b == 0.0) {
        for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
            y[i] = (T)(a * (double)x[i]);
        return;
    }
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        y[i] = (T)(a * (double)x[i] + b * (double)y[i]);
}
template <class S, class D>
__global__ void convert_kernel(int64_t n, const S *__restrict__ s, D *__restrict__ d) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        d[i] = (D)s[i];
}
template <class T>
__global__ void fill_kernel(int64_t n, T v, T *__restrict__ d) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        d[i] = v;
}

// R -= lambda*Fext; partial sum of R^2
__global__ __launch_bounds__(256) void resid_ext_kernel(int64_t n, double lambda, const double *__restrict__ fext,
                                                        double *__restrict__ R) {
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
        R[i] -= lambda * fext[i];
}

// ---------------------------------------------------------------- Dirichlet
// bcflag[d]: 0 free, 1 fixed (value as given), 2 ramped (lambda * value).  One
// impose_dirichlet! call treats the dofs whose flag is in `sel` as constrained
// (sel = 2: the ramped set, 1: the fixed set, 3: both at once).
__device__ __forceinline__ bool isbc(const uint8_t *f, int64_t d, int sel) { return (f[d] & sel) != 0; }

__global__ void bc_dg_kernel(int64_t n, const int64_t *__restrict__ bcdofs, const uint8_t *__restrict__ flag,
                             const double *__restrict__ val, double lambda, const double *__restrict__ U,
                             double *__restrict__ dg) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int64_t d = bcdofs[t];
    const double g = flag[d] == 2 ? lambda * val[d] : val[d];
    dg[d] = g - U[d];
}

// One thread per affected block row: move known columns to the RHS of free rows,
// then zero constrained rows/columns with a unit diagonal (BoundaryConditions.jl:83-101).
__global__ void bc_rows_kernel(int64_t nrows, const int32_t *__restrict__ rows,
                               const int64_t *__restrict__ slice_ptr, const int32_t *__restrict__ rowlen,
                               const int32_t *__restrict__ colind, double *__restrict__ vals,
                               const uint8_t *__restrict__ flag, const double *__restrict__ dg,
                               double *__restrict__ R, int sel) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nrows) return;
    const int64_t i = rows[t];
    const int64_t base = slice_ptr[i >> 6] + (i & 63);
    const int len = rowlen[i];
    bool rb[3];
    double acc[3] = {0, 0, 0};
    for (int r = 0; r < 3; ++r) rb[r] = isbc(flag, 3 * i + r, sel);
    for (int k = 0; k < len; ++k) {
        const int64_t pos = base + (int64_t)k * SELL_C;
        const int64_t j = colind[pos];
        double *v = vals + 9 * pos;
        for (int c = 0; c < 3; ++c) {
            const int64_t cd = 3 * j + c;
            if (!isbc(flag, cd, sel)) continue;
            for (int r = 0; r < 3; ++r)
                if (!rb[r]) acc[r] += v[3 * r + c] * dg[cd];
        }
    }
    for (int r = 0; r < 3; ++r)
        if (!rb[r]) R[3 * i + r] += acc[r];
    for (int k = 0; k < len; ++k) {
        const int64_t pos = base + (int64_t)k * SELL_C;
        const int64_t j = colind[pos];
        double *v = vals + 9 * pos;
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) {
                const int64_t rd = 3 * i + r, cd = 3 * j + c;
                if (rb[r] || isbc(flag, cd, sel)) v[3 * r + c] = (rd == cd && rb[r]) ? 1.0 : 0.0;
            }
    }
}

__global__ void bc_rhs_kernel(int64_t n, const int64_t *__restrict__ bcdofs, const uint8_t *__restrict__ flag,
                              const double *__restrict__ dg, double *__restrict__ R, int sel) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int64_t d = bcdofs[t];
    if (isbc(flag, d, sel)) R[d] = -dg[d];
}

// ---------------------------------------------------------------- block Jacobi (3x3)
__global__ void diag_inv3_kernel(int64_t nb, const int32_t *__restrict__ diagpos, const double *__restrict__ vals,
                                 double *__restrict__ dinv) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nb) return;
    const double *m = vals + 9 * (int64_t)diagpos[i];  // row-major block
    const double a = m[0], b = m[1], c = m[2], d = m[3], e = m[4], f = m[5], g = m[6], h = m[7], k = m[8];
    const double A = e * k - f * h, B = -(d * k - f * g), C = d * h - e * g;
    const double det = a * A + b * B + c * C;
    double *o = dinv + 9 * i;
    if (det == 0.0) {
        for (int t = 0; t < 9; ++t) o[t] = 0.0;
        return;
    }
    const double id = 1.0 / det;
    o[0] = A * id; o[1] = -(b * k - c * h) * id; o[2] = (b * f - c * e) * id;
    o[3] = B * id; o[4] = (a * k - c * g) * id;  o[5] = -(a * f - c * d) * id;
    o[6] = C * id; o[7] = -(a * h - b * g) * id; o[8] = (a * e - b * d) * id;
}

template <class T>
__global__ void block_jacobi3_kernel(int64_t nb, const double *__restrict__ dinv, const T *__restrict__ r,
                                     T *__restrict__ z) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nb) return;
    const double *m = dinv + 9 * i;
    const double r0 = r[3 * i], r1 = r[3 * i + 1], r2 = r[3 * i + 2];
    z[3 * i] = (T)(m[0] * r0 + m[1] * r1 + m[2] * r2);
    z[3 * i + 1] = (T)(m[3] * r0 + m[4] * r1 + m[5] * r2);
    z[3 * i + 2] = (T)(m[6] * r0 + m[7] * r1 + m[8] * r2);
}

inline unsigned grid_for(int64_t n, int bs) {
    int64_t g = (n + bs - 1) / bs;
    if (g < 1) g = 1;
    return (unsigned)g;
}
inline unsigned grid_strided(int64_t n) {
    int64_t g = (n + 255) / 256;
    if (g > 4096) g = 4096;
    if (g < 1) g = 1;
    return (unsigned)g;
}

}  // namespace

// ---------------------------------------------------------------- launchers
void spmv_fine64(const FineMatrix &K, const double *x, double *y, hipStream_t s) {
    hipLaunchKernelGGL((sell_lds_kernel<double, double, SELL_Y>), dim3((unsigned)K.nslices), dim3(SELL_C), 0, s,
                       K.nb, K.slice_ptr.p, K.colind.p, K.v64.p, x, (const double *)nullptr, y, nullptr);
    PF_HIP(hipGetLastError());
}
void spmv_fine32(const FineMatrix &K, const float *x, float *y, hipStream_t s) {
    hipLaunchKernelGGL((sell_lds_kernel<float, float, SELL_Y>), dim3((unsigned)K.nslices), dim3(SELL_C), 0, s,
                       K.nb, K.slice_ptr.p, K.colind.p, K.v32.p, x, (const float *)nullptr, y, nullptr);
    PF_HIP(hipGetLastError());
}
void spmv_fine64_dot(const FineMatrix &K, const double *x, double *y, double *partial, double *out,
                     hipStream_t s) {
    hipLaunchKernelGGL((sell_lds_kernel<double, double, SELL_DOT>), dim3((unsigned)K.nslices), dim3(SELL_C), 0, s,
                       K.nb, K.slice_ptr.p, K.colind.p, K.v64.p, x, (const double *)nullptr, y, partial);
    hipLaunchKernelGGL(finish_sum_kernel, dim3(1), dim3(1024), 0, s, (int)K.nslices, partial, out);
    PF_HIP(hipGetLastError());
}
void spmv_fine32_dot(const FineMatrix &K, const float *x, float *y, double *partial, double *out,
                     hipStream_t s) {
    hipLaunchKernelGGL((sell_lds_kernel<float, float, SELL_DOT>), dim3((unsigned)K.nslices), dim3(SELL_C), 0, s,
                       K.nb, K.slice_ptr.p, K.colind.p, K.v32.p, x, (const float *)nullptr, y, partial);
    hipLaunchKernelGGL(finish_sum_kernel, dim3(1), dim3(1024), 0, s, (int)K.nslices, partial, out);
    PF_HIP(hipGetLastError());
}
void residual_fine64(const FineMatrix &K, const double *x, const double *b, double *r, hipStream_t s) {
    hipLaunchKernelGGL((sell_lds_kernel<double, double, SELL_RES>), dim3((unsigned)K.nslices), dim3(SELL_C), 0, s,
                       K.nb, K.slice_ptr.p, K.colind.p, K.v64.p, x, b, r, nullptr);
    PF_HIP(hipGetLastError());
}

size_t reduce_partials_needed(int64_t nb) {
    size_t a = (size_t)((nb + SELL_C - 1) / SELL_C);
    return a > (size_t)RED_BLOCKS ? a : (size_t)RED_BLOCKS;
}

template <class A, class B>
void dot_dev(int64_t n, const A *a, const B *b, double *partial, double *out, hipStream_t s) {
    hipLaunchKernelGGL((dot_partial_kernel<A, B>), dim3(RED_BLOCKS), dim3(256), 0, s, n, a, b, partial);
    hipLaunchKernelGGL(finish_sum_kernel, dim3(1), dim3(1024), 0, s, RED_BLOCKS, partial, out);
    PF_HIP(hipGetLastError());
}
template void dot_dev<double, double>(int64_t, const double *, const double *, double *, double *, hipStream_t);
template void dot_dev<float, float>(int64_t, const float *, const float *, double *, double *, hipStream_t);

size_t multi_dot_partials() { return (size_t)MV_MAX * RED_BLOCKS; }

void multi_dot(int64_t n, const double *a, const double *const *v, int nv, double *partial, double *out,
               hipStream_t s) {
    for (int j0 = 0; j0 < nv; j0 += MV_MAX) {
        MultiVec V{};
        V.nv = std::min(MV_MAX, nv - j0);
        for (int j = 0; j < V.nv; ++j) V.v[j] = v[j0 + j];
        hipLaunchKernelGGL(multi_dot_partial_kernel, dim3(RED_BLOCKS), dim3(256), 0, s, n, a, V, partial);
        hipLaunchKernelGGL(multi_finish_kernel, dim3(V.nv), dim3(1024), 0, s, RED_BLOCKS, partial, out + j0);
    }
    PF_HIP(hipGetLastError());
}

void multi_axpy2(int64_t n, const double *const *P, const double *const *Q, const double *c, int nv, double *p,
                 double *q, hipStream_t s) {
    for (int j0 = 0; j0 < nv; j0 += MV_MAX) {
        MultiAxpy V{};
        V.nv = std::min(MV_MAX, nv - j0);
        for (int j = 0; j < V.nv; ++j) {
            V.P[j] = P[j0 + j];
            V.Q[j] = Q[j0 + j];
            V.c[j] = c[j0 + j];
        }
        hipLaunchKernelGGL(multi_axpy2_kernel, dim3(grid_strided(n)), dim3(256), 0, s, n, V, p, q);
    }
    PF_HIP(hipGetLastError());
}

void update_dot(int64_t n, double a, const double *p, const double *q, double *x, double *r, double *partial,
                double *out, hipStream_t s) {
    hipLaunchKernelGGL(update_dot_kernel, dim3(RED_BLOCKS), dim3(256), 0, s, n, a, p, q, x, r, partial);
    hipLaunchKernelGGL(finish_sum_kernel, dim3(1), dim3(1024), 0, s, RED_BLOCKS, partial, out);
    PF_HIP(hipGetLastError());
}

template <class T>
void axpy_ratio(int64_t n, const double *num, const double *den, double sign, const T *x, T *y, hipStream_t s) {
    hipLaunchKernelGGL((axpy_ratio_kernel<T>), dim3(grid_strided(n)), dim3(256), 0, s, n, num, den, sign, x, y);
    PF_HIP(hipGetLastError());
}
template <class T>
void xpby_ratio(int64_t n, const double *num, const double *den, const T *z, T *p, hipStream_t s) {
    hipLaunchKernelGGL((xpby_ratio_kernel<T>), dim3(grid_strided(n)), dim3(256), 0, s, n, num, den, z, p);
    PF_HIP(hipGetLastError());
}
void cg_update_r32(int64_t n, const double *num, const double *den, const double *p, const double *q, double *x,
                   double *r, float *r32, hipStream_t s) {
    hipLaunchKernelGGL(cg_update_r32_kernel, dim3(grid_strided(n)), dim3(256), 0, s, n, num, den, p, q, x, r, r32);
    PF_HIP(hipGetLastError());
}
template <class ZT>
void fcg_dots(int64_t n, const double *r, const ZT *z, double *zout, const double *q, double *partial, double *out,
              hipStream_t s) {
    hipLaunchKernelGGL((fcg_dots_kernel<ZT>), dim3(RED_BLOCKS), dim3(256), 0, s, n, r, z, zout, q, partial);
    hipLaunchKernelGGL(multi_finish_kernel, dim3(3), dim3(1024), 0, s, RED_BLOCKS, partial, out);
    PF_HIP(hipGetLastError());
}
template void fcg_dots<double>(int64_t, const double *, const double *, double *, const double *, double *,
                               double *, hipStream_t);
template void fcg_dots<float>(int64_t, const double *, const float *, double *, const double *, double *, double *,
                              hipStream_t);
// r'z, r'r and z'q of FP32 vectors in one pass, accumulated in FP64 (the grid and the
// per-thread order of dot_partial_kernel)
__global__ __launch_bounds__(256) void fcg_dots32_kernel(int64_t n, const float *__restrict__ r,
                                                         const float *__restrict__ z, const float *__restrict__ q,
                                                         double *__restrict__ partial) {
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
        const double ri = r[i], zi = z[i];
        s0 += ri * zi;
        s1 += ri * ri;
        s2 += zi * (double)q[i];
    }
    s0 = block_sum(s0);
    s1 = block_sum(s1);
    s2 = block_sum(s2);
    if (threadIdx.x == 0) {
        partial[blockIdx.x] = s0;
        partial[RED_BLOCKS + blockIdx.x] = s1;
        partial[2 * RED_BLOCKS + blockIdx.x] = s2;
    }
}
void fcg_dots32(int64_t n, const float *r, const float *z, const float *q, double *partial, double *out,
                hipStream_t s) {
    hipLaunchKernelGGL(fcg_dots32_kernel, dim3(RED_BLOCKS), dim3(256), 0, s, n, r, z, q, partial);
    hipLaunchKernelGGL(multi_finish_kernel, dim3(3), dim3(1024), 0, s, RED_BLOCKS, partial, out);
    PF_HIP(hipGetLastError());
}
template <class T>
void cg_update(int64_t n, const double *num, const double *den, const T *p, const T *q, T *x, T *r,
               hipStream_t s) {
    hipLaunchKernelGGL((cg_update_kernel<T>), dim3(grid_strided(n)), dim3(256), 0, s, n, num, den, p, q, x, r);
    PF_HIP(hipGetLastError());
}
// sinv = sqrt|K_ii| (1 where the diagonal is 0), then the FP16 bits of Khat = S K S.
__global__ void half_scale_diag_kernel(int64_t nb, const int32_t *__restrict__ diagpos, const double *__restrict__ v,
                                       float *__restrict__ sinv) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= 3 * nb) return;
    const int64_t i = q / 3;
    const int c = (int)(q % 3);
    const double d = fabs(v[9 * (int64_t)diagpos[i] + 4 * c]);
    sinv[q] = d > 0 ? (float)sqrt(d) : 1.f;
}
__global__ __launch_bounds__(64) void half_scale_values_kernel(int64_t nb, const int64_t *__restrict__ slice_ptr,
                                                               const int32_t *__restrict__ colind,
                                                               const double *__restrict__ v,
                                                               const float *__restrict__ sinv,
                                                               _Float16 *__restrict__ h) {
    const int lane = threadIdx.x;
    const int64_t s = blockIdx.x, i = s * SELL_C + lane;
    const int64_t base = slice_ptr[s];
    const int w = (int)((slice_ptr[s + 1] - base) / SELL_C);
    for (int k = 0; k < w; ++k) {
        const int64_t pos = base + (int64_t)k * SELL_C + lane;
        const int64_t j = colind[pos];
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const double val = i < nb ? v[9 * pos + 3 * r + c] / ((double)sinv[3 * i + r] * sinv[3 * j + c]) : 0.0;
                h[9 * pos + 3 * r + c] = (_Float16)val;
            }
    }
}
// v32 = fl32(v64) and lo16 = bf16(v64 - v32) (round to nearest even) in one pass, one
// thread per slot: its 9 values are contiguous in v64 and v32, and lo16 holds 9 planes of
// 64 per slice column (slice columns start at multiples of 64), so the 64 threads of a
// slice column write each plane as 128 contiguous bytes
__global__ __launch_bounds__(256) void split_copy_kernel(int64_t nslots, const double *__restrict__ v64,
                                                         float *__restrict__ v32, uint16_t *__restrict__ lo16) {
    const int64_t slot = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (slot >= nslots) return;
    const int64_t lane = slot & (SELL_C - 1), base = 9 * (slot - lane) + lane;
#pragma unroll
    for (int m = 0; m < 9; ++m) {
        const double v = v64[9 * slot + m];
        const float h = (float)v;
        const uint32_t u = __float_as_uint((float)(v - (double)h));
        v32[9 * slot + m] = h;
        lo16[base + SELL_C * m] = (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
    }
}
void split_copy(FineMatrix &K, hipStream_t s) {
    if (!K.v32.p) K.v32.alloc(9 * K.nslots);
    if (!K.lo16.p) K.lo16.alloc(9 * K.nslots);
    hipLaunchKernelGGL(split_copy_kernel, dim3((unsigned)((K.nslots + 255) / 256)), dim3(256), 0, s, K.nslots,
                       K.v64.p, K.v32.p, K.lo16.p);
    PF_HIP(hipGetLastError());
}
void spmv_fine_split_dot(const FineMatrix &K, const double *x, double *y, double *partial, double *out,
                         hipStream_t s) {
    hipLaunchKernelGGL(sell_split_kernel<SELL_DOT>, dim3((unsigned)K.nslices), dim3(SELL_C), 0, s, K.nb,
                       K.slice_ptr.p, K.colind.p, K.v32.p, K.lo16.p, x, (const double *)nullptr, y, partial);
    hipLaunchKernelGGL(finish_sum_kernel, dim3(1), dim3(1024), 0, s, (int)K.nslices, partial, out);
    PF_HIP(hipGetLastError());
}
void spmv_fine_split(const FineMatrix &K, const double *x, double *y, hipStream_t s) {
    hipLaunchKernelGGL(sell_split_kernel<SELL_Y>, dim3((unsigned)K.nslices), dim3(SELL_C), 0, s, K.nb,
                       K.slice_ptr.p, K.colind.p, K.v32.p, K.lo16.p, x, (const double *)nullptr, y, nullptr);
    PF_HIP(hipGetLastError());
}
void half_scaled_copy(FineMatrix &K, hipStream_t s) {
    if (!K.sinv.p) K.sinv.alloc(3 * K.nb);
    if (!K.v16.p) K.v16.alloc(9 * K.nslots);
    hipLaunchKernelGGL(half_scale_diag_kernel, dim3(grid_for(3 * K.nb, 256)), dim3(256), 0, s, K.nb, K.diagpos.p,
                       K.v64.p, K.sinv.p);
    hipLaunchKernelGGL(half_scale_values_kernel, dim3((unsigned)K.nslices), dim3(SELL_C), 0, s, K.nb,
                       K.slice_ptr.p, K.colind.p, K.v64.p, K.sinv.p, reinterpret_cast<_Float16 *>(K.v16.p));
    PF_HIP(hipGetLastError());
}
void spmv_fine_half(const FineMatrix &K, const float *x, float *y, hipStream_t s) {
    hipLaunchKernelGGL((sell_half_kernel<SELL_Y>), dim3((unsigned)K.nslices), dim3(SELL_C), 0, s, K.nb,
                       K.slice_ptr.p, K.colind.p, reinterpret_cast<const _Float16 *>(K.v16.p), K.sinv.p, x,
                       (const float *)nullptr, y);
    PF_HIP(hipGetLastError());
}
void residual_fine_half(const FineMatrix &K, const float *x, const float *b, float *r, hipStream_t s) {
    hipLaunchKernelGGL((sell_half_kernel<SELL_RES>), dim3((unsigned)K.nslices), dim3(SELL_C), 0, s, K.nb,
                       K.slice_ptr.p, K.colind.p, reinterpret_cast<const _Float16 *>(K.v16.p), K.sinv.p, x, b, r);
    PF_HIP(hipGetLastError());
}
void axpy_safe_ratio(int64_t n, const double *num, const double *den, double sign, const double *x, double *y,
                     hipStream_t s) {
    hipLaunchKernelGGL(axpy_safe_ratio_kernel, dim3(grid_strided(n)), dim3(256), 0, s, n, num, den, sign, x, y);
    PF_HIP(hipGetLastError());
}
__global__ void axpy_f32_to_f64_kernel(int64_t n, double a, const float *__restrict__ x, double *__restrict__ y) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        y[i] += a * (double)x[i];
}
void axpy_f32_to_f64(int64_t n, double a, const float *x, double *y, hipStream_t s) {
    hipLaunchKernelGGL(axpy_f32_to_f64_kernel, dim3(grid_strided(n)), dim3(256), 0, s, n, a, x, y);
    PF_HIP(hipGetLastError());
}
void cg_update_mixed(int64_t n, const double *num, const double *den, const float *p, const float *q, double *x,
                     float *r, hipStream_t s) {
    hipLaunchKernelGGL(cg_update_mixed_kernel, dim3(grid_strided(n)), dim3(256), 0, s, n, num, den, p, q, x, r);
    PF_HIP(hipGetLastError());
}
template <class T>
void axpby(int64_t n, double a, const T *x, double b, T *y, hipStream_t s) {
    hipLaunchKernelGGL((axpby_kernel<T>), dim3(grid_strided(n)), dim3(256), 0, s, n, a, x, b, y);
    PF_HIP(hipGetLastError());
}
template <class T>
void axpby_negratio(int64_t n, double a, const T *x, const double *num, const double *den, T *y, hipStream_t s) {
    hipLaunchKernelGGL((axpby_negratio_kernel<T>), dim3(grid_strided(n)), dim3(256), 0, s, n, a, x, num, den, y);
    PF_HIP(hipGetLastError());
}
template <class S, class D>
void convert(int64_t n, const S *src, D *dst, hipStream_t s) {
    hipLaunchKernelGGL((convert_kernel<S, D>), dim3(grid_strided(n)), dim3(256), 0, s, n, src, dst);
    PF_HIP(hipGetLastError());
}
template <class T>
void fill(int64_t n, T v, T *d, hipStream_t s) {
    hipLaunchKernelGGL((fill_kernel<T>), dim3(grid_strided(n)), dim3(256), 0, s, n, v, d);
    PF_HIP(hipGetLastError());
}
#define PF_INST(T)                                                                                     \
    template void axpy_ratio<T>(int64_t, const double *, const double *, double, const T *, T *, hipStream_t); \
    template void xpby_ratio<T>(int64_t, const double *, const double *, const T *, T *, hipStream_t);         \
    template void cg_update<T>(int64_t, const double *, const double *, const T *, const T *, T *, T *,       \
                               hipStream_t);                                                              \
    template void axpby<T>(int64_t, double, const T *, double, T *, hipStream_t);                            \
    template void axpby_negratio<T>(int64_t, double, const T *, const double *, const double *, T *, hipStream_t); \
    template void fill<T>(int64_t, T, T *, hipStream_t);
PF_INST(double)
PF_INST(float)
#undef PF_INST
template void convert<double, float>(int64_t, const double *, float *, hipStream_t);
template void convert<float, double>(int64_t, const float *, double *, hipStream_t);

void sub_external(int64_t n, double lambda, const double *fext, double *R, hipStream_t s) {
    hipLaunchKernelGGL(resid_ext_kernel, dim3(grid_strided(n)), dim3(256), 0, s, n, lambda, fext, R);
    PF_HIP(hipGetLastError());
}

void impose_dirichlet_dev(FineMatrix &K, int64_t nbc, const int64_t *bcdofs, const uint8_t *flag,
                          const double *val, double lambda, const double *U, double *dg, int64_t nrows,
                          const int32_t *rows, double *R, int sel, hipStream_t s) {
    if (nbc == 0) return;
    hipLaunchKernelGGL(bc_dg_kernel, dim3(grid_for(nbc, 256)), dim3(256), 0, s, nbc, bcdofs, flag, val, lambda, U,
                       dg);
    hipLaunchKernelGGL(bc_rows_kernel, dim3(grid_for(nrows, 64)), dim3(64), 0, s, nrows, rows, K.slice_ptr.p,
                       K.rowlen.p, K.colind.p, K.v64.p, flag, dg, R, sel);
    hipLaunchKernelGGL(bc_rhs_kernel, dim3(grid_for(nbc, 256)), dim3(256), 0, s, nbc, bcdofs, flag, dg, R, sel);
    PF_HIP(hipGetLastError());
}

void block_diag_inverse3(const FineMatrix &K, double *dinv, hipStream_t s) {
    hipLaunchKernelGGL(diag_inv3_kernel, dim3(grid_for(K.nb, 256)), dim3(256), 0, s, K.nb, K.diagpos.p, K.v64.p,
                       dinv);
    PF_HIP(hipGetLastError());
}
template <class T>
void block_jacobi3(int64_t nb, const double *dinv, const T *r, T *z, hipStream_t s) {
    hipLaunchKernelGGL((block_jacobi3_kernel<T>), dim3(grid_for(nb, 256)), dim3(256), 0, s, nb, dinv, r, z);
    PF_HIP(hipGetLastError());
}
template void block_jacobi3<double>(int64_t, const double *, const double *, double *, hipStream_t);
template void block_jacobi3<float>(int64_t, const double *, const float *, float *, hipStream_t);

}  // namespace pfem
