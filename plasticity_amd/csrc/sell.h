// sell.h -- the fine-level SELL-64 BSR(3x3) product kernel, shared by sparse.hip and
// amg.hip (internal).
//
// One 64-lane workgroup per slice of 64 block rows.  For each slice column k the
// workgroup first copies the 64 contiguous 3x3 blocks (64 x 72 B in FP64, 64 x 36 B in
// FP32) into LDS with 16-byte non-temporal loads -- every global load instruction covers
// 1 KB of consecutive bytes -- then each lane takes its own block from LDS and gathers x.
// Padding slots carry zero values and a valid column, so lanes never branch on the
// row length.
#pragma once
#include "common.h"

namespace pfem {

enum SellMode { SELL_Y = 0, SELL_RES = 1, SELL_DOT = 2 };

// Matrix values and indices are read once per product: load them non-temporally so the
// stream does not evict the gathered vector from L2 (fine FP64 SpMV at 100^3:
// 0.385 -> 0.341 ms).
typedef unsigned int u32x4_t __attribute__((ext_vector_type(4)));
__device__ __forceinline__ void stream_copy16(void *dst, const void *src, int t) {
    reinterpret_cast<u32x4_t *>(dst)[t] = __builtin_nontemporal_load(reinterpret_cast<const u32x4_t *>(src) + t);
}

// Workgroups are dispatched round-robin over the 8 XCDs (each with its own L2): group b
// runs on XCD b % 8.  Remap so that XCD x works through one contiguous eighth of the
// slices; the x-vector neighbours of a slice (+-1, +-row, +-plane) then mostly live in
// the same L2.  Bijective on [0, G).
__device__ __forceinline__ int64_t xcd_contiguous(int64_t b, int64_t G) {
    const int64_t q = G >> 3, r = G & 7, x = b & 7, k = b >> 3;
    return (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + k;
}

// Row block i = s * SELL_C + lane of slice s: y = sum over the slice columns of
// block(k) x(col(k)), the blocks staged in `tile` (SELL_C * 9 values of LDS).
template <class VT, class XT>
__device__ __forceinline__ void sell_slice_product(int64_t s, int lane, const int64_t *__restrict__ slice_ptr,
                                                   const int32_t *__restrict__ colind, const VT *__restrict__ vals,
                                                   const XT *__restrict__ x, VT *tile, VT &y0, VT &y1, VT &y2) {
    constexpr int NV = SELL_C * 9 * (int)sizeof(VT) / 16;  // 16-byte chunks per slice column
    const int64_t base = slice_ptr[s];
    const int w = (int)((slice_ptr[s + 1] - base) / SELL_C);
    y0 = 0, y1 = 0, y2 = 0;
    for (int k = 0; k < w; ++k) {
        const int64_t col0 = base + (int64_t)k * SELL_C;
#pragma unroll
        for (int t = lane; t < NV; t += SELL_C) stream_copy16(tile, vals + 9 * col0, t);
        const int64_t j = __builtin_nontemporal_load(colind + col0 + lane);
        const XT x0 = x[3 * j], x1 = x[3 * j + 1], x2 = x[3 * j + 2];
        __syncthreads();
        const VT *v = tile + 9 * lane;
        y0 += v[0] * (VT)x0 + v[1] * (VT)x1 + v[2] * (VT)x2;
        y1 += v[3] * (VT)x0 + v[4] * (VT)x1 + v[5] * (VT)x2;
        y2 += v[6] * (VT)x0 + v[7] * (VT)x1 + v[8] * (VT)x2;
        __syncthreads();
    }
}

// One Chebyshev step of the fine-level smoother with its product in the same pass:
// t = K d for the lane's block row, then r -= t, d' = c1 d + c2 Dinv r, x += d' -- the
// arithmetic of the separate product + step launches, in the same order (bitwise-equal
// results), without writing and re-reading t.  d is the product's input vector, read by
// every workgroup, so d' goes to d_out; d_out == null (the smoother's last step, after
// which neither r nor d is read again) skips the r and d' stores.
template <class VT, class XT>
__global__ __launch_bounds__(64) void sell_cheb_kernel(int64_t nb, const int64_t *__restrict__ slice_ptr,
                                                       const int32_t *__restrict__ colind,
                                                       const VT *__restrict__ vals, const XT *__restrict__ d,
                                                       const VT *__restrict__ dinv, XT *__restrict__ r,
                                                       XT *__restrict__ d_out, XT *__restrict__ x, double c1,
                                                       double c2) {
    __shared__ __attribute__((aligned(16))) VT tile[SELL_C * 9];
    const int lane = threadIdx.x;
    const int64_t s = xcd_contiguous(blockIdx.x, gridDim.x);
    const int64_t i = s * SELL_C + lane;
    VT y[3];
    sell_slice_product<VT, XT>(s, lane, slice_ptr, colind, vals, d, tile, y[0], y[1], y[2]);
    if (i >= nb) return;
    VT rv[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        rv[c] = (VT)r[3 * i + c] - (VT)(XT)y[c];
        if (d_out) r[3 * i + c] = (XT)rv[c];
    }
    const VT *m = dinv + 9 * i;
#pragma unroll
    for (int rr = 0; rr < 3; ++rr) {
        VT sm = 0;
#pragma unroll
        for (int c = 0; c < 3; ++c) sm = fma(m[3 * rr + c], rv[c], sm);
        const int64_t o = 3 * i + rr;
        const double dn = fma(c1, (double)d[o], c2 * (double)sm);
        if (d_out) d_out[o] = (XT)dn;
        x[o] = (XT)((double)x[o] + dn);
    }
}

template <class VT, class XT, int MODE>
__global__ __launch_bounds__(64) void sell_lds_kernel(int64_t nb, const int64_t *__restrict__ slice_ptr,
                                                      const int32_t *__restrict__ colind,
                                                      const VT *__restrict__ vals, const XT *__restrict__ x,
                                                      const XT *__restrict__ b, XT *__restrict__ y,
                                                      double *__restrict__ partial) {
    __shared__ __attribute__((aligned(16))) VT tile[SELL_C * 9];
    const int lane = threadIdx.x;
    const int64_t s = xcd_contiguous(blockIdx.x, gridDim.x);
    const int64_t i = s * SELL_C + lane;
    VT y0, y1, y2;
    sell_slice_product<VT, XT>(s, lane, slice_ptr, colind, vals, x, tile, y0, y1, y2);
    double d = 0.0;
    if (i < nb) {
        if (MODE == SELL_RES) {
            y[3 * i] = (XT)((VT)b[3 * i] - y0);
            y[3 * i + 1] = (XT)((VT)b[3 * i + 1] - y1);
            y[3 * i + 2] = (XT)((VT)b[3 * i + 2] - y2);
        } else {
            y[3 * i] = (XT)y0;
            y[3 * i + 1] = (XT)y1;
            y[3 * i + 2] = (XT)y2;
            if (MODE == SELL_DOT)
                d = (double)x[3 * i] * (double)y0 + (double)x[3 * i + 1] * (double)y1 +
                    (double)x[3 * i + 2] * (double)y2;
        }
    }
    if (MODE == SELL_DOT) {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) d += __shfl_down(d, off, 64);
        if (lane == 0) partial[s] = d;
    }
}

// The FP64 product from a split copy of the matrix: the FP32 copy hi = fl32(K) (the
// one the FP32 V-cycle reads) plus lo = bf16(K - hi), which carries the next 8
// significant bits.  hi + lo is formed exactly in FP64 (at most 33 significant bits)
// and agrees with K to 2^-33 relative per entry; products and sums are FP64.  54 bytes
// of values per block instead of 72.  lo is stored per slice column as 9 planes of 64
// (plane m holds entry m of the column's 64 blocks), so lane i reads its 9 entries
// from LDS at i, 64 + i, ... without bank conflicts.
__device__ __forceinline__ double bf16_bits_to_double(uint16_t h) {
    return (double)__uint_as_float((uint32_t)h << 16);
}
template <int MODE>
__global__ __launch_bounds__(64) void sell_split_kernel(int64_t nb, const int64_t *__restrict__ slice_ptr,
                                                        const int32_t *__restrict__ colind,
                                                        const float *__restrict__ hi,
                                                        const uint16_t *__restrict__ lo,
                                                        const double *__restrict__ x, const double *__restrict__ b,
                                                        double *__restrict__ y, double *__restrict__ partial) {
    constexpr int NH = SELL_C * 9 * 4 / 16;  // 144 chunks of 16 B per slice column (hi)
    constexpr int NL = SELL_C * 9 * 2 / 16;  // 72 (lo)
    __shared__ __attribute__((aligned(16))) float thi[SELL_C * 9];
    __shared__ __attribute__((aligned(16))) uint16_t tlo[SELL_C * 9];
    const int lane = threadIdx.x;
    const int64_t s = xcd_contiguous(blockIdx.x, gridDim.x);
    const int64_t i = s * SELL_C + lane;
    const int64_t base = slice_ptr[s];
    const int w = (int)((slice_ptr[s + 1] - base) / SELL_C);
    double y0 = 0.0, y1 = 0.0, y2 = 0.0;
    for (int k = 0; k < w; ++k) {
        const int64_t col0 = base + (int64_t)k * SELL_C;
#pragma unroll
        for (int t = lane; t < NH; t += SELL_C) stream_copy16(thi, hi + 9 * col0, t);
        for (int t = lane; t < NL; t += SELL_C) stream_copy16(tlo, lo + 9 * col0, t);
        const int64_t j = __builtin_nontemporal_load(colind + col0 + lane);
        const double x0 = x[3 * j], x1 = x[3 * j + 1], x2 = x[3 * j + 2];
        __syncthreads();
        const float *v = thi + 9 * lane;
        double a[9];
#pragma unroll
        for (int m = 0; m < 9; ++m) a[m] = (double)v[m] + bf16_bits_to_double(tlo[SELL_C * m + lane]);
        y0 += a[0] * x0 + a[1] * x1 + a[2] * x2;
        y1 += a[3] * x0 + a[4] * x1 + a[5] * x2;
        y2 += a[6] * x0 + a[7] * x1 + a[8] * x2;
        __syncthreads();
    }
    double d = 0.0;
    if (i < nb) {
        if (MODE == SELL_RES) {
            y[3 * i] = b[3 * i] - y0;
            y[3 * i + 1] = b[3 * i + 1] - y1;
            y[3 * i + 2] = b[3 * i + 2] - y2;
        } else {
            y[3 * i] = y0;
            y[3 * i + 1] = y1;
            y[3 * i + 2] = y2;
            if (MODE == SELL_DOT) d = x[3 * i] * y0 + x[3 * i + 1] * y1 + x[3 * i + 2] * y2;
        }
    }
    if (MODE == SELL_DOT) {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) d += __shfl_down(d, off, 64);
        if (lane == 0) partial[s] = d;
    }
}

// The same product with an FP16 copy of the diagonally scaled matrix
// Khat = S K S, S = diag(1/sqrt|K_ii|) (unit diagonal, |Khat_ij| <~ 1, so FP16 holds it
// without overflow): y = S^-1 Khat S^-1 x, evaluated as sinv_i * sum_j Khat_ij (sinv_j x_j)
// with sinv = sqrt|K_ii| and FP32 accumulation.  Used for the fine-level products of
// the FP32 V-cycle; 22 B per block instead of 40 B (FP32) or 76 B (FP64).
template <int MODE>
__global__ __launch_bounds__(64) void sell_half_kernel(int64_t nb, const int64_t *__restrict__ slice_ptr,
                                                       const int32_t *__restrict__ colind,
                                                       const _Float16 *__restrict__ vals,
                                                       const float *__restrict__ sinv, const float *__restrict__ x,
                                                       const float *__restrict__ b, float *__restrict__ y) {
    constexpr int NV = SELL_C * 9 * 2 / 16;  // 72 chunks of 16 B per slice column
    __shared__ __attribute__((aligned(16))) _Float16 tile[SELL_C * 9];
    const int lane = threadIdx.x;
    const int64_t s = xcd_contiguous(blockIdx.x, gridDim.x);
    const int64_t i = s * SELL_C + lane;
    const int64_t base = slice_ptr[s];
    const int w = (int)((slice_ptr[s + 1] - base) / SELL_C);
    float y0 = 0.f, y1 = 0.f, y2 = 0.f;
    for (int k = 0; k < w; ++k) {
        const int64_t col0 = base + (int64_t)k * SELL_C;
        for (int t = lane; t < NV; t += SELL_C) stream_copy16(tile, vals + 9 * col0, t);
        const int64_t j = __builtin_nontemporal_load(colind + col0 + lane);
        const float x0 = x[3 * j] * sinv[3 * j], x1 = x[3 * j + 1] * sinv[3 * j + 1],
                    x2 = x[3 * j + 2] * sinv[3 * j + 2];
        __syncthreads();
        const _Float16 *v = tile + 9 * lane;
        y0 += (float)v[0] * x0 + (float)v[1] * x1 + (float)v[2] * x2;
        y1 += (float)v[3] * x0 + (float)v[4] * x1 + (float)v[5] * x2;
        y2 += (float)v[6] * x0 + (float)v[7] * x1 + (float)v[8] * x2;
        __syncthreads();
    }
    if (i < nb) {
        y0 *= sinv[3 * i];
        y1 *= sinv[3 * i + 1];
        y2 *= sinv[3 * i + 2];
        if (MODE == SELL_RES) {
            y[3 * i] = b[3 * i] - y0;
            y[3 * i + 1] = b[3 * i + 1] - y1;
            y[3 * i + 2] = b[3 * i + 2] - y2;
        } else {
            y[3 * i] = y0;
            y[3 * i + 1] = y1;
            y[3 * i + 2] = y2;
        }
    }
}

}  // namespace pfem
