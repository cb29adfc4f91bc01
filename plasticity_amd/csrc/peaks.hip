// peaks.hip -- achievable HBM copy bandwidth and FP64 FMA rate of the device, measured
// with a STREAM-style copy kernel and an FMA loop.  The roofline fractions the
// benchmark reports divide by the published MI355X peaks (8 TB/s, 78.6 TFLOP/s FP64);
// these two measurements say how much of each peak a plain kernel reaches.
#include <algorithm>

#include "common.h"

namespace pfem {
namespace {

typedef unsigned int u32x4_t __attribute__((ext_vector_type(4)));

constexpr int STREAM_U = 8;  // 16-byte words in flight per lane

// Workgroup b streams the contiguous words [b chunk, (b+1) chunk) of n (the last chunk
// is clipped), STREAM_U loads in flight per lane, lanes 16 bytes apart: every load
// instruction of a wave covers 1 KiB of consecutive addresses.
__device__ __forceinline__ void chunk_of(int64_t n, int64_t &lo, int64_t &hi) {
    const int64_t chunk = (n + gridDim.x - 1) / gridDim.x;
    lo = (int64_t)blockIdx.x * chunk;
    hi = lo + chunk < n ? lo + chunk : n;
}

// dst = src over n 16-byte words
__global__ void __launch_bounds__(256) copy16_kernel(const u32x4_t *__restrict__ src, u32x4_t *__restrict__ dst,
                                                     int64_t n) {
    int64_t lo, hi;
    chunk_of(n, lo, hi);
    int64_t i = lo + threadIdx.x;
    for (; i + (STREAM_U - 1) * 256 < hi; i += STREAM_U * 256) {
        u32x4_t v[STREAM_U];
#pragma unroll
        for (int k = 0; k < STREAM_U; ++k) v[k] = __builtin_nontemporal_load(src + i + k * 256);
#pragma unroll
        for (int k = 0; k < STREAM_U; ++k) dst[i + k * 256] = v[k];
    }
    for (; i < hi; i += 256) dst[i] = src[i];
}

// XOR of the n 16-byte words of src, one partial per lane (out: one word per thread):
// the read-only stream rate
__global__ void __launch_bounds__(256) read16_kernel(const u32x4_t *__restrict__ src, int64_t n,
                                                     unsigned *__restrict__ out) {
    int64_t lo, hi;
    chunk_of(n, lo, hi);
    int64_t i = lo + threadIdx.x;
    u32x4_t acc = {0u, 0u, 0u, 0u};
    for (; i + (STREAM_U - 1) * 256 < hi; i += STREAM_U * 256) {
        u32x4_t v[STREAM_U];
#pragma unroll
        for (int k = 0; k < STREAM_U; ++k) v[k] = __builtin_nontemporal_load(src + i + k * 256);
#pragma unroll
        for (int k = 0; k < STREAM_U; ++k) acc ^= v[k];
    }
    for (; i < hi; i += 256) acc ^= src[i];
    out[(int64_t)blockIdx.x * blockDim.x + threadIdx.x] = acc.x ^ acc.y ^ acc.z ^ acc.w;
}

constexpr int FMA_CHAINS = 8;

// FMA_CHAINS independent chains x <- x a + b per lane, `iters` times; the sum is
// stored so that no chain is dead code (out holds one double per thread).
__global__ void __launch_bounds__(256) fma_loop_kernel(double *out, int iters, double a, double b) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double x[FMA_CHAINS];
#pragma unroll
    for (int k = 0; k < FMA_CHAINS; ++k) x[k] = 1e-3 * (double)(threadIdx.x + k);
#pragma unroll 4
    for (int i = 0; i < iters; ++i) {
#pragma unroll
        for (int k = 0; k < FMA_CHAINS; ++k) x[k] = __builtin_fma(x[k], a, b);
    }
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < FMA_CHAINS; ++k) s += x[k];
    out[t] = s;
}

}  // namespace
}  // namespace pfem

using namespace pfem;

extern "C" int pfem_device_peaks(int64_t copy_bytes, int reps, double *copy_gbs, double *read_gbs,
                                 double *fp64_tflops) {
    if (!copy_gbs || !fp64_tflops || reps < 1 || copy_bytes < (1 << 20)) return PFEM_ERR_BAD_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return PFEM_ERR_NO_DEVICE;
    try {
        hipStream_t s;
        PF_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        struct SG { hipStream_t s; ~SG() { (void)hipStreamDestroy(s); } } sg{s};
        hipEvent_t e0, e1;
        PF_HIP(hipEventCreate(&e0));
        PF_HIP(hipEventCreate(&e1));
        struct EG { hipEvent_t a, b; ~EG() { (void)hipEventDestroy(a); (void)hipEventDestroy(b); } } eg{e0, e1};
        int dev = 0, ncu = 0;
        PF_HIP(hipGetDevice(&dev));
        PF_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));

        // copy and read-only stream: the best of `reps` launches each
        const int64_t n16 = copy_bytes / 16;
        dbuf<u32x4_t> src(n16), dst(n16);
        PF_HIP(hipMemsetAsync(src.p, 1, n16 * 16, s));
        const unsigned sgrid = (unsigned)std::min<int64_t>((n16 + 255) / 256, (int64_t)ncu * 16);
        dbuf<unsigned> part((size_t)sgrid * 256);
        double best_copy = 1e30, best_read = 1e30;
        for (int r = 0; r <= reps; ++r) {  // launch 0 of each warms up
            float ms = 0.f;
            PF_HIP(hipEventRecord(e0, s));
            hipLaunchKernelGGL(copy16_kernel, dim3(sgrid), dim3(256), 0, s, src.p, dst.p, n16);
            PF_HIP(hipEventRecord(e1, s));
            PF_HIP(hipEventSynchronize(e1));
            PF_HIP(hipEventElapsedTime(&ms, e0, e1));
            if (r > 0) best_copy = std::min(best_copy, (double)ms);
            PF_HIP(hipEventRecord(e0, s));
            hipLaunchKernelGGL(read16_kernel, dim3(sgrid), dim3(256), 0, s, src.p, n16, part.p);
            PF_HIP(hipEventRecord(e1, s));
            PF_HIP(hipEventSynchronize(e1));
            PF_HIP(hipEventElapsedTime(&ms, e0, e1));
            if (r > 0) best_read = std::min(best_read, (double)ms);
        }
        *copy_gbs = 2.0 * 16.0 * (double)n16 / (best_copy * 1e-3) / 1e9;
        if (read_gbs) *read_gbs = 16.0 * (double)n16 / (best_read * 1e-3) / 1e9;
        double best = 1e30;

        // FP64 FMA loop: 16 blocks of 256 lanes per CU, 2 flops per FMA
        const unsigned fgrid = (unsigned)ncu * 16;
        const int iters = 4096;
        dbuf<double> out((size_t)fgrid * 256);
        for (int r = 0; r <= reps; ++r) {
            PF_HIP(hipEventRecord(e0, s));
            hipLaunchKernelGGL(fma_loop_kernel, dim3(fgrid), dim3(256), 0, s, out.p, iters, 0.999999, 1e-7);
            PF_HIP(hipEventRecord(e1, s));
            PF_HIP(hipEventSynchronize(e1));
            float ms = 0.f;
            PF_HIP(hipEventElapsedTime(&ms, e0, e1));
            if (r > 0) best = std::min(best, (double)ms);
        }
        PF_HIP(hipGetLastError());
        *fp64_tflops = 2.0 * FMA_CHAINS * (double)iters * (double)fgrid * 256.0 / (best * 1e-3) / 1e12;
        return PFEM_OK;
    } catch (const hip_failure &) {
        return PFEM_ERR_HIP;
    }
}
