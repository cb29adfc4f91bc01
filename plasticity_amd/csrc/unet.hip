// unet.hip -- kernels of the learned preconditioner: a bias-free 3D convolutional U-net
// over the structured node grid, on BF16 MFMA with FP32 accumulation.
//
// Tensors are channels-last over a z-major voxel grid: t[((z*ny + y)*nx + x)*C + c].
// Activations are BF16 (stored as uint16), gradients FP32.  Every convolution, forward
// or backward, is an implicit GEMM on BF16 MFMA with FP32 accumulation:
//
//   conv3_tile_kernel<CIN, COUT, MODE>  the 3x3x3 layers (C3, UPC3; forward and
//                             backward data): one workgroup = an output tile of
//                             16 x 4 x 4 voxels whose 18 x 6 x 6 input neighbourhood
//                             (for UPC3 the nearest-neighbour upsample of the coarse
//                             input) is staged in LDS once and read by all 27 taps;
//                             each wave computes one z-plane of the tile, 4 M-tiles of
//                             16 voxels x COUT/16 N-tiles on v_mfma_f32_16x16x32_bf16,
//                             K = 27 taps x CIN in steps of 32, so every B fragment
//                             (weights, L2-resident) feeds 4 MFMAs and every A fragment
//                             (one 16-byte LDS read) COUT/16.
//   conv_kernel<COUT, MODE>   one wave = 16 consecutive voxels of the enumeration grid
//                             x all COUT output channels (COUT/16 accumulator tiles);
//                             the K loop runs over taps x input channels in steps of
//                             16; the A fragment (4 channels of one voxel per lane) is
//                             an 8-byte (BF16) or 16-byte (FP32) load, the B fragment
//                             comes from weights pre-packed in MFMA lane order.
//   wgrad_kernel<MODE>        dW[t][ci][co] = sum_v X(u(v,t))[ci] g(v)[co]: one block per
//                             (voxel chunk, tap, 16x16 channel tile), four waves over
//                             disjoint voxel subsets, per-wave partials reduced in a
//                             fixed order by wreduce_kernel (deterministic).
//
// MFMA 16x16x16 lane layout (wave64): A[i][k] in lane i + 16*(k/4), item k%4;
// B[k][j] in lane j + 16*(k/4), item k%4; D[i][j] in lane j + 16*(i/4), item i%4.
// MFMA 16x16x32 (gfx950): A[i][k] in lane i + 16*(k/8), item k%8; B[k][j] in lane
// j + 16*(k/8), item k%8; D as above (checked on the device with integer operands).
//
// Modes (input voxel u for enumeration voxel v and tap t):
//   C3    3x3x3, stride 1, zero padding            u = v + d(t), same grid
//   C1    1x1x1                                    u = v
//   C2S2  2x2x2, stride 2 (down)                   u = 2v + o(t) on the finer input grid
//   UPC3  3x3x3 on the nearest-neighbour upsample of a coarser input:
//         w = v + d(t) on the output grid, u = w >> 1 on the coarse input grid
//   TS2   transpose of C2S2 for one parity o (blockIdx.y): enumerates the coarse grid,
//         u = v, output voxel 2v + o on the finer grid
#include <hip/hip_runtime.h>

#include "unet.h"

namespace pfem {
namespace {

typedef short s4_t __attribute__((ext_vector_type(4)));
typedef float f4_t __attribute__((ext_vector_type(4)));
typedef __bf16 bf8_t __attribute__((ext_vector_type(8)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t f2bf(float f) {  // round to nearest even
    uint32_t b = __float_as_uint(f);
    b += 0x7fffu + ((b >> 16) & 1u);
    return (uint16_t)(b >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ __forceinline__ s4_t load_a4(const uint16_t *p) {
    const ushort4 v = *reinterpret_cast<const ushort4 *>(p);
    s4_t r;
    r[0] = (short)v.x; r[1] = (short)v.y; r[2] = (short)v.z; r[3] = (short)v.w;
    return r;
}
__device__ __forceinline__ s4_t load_a4(const float *p) {
    const float4 v = *reinterpret_cast<const float4 *>(p);
    s4_t r;
    r[0] = (short)f2bf(v.x); r[1] = (short)f2bf(v.y); r[2] = (short)f2bf(v.z); r[3] = (short)f2bf(v.w);
    return r;
}
__device__ __forceinline__ short load_bf(const uint16_t *p) { return (short)*p; }
__device__ __forceinline__ short load_bf(const float *p) { return (short)f2bf(*p); }
__device__ __forceinline__ void store_out(uint16_t *p, float v) { *p = f2bf(v); }
__device__ __forceinline__ void store_out(float *p, float v) { *p = v; }

__device__ __forceinline__ int mode_taps(int mode) {
    return (mode == NET_C3 || mode == NET_UPC3) ? 27 : (mode == NET_C2S2 ? 8 : 1);
}

// Input voxel of enumeration voxel (x,y,z) for tap t, or -1 (zero padding).
template <int MODE>
__device__ __forceinline__ int64_t input_voxel(int x, int y, int z, int t, const NetDims &din, const NetDims &dout) {
    if (MODE == NET_C1 || MODE == NET_TS2) return ((int64_t)z * din.ny + y) * din.nx + x;
    if (MODE == NET_C2S2) {
        const int ux = 2 * x + (t & 1), uy = 2 * y + ((t >> 1) & 1), uz = 2 * z + (t >> 2);
        if (ux >= din.nx || uy >= din.ny || uz >= din.nz) return -1;
        return ((int64_t)uz * din.ny + uy) * din.nx + ux;
    }
    const int wx = x + t % 3 - 1, wy = y + (t / 3) % 3 - 1, wz = z + t / 9 - 1;
    if (MODE == NET_C3) {
        if (wx < 0 || wy < 0 || wz < 0 || wx >= din.nx || wy >= din.ny || wz >= din.nz) return -1;
        return ((int64_t)wz * din.ny + wy) * din.nx + wx;
    }
    // UPC3: w on the output grid, then the coarse parent
    if (wx < 0 || wy < 0 || wz < 0 || wx >= dout.nx || wy >= dout.ny || wz >= dout.nz) return -1;
    return ((int64_t)(wz >> 1) * din.ny + (wy >> 1)) * din.nx + (wx >> 1);
}

template <int COUT, int MODE, class TI, class TO>
__global__ __launch_bounds__(256) void conv_kernel(ConvArgs a) {
    constexpr int NT = COUT / 16;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const NetDims de = (MODE == NET_TS2) ? a.din : a.dout;  // enumeration grid
    const int64_t ne = (int64_t)de.nx * de.ny * de.nz;
    const int64_t e0 = ((int64_t)blockIdx.x * 4 + wave) * 16;
    if (e0 >= ne) return;  // uniform per wave
    const int m = lane & 15, kq = lane >> 4;
    const int64_t e = e0 + m;
    const bool eok = e < ne;
    int x = 0, y = 0, z = 0;
    if (eok) {
        x = (int)(e % de.nx);
        const int64_t r = e / de.nx;
        y = (int)(r % de.ny);
        z = (int)(r / de.ny);
    }
    const int par = blockIdx.y;
    const int KC = a.cin / 16;
    const int ntaps = mode_taps(MODE);
    const TI *in = static_cast<const TI *>(a.in);
    const s4_t *wpk = reinterpret_cast<const s4_t *>(a.w);
    f4_t acc[NT];
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) acc[nt] = f4_t{0.f, 0.f, 0.f, 0.f};
    for (int t = 0; t < ntaps; ++t) {
        const int64_t u = eok ? input_voxel<MODE>(x, y, z, t, a.din, a.dout) : -1;
        const TI *src = in + (u >= 0 ? u : 0) * a.cin + kq * 4;
        const s4_t *wt = wpk + (int64_t)((MODE == NET_TS2 ? par : t) * KC) * NT * 64 + lane;
        for (int kc = 0; kc < KC; ++kc) {
            s4_t af = {0, 0, 0, 0};
            if (u >= 0) af = load_a4(src + kc * 16);
#pragma unroll
            for (int nt = 0; nt < NT; ++nt)
                acc[nt] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(af, wt[(kc * NT + nt) * 64], acc[nt], 0, 0, 0);
        }
    }
    const uint16_t *skip = static_cast<const uint16_t *>(a.skip);
    TO *out = static_cast<TO *>(a.out);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int64_t ee = e0 + 4 * kq + j;
        if (ee >= ne) continue;
        int64_t o = ee;
        if (MODE == NET_TS2) {
            const int cx = (int)(ee % de.nx);
            const int64_t r = ee / de.nx;
            const int cy = (int)(r % de.ny), cz = (int)(r / de.ny);
            const int fx = 2 * cx + (par & 1), fy = 2 * cy + ((par >> 1) & 1), fz = 2 * cz + (par >> 2);
            if (fx >= a.dout.nx || fy >= a.dout.ny || fz >= a.dout.nz) continue;
            o = ((int64_t)fz * a.dout.ny + fy) * a.dout.nx + fx;
        }
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) {
            float v = acc[nt][j];
            const int64_t idx = o * COUT + nt * 16 + m;
            if (skip) v += bf2f(skip[idx]);
            if (a.relu) v = fmaxf(v, 0.f);
            store_out(out + idx, v);
        }
    }
}


// ---------------------------------------------------------------- LDS-tiled 3x3x3
constexpr int T3X = 16, T3Y = 4, T3Z = 4;                      // output tile
constexpr int H3X = T3X + 2, H3Y = T3Y + 2, H3Z = T3Z + 2;     // staged neighbourhood
constexpr int H3VOX = H3X * H3Y * H3Z;

// 8 consecutive channels of one voxel as 8 BF16 (16 bytes)
__device__ __forceinline__ u32x4 load_bf8(const uint16_t *p) { return *reinterpret_cast<const u32x4 *>(p); }
__device__ __forceinline__ u32x4 load_bf8(const float *p) {
    const float4 lo = *reinterpret_cast<const float4 *>(p), hi = *reinterpret_cast<const float4 *>(p + 4);
    u32x4 r;
    r[0] = (uint32_t)f2bf(lo.x) | ((uint32_t)f2bf(lo.y) << 16);
    r[1] = (uint32_t)f2bf(lo.z) | ((uint32_t)f2bf(lo.w) << 16);
    r[2] = (uint32_t)f2bf(hi.x) | ((uint32_t)f2bf(hi.y) << 16);
    r[3] = (uint32_t)f2bf(hi.z) | ((uint32_t)f2bf(hi.w) << 16);
    return r;
}

// Output tile (blockIdx.x*16.., blockIdx.y*4.., blockIdx.z*4..) of a 3x3x3 convolution
// on the output grid a.dout.  Tile voxel (lx, ly, lz) of the LDS neighbourhood is grid
// voxel w = origin + (lx, ly, lz) - 1: for C3 the input at w, for UPC3 the coarse input
// at w >> 1; zero where w lies outside the output grid (the zero padding of the
// upsampled tensor).  The voxel stride is CIN + 8 BF16 (16 bytes of padding), so the
// 16 lanes reading 16 consecutive voxels with 16-byte LDS loads touch disjoint banks.
template <int CIN, int COUT, int MODE, class TI, class TO>
__global__ __launch_bounds__(256) void conv3_tile_kernel(ConvArgs a) {
    constexpr int NT = COUT / 16;
    constexpr int VS = CIN + 8;
    constexpr int NS = (27 * CIN + 31) / 32;  // K steps
    constexpr int CPV = CIN / 8;              // 16-byte chunks per voxel
    __shared__ __attribute__((aligned(16))) uint16_t tile[H3VOX * VS];
    const int x0 = blockIdx.x * T3X, y0 = blockIdx.y * T3Y, z0 = blockIdx.z * T3Z;
    const NetDims dout = a.dout, din = a.din;
    const TI *in = static_cast<const TI *>(a.in);
    for (int q = threadIdx.x; q < H3VOX * CPV; q += 256) {
        const int v = q / CPV, part = q - v * CPV;
        const int lx = v % H3X, ly = (v / H3X) % H3Y, lz = v / (H3X * H3Y);
        const int wx = x0 + lx - 1, wy = y0 + ly - 1, wz = z0 + lz - 1;
        u32x4 val = {0u, 0u, 0u, 0u};
        if (wx >= 0 && wy >= 0 && wz >= 0 && wx < dout.nx && wy < dout.ny && wz < dout.nz) {
            const int64_t u = MODE == NET_UPC3 ? ((int64_t)(wz >> 1) * din.ny + (wy >> 1)) * din.nx + (wx >> 1)
                                               : ((int64_t)wz * din.ny + wy) * din.nx + wx;
            val = load_bf8(in + u * CIN + part * 8);
        }
        *reinterpret_cast<u32x4 *>(&tile[v * VS + part * 8]) = val;
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int m = lane & 15, g = lane >> 4;
    const u32x4 *wp = reinterpret_cast<const u32x4 *>(a.w32) + lane;
    f4_t acc[T3Y][NT];
#pragma unroll
    for (int my = 0; my < T3Y; ++my)
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) acc[my][nt] = f4_t{0.f, 0.f, 0.f, 0.f};
    for (int st = 0; st < NS; ++st) {
        const int k = st * 32 + g * 8;
        const int t = k / CIN, c0 = k % CIN;  // this lane's tap and first channel
        u32x4 bw[NT];
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) bw[nt] = wp[(st * NT + nt) * 64];
        const int tt = t < 27 ? t : 0;
        const int dx = tt % 3, dy = (tt / 3) % 3, dz = tt / 9;
        const int base = ((wave + dz) * H3Y + dy) * H3X + dx + m;
#pragma unroll
        for (int my = 0; my < T3Y; ++my) {
            u32x4 av = *reinterpret_cast<const u32x4 *>(&tile[(base + my * H3X) * VS + c0]);
            if (t >= 27) av = u32x4{0u, 0u, 0u, 0u};  // K padding of the last step
            const bf8_t af = __builtin_bit_cast(bf8_t, av);
#pragma unroll
            for (int nt = 0; nt < NT; ++nt)
                acc[my][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af, __builtin_bit_cast(bf8_t, bw[nt]),
                                                                      acc[my][nt], 0, 0, 0);
        }
    }
    const uint16_t *skip = static_cast<const uint16_t *>(a.skip);
    TO *out = static_cast<TO *>(a.out);
    const int z = z0 + wave;
    if (z >= dout.nz) return;
#pragma unroll
    for (int my = 0; my < T3Y; ++my) {
        const int y = y0 + my;
        if (y >= dout.ny) continue;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int x = x0 + 4 * g + j;
            if (x >= dout.nx) continue;
            const int64_t o = ((int64_t)z * dout.ny + y) * dout.nx + x;
#pragma unroll
            for (int nt = 0; nt < NT; ++nt) {
                float v = acc[my][nt][j];
                const int64_t idx = o * COUT + nt * 16 + m;
                if (skip) v += bf2f(skip[idx]);
                if (a.relu) v = fmaxf(v, 0.f);
                store_out(out + idx, v);
            }
        }
    }
}

// W[t][ci][co] (FP32, the layer's own cin x cout) -> B fragments of the 16x16x32 MFMA for
// the GEMM K = (tap, ci') (27 taps x cinp, 32 per step, zero beyond), N = co':
//   pk[((s NT + nt) 64 + l) 8 + q] = W'(k = 32 s + 8 (l / 16) + q, n = 16 nt + l % 16),
// transpose: ci' = co, co' = ci (backward data); flip: tap t reads tap 26 - t.
__global__ void pack32_kernel(int cin, int cout, const float *__restrict__ W, uint16_t *__restrict__ pk,
                              int transpose, int flip) {
    const int cinp = transpose ? cout : cin, coutp = transpose ? cin : cout;
    const int NT = coutp / 16, K = 27 * cinp, NS = (K + 31) / 32;
    const int64_t total = (int64_t)NS * NT * 512;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < total;
         q += (int64_t)gridDim.x * blockDim.x) {
        const int qi = (int)(q & 7), l = (int)((q >> 3) & 63);
        const int64_t r = q >> 9;
        const int nt = (int)(r % NT), st = (int)(r / NT);
        const int k = 32 * st + 8 * (l >> 4) + qi, n = 16 * nt + (l & 15);
        float v = 0.f;
        if (k < K) {
            const int t = k / cinp, kp = k % cinp;
            const int ts = flip ? 26 - t : t;
            const int ci = transpose ? n : kp, co = transpose ? kp : n;
            v = W[((int64_t)ts * cin + ci) * cout + co];
        }
        pk[q] = f2bf(v);
    }
}

template <int MODE, class TI, class TG>
__global__ __launch_bounds__(256) void wgrad_kernel(WgradArgs a) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int t = blockIdx.y;
    const int NTo = a.cout / 16;
    const int cit = blockIdx.z / NTo, cot = blockIdx.z % NTo;
    const NetDims &de = a.dout;
    const int64_t ne = (int64_t)de.nx * de.ny * de.nz;
    const int64_t b0 = (int64_t)blockIdx.x * a.chunk;
    const int64_t b1 = b0 + a.chunk < ne ? b0 + a.chunk : ne;
    const int i = lane & 15, kq = lane >> 4;
    const TI *X = static_cast<const TI *>(a.x);
    const TG *G = static_cast<const TG *>(a.g);
    f4_t acc = {0.f, 0.f, 0.f, 0.f};
    for (int64_t v0 = b0 + wave * 16; v0 < b1; v0 += 64) {
        s4_t af = {0, 0, 0, 0}, bfr = {0, 0, 0, 0};
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int64_t v = v0 + 4 * kq + j;
            if (v < b1) {
                const int x = (int)(v % de.nx);
                const int64_t r = v / de.nx;
                const int y = (int)(r % de.ny), z = (int)(r / de.ny);
                const int64_t u = input_voxel<MODE>(x, y, z, t, a.din, a.dout);
                if (u >= 0) af[j] = load_bf(X + u * a.cin + cit * 16 + i);
                bfr[j] = load_bf(G + v * a.cout + cot * 16 + i);
            }
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(af, bfr, acc, 0, 0, 0);
    }
    const int ntaps = mode_taps(MODE);
    float *P = a.partial + ((((int64_t)blockIdx.x * 4 + wave) * ntaps + t) * a.cin + cit * 16) * a.cout + cot * 16;
#pragma unroll
    for (int j = 0; j < 4; ++j) P[(4 * kq + j) * a.cout + i] = acc[j];
}

// dW[k] (+)= sum over nparts partial slabs, in order.
__global__ void wreduce_kernel(int64_t nw, int nparts, const float *__restrict__ partial, float *__restrict__ dw,
                               int accumulate, float scale) {
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < nw; k += (int64_t)gridDim.x * blockDim.x) {
        float s = 0.f;
        for (int p = 0; p < nparts; ++p) s += partial[(int64_t)p * nw + k];
        dw[k] = accumulate ? dw[k] + scale * s : scale * s;
    }
}

// Packs master weights W[t][ci][co] (FP32) into MFMA B-fragment order for a layer
// whose GEMM has K = ci' (16-chunks), N = co' (16-tiles).  transpose: ci' = co, co' = ci
// (backward data); flip: tap t' reads tap ntaps-1-t (point reflection of a 3x3x3 stencil).
__global__ void pack_kernel(int ntaps, int cin, int cout, const float *__restrict__ W, uint16_t *__restrict__ pk,
                            int transpose, int flip) {
    const int cinp = transpose ? cout : cin, coutp = transpose ? cin : cout;
    const int KC = cinp / 16, NT = coutp / 16;
    const int64_t total = (int64_t)ntaps * KC * NT * 64 * 4;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < total;
         q += (int64_t)gridDim.x * blockDim.x) {
        const int j = (int)(q & 3), l = (int)((q >> 2) & 63);
        int64_t r = q >> 8;
        const int nt = (int)(r % NT);
        r /= NT;
        const int kc = (int)(r % KC);
        const int t = (int)(r / KC);
        const int kp = kc * 16 + 4 * (l >> 4) + j;  // K index (ci')
        const int np = nt * 16 + (l & 15);          // N index (co')
        const int ts = flip ? ntaps - 1 - t : t;
        const int ci = transpose ? np : kp, co = transpose ? kp : np;
        pk[q] = f2bf(W[((int64_t)ts * cin + ci) * cout + co]);
    }
}

// d = (d + extra) * (y > 0), the ReLU backward with an optional skip-path gradient.
__global__ void relu_back_kernel(int64_t n, float *__restrict__ d, const float *__restrict__ extra,
                                 const uint16_t *__restrict__ y) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        float v = d[i];
        if (extra) v += extra[i];
        d[i] = (y[i] != 0 && !(y[i] & 0x8000u)) ? v : 0.f;
    }
}

// Transpose of the nearest-neighbour upsample: coarse[c] = sum of the (up to 8) fine
// children 2c + o inside the fine grid.
__global__ void pool_sum_kernel(NetDims fine, NetDims coarse, int C, const float *__restrict__ df,
                                float *__restrict__ dc) {
    const int64_t n = (int64_t)coarse.nx * coarse.ny * coarse.nz * C;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
        const int c = (int)(q % C);
        int64_t v = q / C;
        const int x = (int)(v % coarse.nx);
        v /= coarse.nx;
        const int y = (int)(v % coarse.ny), z = (int)(v / coarse.ny);
        float s = 0.f;
        for (int o = 0; o < 8; ++o) {
            const int fx = 2 * x + (o & 1), fy = 2 * y + ((o >> 1) & 1), fz = 2 * z + (o >> 2);
            if (fx < fine.nx && fy < fine.ny && fz < fine.nz)
                s += df[(((int64_t)fz * fine.ny + fy) * fine.nx + fx) * C + c];
        }
        dc[q] = s;
    }
}

// 3-component nodal vector (FP64, constrained DOFs zeroed through `mask`) scaled by
// `scale` (read from device memory) -> 16-channel BF16 input.
__global__ void input_from_dofs_kernel(int64_t nvox, const double *__restrict__ r, const uint8_t *__restrict__ mask,
                                       const double *__restrict__ scale, uint16_t *__restrict__ x16) {
    const double s = *scale > 0 ? 1.0 / *scale : 0.0;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < nvox * 16;
         q += (int64_t)gridDim.x * blockDim.x) {
        const int c = (int)(q & 15);
        const int64_t v = q >> 4;
        float val = 0.f;
        if (c < 3 && !(mask && mask[3 * v + c])) val = (float)(r[3 * v + c] * s);
        x16[q] = f2bf(val);
    }
}
// BF16 3-channel sample -> 16-channel BF16 input.
__global__ void input_from_pairs_kernel(int64_t nvox, const uint16_t *__restrict__ p3, uint16_t *__restrict__ x16) {
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < nvox * 16;
         q += (int64_t)gridDim.x * blockDim.x) {
        const int c = (int)(q & 15);
        const int64_t v = q >> 4;
        x16[q] = c < 3 ? p3[3 * v + c] : (uint16_t)0;
    }
}
// 16-channel FP32 output -> nodal FP64 vector; constrained DOFs get r (identity rows),
// or 0 when r is null.
__global__ void output_to_dofs_kernel(int64_t nvox, const float *__restrict__ o16, const uint8_t *__restrict__ mask,
                                      const double *__restrict__ r, double *__restrict__ z) {
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < nvox * 3;
         q += (int64_t)gridDim.x * blockDim.x) {
        const int64_t v = q / 3;
        const int c = (int)(q % 3);
        z[q] = (mask && mask[q]) ? (r ? r[q] : 0.0) : (double)o16[v * 16 + c];
    }
}
// Nodal FP64 vector (constrained DOFs zeroed) -> BF16 3-channel sample scaled by 1/scale.
__global__ void dofs_to_pair_kernel(int64_t n, const double *__restrict__ v, const uint8_t *__restrict__ mask,
                                    const double *__restrict__ scale, uint16_t *__restrict__ p3) {
    const double s = *scale > 0 ? 1.0 / *scale : 0.0;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x)
        p3[q] = f2bf((mask && mask[q]) ? 0.f : (float)(v[q] * s));
}

// max |v| over free entries: per-block maxima, then one block.
__global__ __launch_bounds__(256) void absmax_partial_kernel(int64_t n, const double *__restrict__ v,
                                                             const uint8_t *__restrict__ mask,
                                                             double *__restrict__ partial) {
    __shared__ double sh[256];
    double m = 0.0;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x)
        if (!(mask && mask[q])) m = fmax(m, fabs(v[q]));
    sh[threadIdx.x] = m;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}
__global__ __launch_bounds__(256) void max_finish_kernel(int n, const double *__restrict__ partial,
                                                         double *__restrict__ out) {
    __shared__ double sh[256];
    double m = 0.0;
    for (int q = threadIdx.x; q < n; q += 256) m = fmax(m, partial[q]);
    sh[threadIdx.x] = m;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = sh[0];
}

// Loss 1 - cos(o, z) over the 3 used channels: per-block partial sums of (o.z, o.o, z.z).
__global__ __launch_bounds__(256) void cos_partial_kernel(int64_t nvox, const float *__restrict__ o16,
                                                          const uint16_t *__restrict__ z3, double *__restrict__ partial) {
    __shared__ double sh[3][256];
    double oz = 0, oo = 0, zz = 0;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < nvox * 3;
         q += (int64_t)gridDim.x * blockDim.x) {
        const double o = o16[(q / 3) * 16 + q % 3], z = bf2f(z3[q]);
        oz += o * z;
        oo += o * o;
        zz += z * z;
    }
    sh[0][threadIdx.x] = oz;
    sh[1][threadIdx.x] = oo;
    sh[2][threadIdx.x] = zz;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
            for (int k = 0; k < 3; ++k) sh[k][threadIdx.x] += sh[k][threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int k = 0; k < 3; ++k) partial[3 * blockIdx.x + k] = sh[k][0];
}
__global__ __launch_bounds__(256) void cos_finish_kernel(int nb, const double *__restrict__ partial,
                                                         double *__restrict__ sums, double *__restrict__ loss_acc) {
    __shared__ double sh[3][256];
    double a[3] = {0, 0, 0};
    for (int q = threadIdx.x; q < nb; q += 256)
        for (int k = 0; k < 3; ++k) a[k] += partial[3 * q + k];
    for (int k = 0; k < 3; ++k) sh[k][threadIdx.x] = a[k];
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
            for (int k = 0; k < 3; ++k) sh[k][threadIdx.x] += sh[k][threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        for (int k = 0; k < 3; ++k) sums[k] = sh[k][0];
        const double den = sqrt(sh[1][0] * sh[2][0]);
        if (loss_acc) *loss_acc += den > 0 ? 1.0 - sh[0][0] / den : 1.0;
    }
}
// dL/do for L = 1 - o.z / (|o||z|), channels >= 3 get 0.
__global__ void cos_grad_kernel(int64_t nvox, const float *__restrict__ o16, const uint16_t *__restrict__ z3,
                                const double *__restrict__ sums, float *__restrict__ d16) {
    const double oz = sums[0], oo = sums[1], zz = sums[2];
    const double no = sqrt(oo), nz = sqrt(zz);
    const double a = (no > 0 && nz > 0) ? -1.0 / (no * nz) : 0.0;
    const double b = (no > 0 && nz > 0) ? oz / (no * no * no * nz) : 0.0;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < nvox * 16;
         q += (int64_t)gridDim.x * blockDim.x) {
        const int c = (int)(q & 15);
        const int64_t v = q >> 4;
        d16[q] = c < 3 ? (float)(a * bf2f(z3[3 * v + c]) + b * o16[q]) : 0.f;
    }
}


// z += scale * o on the free DOFs (constrained DOFs untouched): the network's correction
// of a V-cycle output, rescaled by the factor its input was divided by.
__global__ void output_add_scaled_kernel(int64_t nvox, const float *__restrict__ o16, const uint8_t *__restrict__ mask,
                                         const double *__restrict__ scale, double *__restrict__ z) {
    const double sc = *scale;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < nvox * 3;
         q += (int64_t)gridDim.x * blockDim.x) {
        if (mask && mask[q]) continue;
        z[q] += sc * (double)o16[(q / 3) * 16 + q % 3];
    }
}
// (a - b) / scale -> BF16 3-channel sample (constrained DOFs 0)
__global__ void diff_to_pair_kernel(int64_t n, const double *__restrict__ a, const double *__restrict__ b,
                                    const uint8_t *__restrict__ mask, const double *__restrict__ scale,
                                    uint16_t *__restrict__ p3) {
    const double s = *scale > 0 ? 1.0 / *scale : 0.0;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x)
        p3[q] = f2bf((mask && mask[q]) ? 0.f : (float)((a[q] - b[q]) * s));
}

// Loss ||o - t||^2 / ||t||^2 over the 3 used channels: per-block partial sums of
// ((o - t).(o - t), t.t).
__global__ __launch_bounds__(256) void mse_partial_kernel(int64_t nvox, const float *__restrict__ o16,
                                                          const uint16_t *__restrict__ t3, double *__restrict__ partial) {
    __shared__ double sh[2][256];
    double dd = 0, tt = 0;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < nvox * 3;
         q += (int64_t)gridDim.x * blockDim.x) {
        const double t = bf2f(t3[q]), d = o16[(q / 3) * 16 + q % 3] - t;
        dd += d * d;
        tt += t * t;
    }
    sh[0][threadIdx.x] = dd;
    sh[1][threadIdx.x] = tt;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
            for (int k = 0; k < 2; ++k) sh[k][threadIdx.x] += sh[k][threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int k = 0; k < 2; ++k) partial[2 * blockIdx.x + k] = sh[k][0];
}
__global__ __launch_bounds__(256) void mse_finish_kernel(int nb, const double *__restrict__ partial,
                                                         double *__restrict__ sums, double *__restrict__ loss_acc) {
    __shared__ double sh[2][256];
    double a[2] = {0, 0};
    for (int q = threadIdx.x; q < nb; q += 256)
        for (int k = 0; k < 2; ++k) a[k] += partial[2 * q + k];
    for (int k = 0; k < 2; ++k) sh[k][threadIdx.x] = a[k];
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
            for (int k = 0; k < 2; ++k) sh[k][threadIdx.x] += sh[k][threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        sums[0] = sh[0][0];
        sums[1] = sh[1][0];
        if (loss_acc) *loss_acc += sh[1][0] > 0 ? sh[0][0] / sh[1][0] : 0.0;
    }
}
// dL/do = 2 (o - t) / ||t||^2, channels >= 3 get 0
__global__ void mse_grad_kernel(int64_t nvox, const float *__restrict__ o16, const uint16_t *__restrict__ t3,
                                const double *__restrict__ sums, float *__restrict__ d16) {
    const double a = sums[1] > 0 ? 2.0 / sums[1] : 0.0;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < nvox * 16;
         q += (int64_t)gridDim.x * blockDim.x) {
        const int c = (int)(q & 15);
        const int64_t v = q >> 4;
        d16[q] = c < 3 ? (float)(a * (o16[q] - bf2f(t3[3 * v + c]))) : 0.f;
    }
}

// Adam (Kingma & Ba 2015) on FP32 master weights.  Padding weights (input channels 3..15
// of the first layer, output channels 3..15 of the last) start at zero and receive
// exactly zero gradient, so they stay zero.
__global__ void adam_kernel(int64_t n, float *__restrict__ w, const float *__restrict__ g, float *__restrict__ m,
                            float *__restrict__ v, float lr, float b1, float b2, float eps, float c1, float c2) {
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) {
        const float gk = g[k];
        const float mk = b1 * m[k] + (1.f - b1) * gk;
        const float vk = b2 * v[k] + (1.f - b2) * gk * gk;
        m[k] = mk;
        v[k] = vk;
        w[k] -= lr * (mk / c1) / (sqrtf(vk / c2) + eps);
    }
}

inline unsigned grid1(int64_t n, int bs = 256) {
    int64_t g = (n + bs - 1) / bs;
    return (unsigned)(g < 1 ? 1 : (g > 65535 * 4 ? 65535 * 4 : g));
}

template <int COUT, int MODE, class TI, class TO>
void launch_conv_t(const ConvArgs &a, hipStream_t s) {
    const NetDims de = (MODE == NET_TS2) ? a.din : a.dout;
    const int64_t ne = (int64_t)de.nx * de.ny * de.nz;
    const int64_t waves = (ne + 15) / 16;
    dim3 grid((unsigned)((waves + 3) / 4), MODE == NET_TS2 ? 8 : 1);
    hipLaunchKernelGGL((conv_kernel<COUT, MODE, TI, TO>), grid, dim3(256), 0, s, a);
}

template <int CIN, int COUT, int MODE, class TI, class TO>
void launch_conv3_tile_t(const ConvArgs &a, hipStream_t s) {
    dim3 grid((unsigned)((a.dout.nx + T3X - 1) / T3X), (unsigned)((a.dout.ny + T3Y - 1) / T3Y),
              (unsigned)((a.dout.nz + T3Z - 1) / T3Z));
    hipLaunchKernelGGL((conv3_tile_kernel<CIN, COUT, MODE, TI, TO>), grid, dim3(256), 0, s, a);
}

template <int MODE, class TI, class TG>
void launch_wgrad_t(const WgradArgs &a, int nchunks, hipStream_t s) {
    dim3 grid((unsigned)nchunks, mode_taps_host(MODE), (unsigned)((a.cin / 16) * (a.cout / 16)));
    hipLaunchKernelGGL((wgrad_kernel<MODE, TI, TG>), grid, dim3(256), 0, s, a);
}

}  // namespace

int mode_taps_host(int mode) { return (mode == NET_C3 || mode == NET_UPC3) ? 27 : (mode == NET_C2S2 ? 8 : 1); }

// Dispatch over (COUT, MODE, input type, output type).  in_f32: input tensor FP32
// (gradients) else BF16; out_f32: output FP32 else BF16.
void launch_conv(const ConvArgs &a, bool in_f32, bool out_f32, hipStream_t s) {
    if (a.w32 && (a.mode == NET_C3 || a.mode == NET_UPC3) && in_f32 == out_f32) {
#define PF_T3(CI, CO, TI, TO)                                                                         \
    if (a.mode == NET_C3) launch_conv3_tile_t<CI, CO, NET_C3, TI, TO>(a, s);                          \
    else launch_conv3_tile_t<CI, CO, NET_UPC3, TI, TO>(a, s);
#define PF_T3_CO(CI, TI, TO)                                                                          \
    switch (a.cout) {                                                                                 \
        case 16: PF_T3(CI, 16, TI, TO) break;                                                         \
        case 32: PF_T3(CI, 32, TI, TO) break;                                                         \
        case 64: PF_T3(CI, 64, TI, TO) break;                                                         \
        default: throw status_failure(PFEM_ERR_BAD_ARG, "conv cout");                                 \
    }
#define PF_T3_CI(TI, TO)                                                                              \
    switch (a.cin) {                                                                                  \
        case 16: PF_T3_CO(16, TI, TO) break;                                                          \
        case 32: PF_T3_CO(32, TI, TO) break;                                                          \
        case 64: PF_T3_CO(64, TI, TO) break;                                                          \
        default: throw status_failure(PFEM_ERR_BAD_ARG, "conv cin");                                  \
    }
        if (in_f32) { PF_T3_CI(float, float) }
        else { PF_T3_CI(uint16_t, uint16_t) }
#undef PF_T3_CI
#undef PF_T3_CO
#undef PF_T3
        PF_HIP(hipGetLastError());
        return;
    }
#define PF_CONV_MODES(CO, TI, TO)                                                  \
    switch (a.mode) {                                                              \
        case NET_C3: launch_conv_t<CO, NET_C3, TI, TO>(a, s); break;               \
        case NET_C1: launch_conv_t<CO, NET_C1, TI, TO>(a, s); break;               \
        case NET_C2S2: launch_conv_t<CO, NET_C2S2, TI, TO>(a, s); break;           \
        case NET_UPC3: launch_conv_t<CO, NET_UPC3, TI, TO>(a, s); break;           \
        case NET_TS2: launch_conv_t<CO, NET_TS2, TI, TO>(a, s); break;             \
        default: throw status_failure(PFEM_ERR_BAD_ARG, "conv mode");              \
    }
#define PF_CONV_TYPES(CO)                                                          \
    if (!in_f32 && !out_f32) { PF_CONV_MODES(CO, uint16_t, uint16_t) }             \
    else if (!in_f32 && out_f32) { PF_CONV_MODES(CO, uint16_t, float) }            \
    else if (in_f32 && out_f32) { PF_CONV_MODES(CO, float, float) }                \
    else { PF_CONV_MODES(CO, float, uint16_t) }
    if (a.cin % 16 || a.cin < 16 || a.cin > 64) throw status_failure(PFEM_ERR_BAD_ARG, "conv cin");
    switch (a.cout) {
        case 16: PF_CONV_TYPES(16) break;
        case 32: PF_CONV_TYPES(32) break;
        case 64: PF_CONV_TYPES(64) break;
        default: throw status_failure(PFEM_ERR_BAD_ARG, "conv cout");
    }
#undef PF_CONV_TYPES
#undef PF_CONV_MODES
    PF_HIP(hipGetLastError());
}

void launch_wgrad(const WgradArgs &a, int nchunks, bool x_f32, bool g_f32, hipStream_t s) {
#define PF_WG_MODES(TI, TG)                                                        \
    switch (a.mode) {                                                              \
        case NET_C3: launch_wgrad_t<NET_C3, TI, TG>(a, nchunks, s); break;         \
        case NET_C1: launch_wgrad_t<NET_C1, TI, TG>(a, nchunks, s); break;         \
        case NET_C2S2: launch_wgrad_t<NET_C2S2, TI, TG>(a, nchunks, s); break;     \
        case NET_UPC3: launch_wgrad_t<NET_UPC3, TI, TG>(a, nchunks, s); break;     \
        default: throw status_failure(PFEM_ERR_BAD_ARG, "wgrad mode");             \
    }
    if (!x_f32 && g_f32) { PF_WG_MODES(uint16_t, float) }
    else if (!x_f32 && !g_f32) { PF_WG_MODES(uint16_t, uint16_t) }
    else if (x_f32 && g_f32) { PF_WG_MODES(float, float) }
    else { PF_WG_MODES(float, uint16_t) }
#undef PF_WG_MODES
    PF_HIP(hipGetLastError());
}

void launch_wreduce(int64_t nw, int nparts, const float *partial, float *dw, bool accumulate, float scale,
                    hipStream_t s) {
    hipLaunchKernelGGL(wreduce_kernel, dim3(grid1(nw)), dim3(256), 0, s, nw, nparts, partial, dw, accumulate ? 1 : 0,
                       scale);
    PF_HIP(hipGetLastError());
}

void launch_pack(int ntaps, int cin, int cout, const float *W, uint16_t *pk, bool transpose, bool flip,
                 hipStream_t s) {
    const int64_t total = (int64_t)ntaps * cin * cout;
    hipLaunchKernelGGL(pack_kernel, dim3(grid1(total)), dim3(256), 0, s, ntaps, cin, cout, W, pk, transpose ? 1 : 0,
                       flip ? 1 : 0);
    PF_HIP(hipGetLastError());
}

int64_t pack32_elems(int cinp, int coutp) { return (int64_t)((27 * cinp + 31) / 32) * (coutp / 16) * 512; }

void launch_pack32(int cin, int cout, const float *W, uint16_t *pk, bool transpose, bool flip, hipStream_t s) {
    const int64_t total = pack32_elems(transpose ? cout : cin, transpose ? cin : cout);
    hipLaunchKernelGGL(pack32_kernel, dim3(grid1(total)), dim3(256), 0, s, cin, cout, W, pk, transpose ? 1 : 0,
                       flip ? 1 : 0);
    PF_HIP(hipGetLastError());
}

void launch_relu_back(int64_t n, float *d, const float *extra, const uint16_t *y, hipStream_t s) {
    hipLaunchKernelGGL(relu_back_kernel, dim3(grid1(n)), dim3(256), 0, s, n, d, extra, y);
    PF_HIP(hipGetLastError());
}

void launch_pool_sum(NetDims fine, NetDims coarse, int C, const float *df, float *dc, hipStream_t s) {
    const int64_t n = (int64_t)coarse.nx * coarse.ny * coarse.nz * C;
    hipLaunchKernelGGL(pool_sum_kernel, dim3(grid1(n)), dim3(256), 0, s, fine, coarse, C, df, dc);
    PF_HIP(hipGetLastError());
}

void launch_absmax(int64_t n, const double *v, const uint8_t *mask, double *partial, double *out, hipStream_t s) {
    hipLaunchKernelGGL(absmax_partial_kernel, dim3(NET_RED_BLOCKS), dim3(256), 0, s, n, v, mask, partial);
    hipLaunchKernelGGL(max_finish_kernel, dim3(1), dim3(256), 0, s, NET_RED_BLOCKS, partial, out);
    PF_HIP(hipGetLastError());
}

void launch_input_from_dofs(int64_t nvox, const double *r, const uint8_t *mask, const double *scale, uint16_t *x16,
                            hipStream_t s) {
    hipLaunchKernelGGL(input_from_dofs_kernel, dim3(grid1(nvox * 16)), dim3(256), 0, s, nvox, r, mask, scale, x16);
    PF_HIP(hipGetLastError());
}
void launch_input_from_pairs(int64_t nvox, const uint16_t *p3, uint16_t *x16, hipStream_t s) {
    hipLaunchKernelGGL(input_from_pairs_kernel, dim3(grid1(nvox * 16)), dim3(256), 0, s, nvox, p3, x16);
    PF_HIP(hipGetLastError());
}
void launch_output_to_dofs(int64_t nvox, const float *o16, const uint8_t *mask, const double *r, double *z,
                           hipStream_t s) {
    hipLaunchKernelGGL(output_to_dofs_kernel, dim3(grid1(nvox * 3)), dim3(256), 0, s, nvox, o16, mask, r, z);
    PF_HIP(hipGetLastError());
}
void launch_dofs_to_pair(int64_t n, const double *v, const uint8_t *mask, const double *scale, uint16_t *p3,
                         hipStream_t s) {
    hipLaunchKernelGGL(dofs_to_pair_kernel, dim3(grid1(n)), dim3(256), 0, s, n, v, mask, scale, p3);
    PF_HIP(hipGetLastError());
}

void launch_cos_loss(int64_t nvox, const float *o16, const uint16_t *z3, double *partial, double *sums,
                     double *loss_acc, float *d16, hipStream_t s) {
    hipLaunchKernelGGL(cos_partial_kernel, dim3(NET_RED_BLOCKS), dim3(256), 0, s, nvox, o16, z3, partial);
    hipLaunchKernelGGL(cos_finish_kernel, dim3(1), dim3(256), 0, s, NET_RED_BLOCKS, partial, sums, loss_acc);
    if (d16) hipLaunchKernelGGL(cos_grad_kernel, dim3(grid1(nvox * 16)), dim3(256), 0, s, nvox, o16, z3, sums, d16);
    PF_HIP(hipGetLastError());
}

void launch_output_add_scaled(int64_t nvox, const float *o16, const uint8_t *mask, const double *scale, double *z,
                              hipStream_t s) {
    hipLaunchKernelGGL(output_add_scaled_kernel, dim3(grid1(nvox * 3)), dim3(256), 0, s, nvox, o16, mask, scale, z);
    PF_HIP(hipGetLastError());
}
void launch_diff_to_pair(int64_t n, const double *a, const double *b, const uint8_t *mask, const double *scale,
                         uint16_t *p3, hipStream_t s) {
    hipLaunchKernelGGL(diff_to_pair_kernel, dim3(grid1(n)), dim3(256), 0, s, n, a, b, mask, scale, p3);
    PF_HIP(hipGetLastError());
}
void launch_mse_loss(int64_t nvox, const float *o16, const uint16_t *t3, double *partial, double *sums,
                     double *loss_acc, float *d16, hipStream_t s) {
    hipLaunchKernelGGL(mse_partial_kernel, dim3(NET_RED_BLOCKS), dim3(256), 0, s, nvox, o16, t3, partial);
    hipLaunchKernelGGL(mse_finish_kernel, dim3(1), dim3(256), 0, s, NET_RED_BLOCKS, partial, sums, loss_acc);
    if (d16) hipLaunchKernelGGL(mse_grad_kernel, dim3(grid1(nvox * 16)), dim3(256), 0, s, nvox, o16, t3, sums, d16);
    PF_HIP(hipGetLastError());
}

void launch_adam(int64_t n, float *w, const float *g, float *m, float *v, float lr, float b1, float b2, float eps,
                 int step, hipStream_t s) {
    const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
    hipLaunchKernelGGL(adam_kernel, dim3(grid1(n)), dim3(256), 0, s, n, w, g, m, v, lr, b1, b2, eps, c1, c2);
    PF_HIP(hipGetLastError());
}

}  // namespace pfem
