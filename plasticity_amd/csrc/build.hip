// build.hip -- model set-up on the device (Assembly.jl:62-206): node adjacency, the
// SELL-64 3x3-BSR pattern of K, the element scatter slots, the element colouring, and
// the box mesh generator with its seeded jitter (Mesh.jl:30-59).
//
//   incidence   node -> element lists by count (atomics), exclusive scan, fill
//   adjacency   one 64-lane wave per node: the 8 x (incident elements) candidate node
//               ids are staged in LDS; a candidate leads if no earlier candidate has its
//               id, and a leader's position in the sorted row is the number of leaders
//               with a smaller id -- the sorted, de-duplicated neighbour list of
//               _node_adjacency (Assembly.jl:62-82) without a sort
//   SELL        slice width = longest row of its 64 rows (wave max), slice offsets by
//               scan, column indices written column-major within the slice
//   slots       one thread per (element, node a): binary search of each node b in row a
//   colours     structured box numbering (Mesh.jl:30-59) -> the parity colours of
//               element_colors (Assembly.jl:172-188) and each element's position in its
//               colour in closed form; other meshes use the host's greedy colouring
#include <hip/hip_runtime.h>

#include <hipcub/hipcub.hpp>

#include "build.h"

namespace pfem {
namespace {

constexpr int MAXC = 8 * BUILD_MAXINC;  // candidates per node

__global__ void count_incidence_kernel(int64_t n, const int64_t *__restrict__ conn,
                                       unsigned long long *__restrict__ cnt) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        atomicAdd(&cnt[conn[i]], 1ull);
}

__global__ void fill_incidence_kernel(int64_t n, const int64_t *__restrict__ conn,
                                      unsigned long long *__restrict__ pos, int32_t *__restrict__ inc) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const unsigned long long p = atomicAdd(&pos[conn[i]], 1ull);
        inc[p] = (int32_t)(i >> 3);
    }
}

// One wave (block of 64) per node, grid-stride over nodes.
__global__ __launch_bounds__(64) void adjacency_kernel(int64_t nn, const int64_t *__restrict__ conn,
                                                       const int64_t *__restrict__ iptr,
                                                       const int32_t *__restrict__ inc,
                                                       int32_t *__restrict__ rowlen, int32_t *__restrict__ nbr,
                                                       int *__restrict__ err) {
    __shared__ int32_t cand[MAXC];
    __shared__ int32_t lead[MAXC];
    const int lane = threadIdx.x;
    for (int64_t node = blockIdx.x; node < nn; node += gridDim.x) {
        const int64_t b = iptr[node];
        const int64_t ninc = iptr[node + 1] - b;
        if (ninc > BUILD_MAXINC) {  // uniform over the block
            if (lane == 0) {
                atomicOr(err, 1);
                rowlen[node] = 0;
            }
            continue;
        }
        const int C = 8 * (int)ninc;
        for (int t = lane; t < C; t += 64) cand[t] = (int32_t)conn[8 * (int64_t)inc[b + (t >> 3)] + (t & 7)];
        __syncthreads();
        for (int t = lane; t < C; t += 64) {
            const int32_t v = cand[t];
            int first = 1;
            for (int u = 0; u < t; ++u)
                if (cand[u] == v) { first = 0; break; }
            lead[t] = first;
        }
        __syncthreads();
        int mine = 0;
        for (int t = lane; t < C; t += 64) {
            if (!lead[t]) continue;
            ++mine;
            const int32_t v = cand[t];
            int r = 0;
            for (int u = 0; u < C; ++u) r += (lead[u] && cand[u] < v) ? 1 : 0;
            if (r < BUILD_MAXNB) nbr[node * BUILD_MAXNB + r] = v;
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) mine += __shfl_xor(mine, off, 64);
        if (lane == 0) {
            rowlen[node] = mine;
            if (mine > BUILD_MAXNB) atomicOr(err, 2);
        }
        __syncthreads();  // cand/lead are rewritten for the next node
    }
}

// width[sl] = 64 * (longest row of slice sl); one wave per slice.
__global__ __launch_bounds__(64) void slice_width_kernel(int64_t nn, int64_t nslices, const int32_t *__restrict__ rowlen,
                                                         int64_t *__restrict__ width) {
    for (int64_t sl = blockIdx.x; sl < nslices; sl += gridDim.x) {
        const int64_t r = sl * SELL_C + threadIdx.x;
        int w = r < nn ? rowlen[r] : 0;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) w = max(w, __shfl_xor(w, off, 64));
        if (threadIdx.x == 0) width[sl] = (int64_t)SELL_C * w;
    }
}

// Column indices in SELL order; padding entries point at the row itself (value 0).
__global__ void sell_fill_kernel(int64_t nn, int64_t nrows_padded, const int64_t *__restrict__ slice_ptr,
                                 const int32_t *__restrict__ rowlen, const int32_t *__restrict__ nbr,
                                 int32_t *__restrict__ colind, int32_t *__restrict__ diagpos) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nrows_padded;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t sl = i / SELL_C, l = i % SELL_C;
        const int64_t w = (slice_ptr[sl + 1] - slice_ptr[sl]) / SELL_C;
        const int len = i < nn ? rowlen[i] : 0;
        const int32_t pad = (int32_t)(i < nn ? i : nn - 1);
        for (int64_t k = 0; k < w; ++k) {
            const int64_t pos = slice_ptr[sl] + k * SELL_C + l;
            const int32_t v = k < len ? nbr[i * BUILD_MAXNB + k] : pad;
            colind[pos] = v;
            if (k < len && v == i) diagpos[i] = (int32_t)pos;
        }
    }
}

__global__ void slots_kernel(int64_t ne, const int64_t *__restrict__ conn, const int64_t *__restrict__ slice_ptr,
                             const int32_t *__restrict__ rowlen, const int32_t *__restrict__ nbr,
                             int32_t *__restrict__ slots) {
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < 8 * ne; t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t e = t >> 3;
        const int a = (int)(t & 7);
        const int64_t na = conn[8 * e + a];
        const int32_t *row = nbr + na * BUILD_MAXNB;
        const int len = rowlen[na];
        const int64_t sb = slice_ptr[na / SELL_C] + na % SELL_C;
        for (int bb = 0; bb < 8; ++bb) {
            const int32_t v = (int32_t)conn[8 * e + bb];
            int lo = 0, hi = len;  // lower_bound
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (row[mid] < v) lo = mid + 1;
                else hi = mid;
            }
            slots[64 * e + 8 * a + bb] = (int32_t)(sb + (int64_t)lo * SELL_C);
        }
    }
}

// Does conn follow box_mesh's numbering (Mesh.jl:30-59) for an nx x ny x nz grid?
__global__ void box_check_kernel(int64_t ne, int64_t nx, int64_t ny, const int64_t *__restrict__ conn,
                                 int *__restrict__ bad) {
    const int64_t px = nx + 1, pxy = (nx + 1) * (ny + 1);
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < ne; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = e % nx, j = (e / nx) % ny, k = e / (nx * ny);
        const int64_t n0 = i + px * j + pxy * k;
        const int64_t ref[8] = {n0, n0 + 1, n0 + 1 + px, n0 + px, n0 + pxy, n0 + 1 + pxy, n0 + 1 + px + pxy, n0 + px + pxy};
        bool ok = true;
#pragma unroll
        for (int a = 0; a < 8; ++a) ok = ok && conn[8 * e + a] == ref[a];
        if (!ok) atomicOr(bad, 1);
    }
}

// Parity colour class of element (i, j, k) and its position among the earlier elements
// of the same class (elements numbered i fastest, Mesh.jl:48-57).
__device__ __forceinline__ int64_t count_par(int64_t m, int p) { return (m + 1 - p) / 2; }  // #x < m, x%2 == p

__global__ void parity_colour_kernel(int64_t ne, int64_t nx, int64_t ny, int64_t nz, const int32_t *__restrict__ cmap,
                                     const int64_t *__restrict__ cptr, int32_t *__restrict__ colour,
                                     int64_t *__restrict__ order) {
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < ne; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = e % nx, j = (e / nx) % ny, k = e / (nx * ny);
        const int pi = (int)(i & 1), pj = (int)(j & 1), pk = (int)(k & 1);
        const int c = cmap[pi + 2 * pj + 4 * pk];
        const int64_t cx = count_par(nx, pi), cy = count_par(ny, pj);
        const int64_t pos = count_par(k, pk) * cy * cx + count_par(j, pj) * cx + count_par(i, pi);
        colour[e] = c;
        order[cptr[c] + pos] = e;
    }
}

__global__ void identity6_kernel(int64_t ngp, double *__restrict__ cp) {
    for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < ngp; g += (int64_t)gridDim.x * blockDim.x) {
        double *c = cp + 6 * g;
        c[0] = 1.0; c[1] = 1.0; c[2] = 1.0; c[3] = 0.0; c[4] = 0.0; c[5] = 0.0;
    }
}

// splitmix64 (Steele, Lea & Flood, OOPSLA 2014) and the jitter draw
// u = splitmix64(splitmix64(seed) ^ (3 node + comp)) >> 11, times 2^-53 -- the host
// generator's stream, so the device mesh is bit-identical to it.
__device__ __forceinline__ uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

__global__ void box_nodes_kernel(double lx, double ly, double lz, int64_t nx, int64_t ny, int64_t nz, double jit,
                                 uint64_t seed, double *__restrict__ nodes) {
#pragma clang fp contract(off)  // x + d*h rounded twice, as on the host
    const int64_t nn = (nx + 1) * (ny + 1) * (nz + 1);
    const uint64_t s0 = splitmix64(seed);
    const double L[3] = {lx, ly, lz};
    const int64_t cnt[3] = {nx, ny, nz};
    for (int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; n < nn; n += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = n % (nx + 1), j = (n / (nx + 1)) % (ny + 1), k = n / ((nx + 1) * (ny + 1));
        const int64_t idx[3] = {i, j, k};
        const bool interior = i > 0 && i < nx && j > 0 && j < ny && k > 0 && k < nz;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            double x = L[d] * (double)idx[d] / (double)cnt[d];
            if (jit != 0.0 && interior) {
                const uint64_t r = splitmix64(s0 ^ (uint64_t)(3 * n + d));
                const double u = (double)(r >> 11) * 0x1.0p-53;
                x = x + (2.0 * u - 1.0) * jit * (L[d] / (double)cnt[d]);
            }
            nodes[3 * n + d] = x;
        }
    }
}

__global__ void box_conn_kernel(int64_t nx, int64_t ny, int64_t nz, int64_t *__restrict__ conn) {
    const int64_t ne = nx * ny * nz, px = nx + 1, pxy = (nx + 1) * (ny + 1);
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < ne; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = e % nx, j = (e / nx) % ny, k = e / (nx * ny);
        const int64_t n0 = i + px * j + pxy * k;
        int64_t *c = conn + 8 * e;
        c[0] = n0; c[1] = n0 + 1; c[2] = n0 + 1 + px; c[3] = n0 + px;
        c[4] = n0 + pxy; c[5] = n0 + 1 + pxy; c[6] = n0 + 1 + px + pxy; c[7] = n0 + px + pxy;
    }
}

unsigned grid_of(int64_t n, int block) {
    int64_t g = (n + block - 1) / block;
    if (g > 65536) g = 65536;
    return (unsigned)(g < 1 ? 1 : g);
}

// out[0] = 0, out[i + 1] = in[0] + ... + in[i]
void inclusive_offsets(const int64_t *in, int64_t *out, int64_t n, hipStream_t s) {
    PF_HIP(hipMemsetAsync(out, 0, sizeof(int64_t), s));
    size_t tb = 0;
    PF_HIP(hipcub::DeviceScan::InclusiveSum(nullptr, tb, in, out + 1, (int)n, s));
    dbuf<unsigned char> tmp(tb);
    PF_HIP(hipcub::DeviceScan::InclusiveSum(tmp.p, tb, in, out + 1, (int)n, s));
    PF_HIP(hipStreamSynchronize(s));  // tmp is freed on return
}

}  // namespace

int build_pattern_device(int64_t nn, int64_t ne, const int64_t *d_conn, DevicePattern &P, hipStream_t s) {
    if (8 * ne >= ((int64_t)1 << 31) || nn >= ((int64_t)1 << 31)) return BUILD_TOO_LARGE;
    dbuf<unsigned long long> cnt(nn);
    cnt.zero(s);
    hipLaunchKernelGGL(count_incidence_kernel, dim3(grid_of(8 * ne, 256)), dim3(256), 0, s, 8 * ne, d_conn, cnt.p);
    PF_HIP(hipGetLastError());
    dbuf<int64_t> iptr(nn + 1);
    inclusive_offsets(reinterpret_cast<const int64_t *>(cnt.p), iptr.p, nn, s);
    PF_HIP(hipMemcpyAsync(cnt.p, iptr.p, nn * sizeof(int64_t), hipMemcpyDeviceToDevice, s));  // fill cursors
    dbuf<int32_t> inc(8 * ne);
    hipLaunchKernelGGL(fill_incidence_kernel, dim3(grid_of(8 * ne, 256)), dim3(256), 0, s, 8 * ne, d_conn, cnt.p, inc.p);
    PF_HIP(hipGetLastError());
    cnt.release();

    dbuf<int> err(1);
    err.zero(s);
    P.rowlen.alloc(nn);
    dbuf<int32_t> nbr((size_t)nn * BUILD_MAXNB);
    hipLaunchKernelGGL(adjacency_kernel, dim3(grid_of(nn, 1)), dim3(64), 0, s,
                       nn, d_conn, iptr.p, inc.p, P.rowlen.p, nbr.p, err.p);
    PF_HIP(hipGetLastError());
    int herr = 0;
    err.download(&herr, 1, s);
    PF_HIP(hipStreamSynchronize(s));
    if (herr) return BUILD_TOO_DENSE;
    inc.release();
    iptr.release();

    P.nslices = (nn + SELL_C - 1) / SELL_C;
    dbuf<int64_t> width(P.nslices);
    hipLaunchKernelGGL(slice_width_kernel, dim3(grid_of(P.nslices, 1)),
                       dim3(64), 0, s, nn, P.nslices, P.rowlen.p, width.p);
    PF_HIP(hipGetLastError());
    P.slice_ptr.alloc(P.nslices + 1);
    inclusive_offsets(width.p, P.slice_ptr.p, P.nslices, s);
    PF_HIP(hipMemcpyAsync(&P.nslots, P.slice_ptr.p + P.nslices, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    PF_HIP(hipStreamSynchronize(s));
    if (P.nslots >= ((int64_t)1 << 31)) return BUILD_TOO_LARGE;
    P.colind.alloc(P.nslots);
    P.diagpos.alloc(nn);
    hipLaunchKernelGGL(sell_fill_kernel, dim3(grid_of(P.nslices * SELL_C, 256)), dim3(256), 0, s, nn,
                       P.nslices * SELL_C, P.slice_ptr.p, P.rowlen.p, nbr.p, P.colind.p, P.diagpos.p);
    PF_HIP(hipGetLastError());
    P.slots.alloc(64 * ne);
    hipLaunchKernelGGL(slots_kernel, dim3(grid_of(8 * ne, 256)), dim3(256), 0, s, ne, d_conn, P.slice_ptr.p,
                       P.rowlen.p, nbr.p, P.slots.p);
    PF_HIP(hipGetLastError());
    PF_HIP(hipStreamSynchronize(s));  // nbr is freed on return
    return BUILD_OK;
}

int box_dims_from_conn(int64_t nn, int64_t ne, const int64_t *h_conn0, const int64_t *d_conn, int64_t dims[3],
                       hipStream_t s) {
    // element 0 of an nx x ny x nz box: conn = n0 + {0, 1, 1+px, px, pxy, ...}, n0 = 0
    if (h_conn0[0] != 0 || h_conn0[1] != 1) return 0;
    const int64_t px = h_conn0[3], pxy = h_conn0[4];
    if (px < 2 || pxy < 2 * px || pxy % px) return 0;
    const int64_t nx = px - 1, ny = pxy / px - 1;
    if (ne % (nx * ny)) return 0;
    const int64_t nz = ne / (nx * ny);
    if ((nx + 1) * (ny + 1) * (nz + 1) != nn) return 0;
    dbuf<int> bad(1);
    bad.zero(s);
    hipLaunchKernelGGL(box_check_kernel, dim3(grid_of(ne, 256)), dim3(256), 0, s, ne, nx, ny, d_conn, bad.p);
    PF_HIP(hipGetLastError());
    int hb = 0;
    bad.download(&hb, 1, s);
    PF_HIP(hipStreamSynchronize(s));
    if (hb) return 0;
    dims[0] = nx; dims[1] = ny; dims[2] = nz;
    return 1;
}

int parity_colours_device(const int64_t dims[3], int32_t *d_colour, int64_t *d_order, std::vector<int64_t> &cptr,
                          hipStream_t s) {
    const int64_t nx = dims[0], ny = dims[1], nz = dims[2];
    const int64_t n[3] = {nx, ny, nz};
    int32_t cmap[8];
    int64_t csize[8];
    int nc = 0;
    for (int p = 0; p < 8; ++p) {  // element_colors (Assembly.jl:172-188): classes in parity order, empties dropped
        int64_t size = 1;
        for (int d = 0; d < 3; ++d) size *= (n[d] + 1 - ((p >> d) & 1)) / 2;
        cmap[p] = size > 0 ? nc : -1;
        if (size > 0) csize[nc++] = size;
    }
    cptr.assign(nc + 1, 0);
    for (int c = 0; c < nc; ++c) cptr[c + 1] = cptr[c] + csize[c];
    dbuf<int32_t> dmap(8);
    dbuf<int64_t> dptr(nc + 1);
    dmap.upload(cmap, 8, s);
    dptr.upload(cptr.data(), nc + 1, s);
    const int64_t ne = nx * ny * nz;
    hipLaunchKernelGGL(parity_colour_kernel, dim3(grid_of(ne, 256)), dim3(256), 0, s, ne, nx, ny, nz, dmap.p, dptr.p,
                       d_colour, d_order);
    PF_HIP(hipGetLastError());
    PF_HIP(hipStreamSynchronize(s));
    return nc;
}

void fill_identity6(int64_t ngp, double *d_cp, hipStream_t s) {
    hipLaunchKernelGGL(identity6_kernel, dim3(grid_of(ngp, 256)), dim3(256), 0, s, ngp, d_cp);
    PF_HIP(hipGetLastError());
}

void box_mesh_device(double lx, double ly, double lz, int64_t nx, int64_t ny, int64_t nz, double jitter,
                     uint64_t seed, double *d_nodes, int64_t *d_conn, hipStream_t s) {
    const int64_t nn = (nx + 1) * (ny + 1) * (nz + 1), ne = nx * ny * nz;
    hipLaunchKernelGGL(box_nodes_kernel, dim3(grid_of(nn, 256)), dim3(256), 0, s, lx, ly, lz, nx, ny, nz, jitter,
                       seed, d_nodes);
    PF_HIP(hipGetLastError());
    hipLaunchKernelGGL(box_conn_kernel, dim3(grid_of(ne, 256)), dim3(256), 0, s, nx, ny, nz, d_conn);
    PF_HIP(hipGetLastError());
}

}  // namespace pfem
