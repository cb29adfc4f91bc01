// batch_noalloc.hip -- pfem_element_batch without device allocation.
//
// The same per-Gauss-point and tangent-contraction kernels as pfem_element_batch
// (element.hip), with every device buffer -- the staged inputs and outputs, the
// per-Gauss-point packets, the centroid gradients, the material and the inversion key --
// carved out of one file-scope device array.  Each GPU holds its own copy of the array;
// its address is looked up once per device (hipGetSymbolAddress) and only that address
// is handed to kernels and copies.  Calls run on the null stream and are synchronous;
// calls that overlap on one GPU share the array and race.
#include <mutex>

#include "common.h"

using namespace pfem;

namespace {

constexpr int64_t NA_MAX = PFEM_BATCH_NOALLOC_MAX;
// doubles per element: X 24, U 24, conn 8 (int64), epsp 48, beta 48, abar 8, cpinv 48,
// sig 48, packets 8 * PK, reference gradients 8 * 24, g0 24, Fe 24, Ke 576
constexpr int64_t NA_PER_ELEM = 24 + 24 + 8 + 48 + 48 + 8 + 48 + 48 + 8 * PK + 8 * 24 + 24 + 24 + 576;
constexpr size_t NA_HEAD = 768;  // material, inversion key, det F of the reported point
constexpr size_t NA_BYTES = NA_HEAD + (size_t)NA_MAX * NA_PER_ELEM * sizeof(double);

__device__ __align__(256) unsigned char scratch[NA_BYTES];

std::mutex g_mu;
constexpr int NA_DEVICES = 64;
unsigned char *g_ptr[NA_DEVICES] = {};

__global__ void iota_kernel(int64_t n, int64_t *__restrict__ d) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        d[i] = i;
}
__global__ void fill_key_kernel(unsigned long long *k) { *k = ~0ull; }
__global__ void zero_kernel(int64_t n, double *__restrict__ d) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        d[i] = 0.0;
}

// This device's copy of the array (looked up on the first call on each device).
int scratch_base(unsigned char **out) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= NA_DEVICES) return PFEM_ERR_NO_DEVICE;
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) return PFEM_ERR_HIP;
        g_ptr[dev] = static_cast<unsigned char *>(p);
    }
    *out = g_ptr[dev];
    return PFEM_OK;
}

}  // namespace

extern "C" int pfem_element_batch_noalloc(int32_t kind, const pfem_material *mat, int64_t nelem, const double *Xe,
                                          const double *ue, double *epsp, double *beta, double *abar, double *cpinv,
                                          int commit, double *Fe, double *Ke, double *sig, pfem_fail *fail) {
    if (kind != PFEM_HEX8_FINITE && kind != PFEM_HEX8_FINITE_FBAR && kind != PFEM_HEX8_SMALL) return PFEM_ERR_BAD_ARG;
    if (!mat || nelem <= 0 || nelem > NA_MAX || !Xe || !ue || !epsp || !abar || (!cpinv && kind != PFEM_HEX8_SMALL) ||
        !Fe || !Ke || !sig)
        return PFEM_ERR_BAD_ARG;
    DevMat dm;
    int rc = make_devmat(*mat, dm);
    if (rc) return rc;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return PFEM_ERR_NO_DEVICE;
    unsigned char *base = nullptr;
    if ((rc = scratch_base(&base))) return rc;
    try {
        const hipStream_t s = nullptr;  // the null stream: no stream object is created
        const int64_t ne = nelem, ngp = 8 * ne;
        DevMat *md = reinterpret_cast<DevMat *>(base);
        auto *key = reinterpret_cast<unsigned long long *>(base + 256);
        double *Jout = reinterpret_cast<double *>(base + 512);
        double *p = reinterpret_cast<double *>(base + NA_HEAD);
        auto take = [&](int64_t count) { double *q = p; p += count; return q; };
        double *X = take(24 * ne), *U = take(24 * ne);
        int64_t *cn = reinterpret_cast<int64_t *>(take(8 * ne));
        double *ep = take(6 * ngp), *bt = take(6 * ngp), *ab = take(ngp), *cp = take(6 * ngp), *sg = take(6 * ngp);
        double *pk = take(ngp * PK), *gr = take(ngp * 24), *g0 = take(24 * ne), *fe = take(24 * ne),
               *ke = take(576 * ne);
        const size_t d = sizeof(double);
        PF_HIP(hipMemcpyAsync(md, &dm, sizeof dm, hipMemcpyHostToDevice, s));
        PF_HIP(hipMemcpyAsync(X, Xe, 24 * ne * d, hipMemcpyHostToDevice, s));
        PF_HIP(hipMemcpyAsync(U, ue, 24 * ne * d, hipMemcpyHostToDevice, s));
        PF_HIP(hipMemcpyAsync(ep, epsp, 6 * ngp * d, hipMemcpyHostToDevice, s));
        PF_HIP(hipMemcpyAsync(ab, abar, ngp * d, hipMemcpyHostToDevice, s));
        const unsigned grid = (unsigned)std::min<int64_t>(1024, (24 * ne + 255) / 256);
        // the array keeps the last call's contents: everything read before it is written
        // is uploaded or cleared here
        if (cpinv) PF_HIP(hipMemcpyAsync(cp, cpinv, 6 * ngp * d, hipMemcpyHostToDevice, s));
        else hipLaunchKernelGGL(zero_kernel, dim3(grid), dim3(256), 0, s, 6 * ngp, cp);
        if (beta) PF_HIP(hipMemcpyAsync(bt, beta, 6 * ngp * d, hipMemcpyHostToDevice, s));
        else hipLaunchKernelGGL(zero_kernel, dim3(grid), dim3(256), 0, s, 6 * ngp, bt);
        hipLaunchKernelGGL(zero_kernel, dim3(grid), dim3(256), 0, s, 6 * ngp, sg);
        hipLaunchKernelGGL(iota_kernel, dim3(grid), dim3(256), 0, s, 8 * ne, cn);
        hipLaunchKernelGGL(fill_key_kernel, dim3(1), dim3(1), 0, s, key);
        PF_HIP(hipGetLastError());
        ElementArgs a{};
        a.nelem = ne; a.X = X; a.conn = cn; a.U = U; a.matid = nullptr; a.mats = md;
        a.kind = kind; a.fbar = kind == PFEM_HEX8_FINITE_FBAR; a.kin = dm.Hkin > 0; a.commit = commit;
        a.epsp = ep; a.beta = bt; a.abar = ab; a.cpinv = cp;
        a.sig = sg; a.packet = pk; a.g0 = g0; a.fail_key = key;
        a.geo = gr; a.geo_write = 1;
        launch_gp_kernel(a, 0, ne, s);
        unsigned long long hk = 0;
        PF_HIP(hipMemcpyAsync(&hk, key, sizeof hk, hipMemcpyDeviceToHost, s));
        PF_HIP(hipStreamSynchronize(s));
        if (hk != ~0ull) {
            launch_detF_kernel(a, (int64_t)hk, Jout, s);
            double hJ = 0;
            PF_HIP(hipMemcpyAsync(&hJ, Jout, d, hipMemcpyDeviceToHost, s));
            PF_HIP(hipStreamSynchronize(s));
            if (fail) { fail->e = (int64_t)(hk / 8); fail->g = (int32_t)(hk % 8); fail->J = hJ; }
            return PFEM_ERR_INVERTED;
        }
        launch_dense_ke_kernel(a, ne, fe, ke, s);
        PF_HIP(hipMemcpyAsync(Fe, fe, 24 * ne * d, hipMemcpyDeviceToHost, s));
        PF_HIP(hipMemcpyAsync(Ke, ke, 576 * ne * d, hipMemcpyDeviceToHost, s));
        PF_HIP(hipMemcpyAsync(sig, sg, 6 * ngp * d, hipMemcpyDeviceToHost, s));
        if (commit) {
            PF_HIP(hipMemcpyAsync(epsp, ep, 6 * ngp * d, hipMemcpyDeviceToHost, s));
            PF_HIP(hipMemcpyAsync(abar, ab, ngp * d, hipMemcpyDeviceToHost, s));
            if (cpinv) PF_HIP(hipMemcpyAsync(cpinv, cp, 6 * ngp * d, hipMemcpyDeviceToHost, s));
            if (beta) PF_HIP(hipMemcpyAsync(beta, bt, 6 * ngp * d, hipMemcpyDeviceToHost, s));
        }
        PF_HIP(hipStreamSynchronize(s));
        return PFEM_OK;
    } catch (const hip_failure &e) {
        if (fail) fail->hip = (int32_t)e.code;
        return PFEM_ERR_HIP;
    }
}
