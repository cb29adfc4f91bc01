// common.h -- internal declarations shared by the libpfem translation units.
// Not part of the C-ABI (include/pfem.h is).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "pfem.h"

namespace pfem {

// HIP error -> exception carrying the code; the C-ABI layer maps it to PFEM_ERR_HIP.
struct hip_failure : std::runtime_error {
    hipError_t code;
    hip_failure(hipError_t c, const char *where)
        : std::runtime_error(std::string(where) + ": " + hipGetErrorString(c)), code(c) {}
};
struct status_failure : std::runtime_error {
    int status;
    status_failure(int s, const std::string &m) : std::runtime_error(m), status(s) {}
};

#define PF_HIP(call)                                                        \
    do {                                                                    \
        hipError_t e_ = (call);                                             \
        if (e_ != hipSuccess) throw ::pfem::hip_failure(e_, #call);         \
    } while (0)

// Device buffer owned by a context (RAII).
template <class T>
struct dbuf {
    T *p = nullptr;
    size_t n = 0;
    dbuf() = default;
    explicit dbuf(size_t count) { alloc(count); }
    dbuf(const dbuf &) = delete;
    dbuf &operator=(const dbuf &) = delete;
    dbuf(dbuf &&o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
    dbuf &operator=(dbuf &&o) noexcept {
        if (this != &o) { release(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; }
        return *this;
    }
    ~dbuf() { release(); }
    void alloc(size_t count) {
        release();
        n = count;
        if (count) PF_HIP(hipMalloc(&p, count * sizeof(T)));
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    size_t bytes() const { return n * sizeof(T); }
    void upload(const T *h, size_t count, hipStream_t s) {
        if (count) PF_HIP(hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, s));
    }
    void download(T *h, size_t count, hipStream_t s) const {
        if (count) PF_HIP(hipMemcpyAsync(h, p, count * sizeof(T), hipMemcpyDeviceToHost, s));
    }
    void zero(hipStream_t s) {
        if (n) PF_HIP(hipMemsetAsync(p, 0, n * sizeof(T), s));
    }
};

// Material constants as used on the device (derived like J2Material, Materials.jl:104-107).
struct DevMat {
    double E, nu, sy0, Hiso, Hkin, ssat, delta;
    double G, K, lam;
};
int make_devmat(const pfem_material &m, DevMat &d);  // PFEM_OK or PFEM_ERR_BAD_ARG / UNSUPPORTED

// Fine-level stiffness in sliced block-ELL ("SELL-64") layout: 64 consecutive block
// rows form a slice; slice s stores its blocks column by column, position
// pos(i, k) = slice_ptr[i/64] + 64*k + i%64, each block 9 contiguous values (row-major).
constexpr int SELL_C = 64;

struct FineMatrix {
    int64_t nb = 0;         // block rows (= nodes)
    int64_t nslots = 0;     // padded block slots
    int64_t nnzb = 0;       // true blocks
    int64_t nslices = 0;
    dbuf<int64_t> slice_ptr;  // nslices + 1
    dbuf<int32_t> rowlen;     // nb
    dbuf<int32_t> colind;     // nslots (padding -> own row, value 0)
    dbuf<int32_t> diagpos;    // nb: slot of the diagonal block
    dbuf<double> v64;         // 9 * nslots
    dbuf<float> v32;          // 9 * nslots (FP32-IR copy)
    dbuf<uint16_t> lo16;      // 9 * nslots, BF16 bits of K - v32, 9 planes of 64 per slice column
    dbuf<uint16_t> v16;       // 9 * nslots, FP16 bits of S K S (FP32 V-cycle fine level)
    dbuf<float> sinv;         // 3 * nb, sqrt|K_ii| (S^-1)
    std::vector<int64_t> h_slice_ptr;
    std::vector<int32_t> h_rowlen, h_colind;  // host copies of the pattern
};

// Element kernels (element.hip)
struct ElementArgs {
    int64_t nelem;
    const double *X;        // 3 x nnodes
    const int64_t *conn;    // 8 x nelem
    const double *U;        // ndof
    const int32_t *matid;   // nelem or null
    const DevMat *mats;     // device array
    int kind;               // pfem_element_kind: 0 finite, 1 finite F-bar, 2 small strain
    int fbar;               // kind == 1
    int kin;                // some material has Hkin > 0 (objective back-stress branch)
    int commit;
    double *epsp, *abar, *cpinv, *sig;  // Gauss state (6 x ngp / ngp)
    double *beta;           // back stress (6 x ngp); reference configuration for finite strain
    double *packet;         // per GP packet (pk_len doubles), 8 (i - e_begin) + g for list position i
    double *g0;             // per element 24
    unsigned long long *fail_key;  // min (8e+g) of inverted GPs, ~0ull if none
    const int64_t *elist;   // element of list position i (null: element i)
    double *geo;            // per GP 24: reference gradients dN/dX (written when geo_write)
    int geo_write;
};
// Gauss-point packet, written by the Gauss-point kernels and read by the contraction:
//   [0, na)           s w A: A = dP/dFbar (s the F-bar factor, w the weight; the
//                     small-strain kernel's A holds w D).  Where A has major symmetry
//                     (every case but the objective back stress of finite-strain
//                     kinematic hardening) its upper triangle, packed by columns:
//                     (r, c), r <= c, at c (c + 1) / 2 + r (na = 45); else all 81, r + 9c.
//   [na, na + 9)      (s w / 3) A vec(F)   (F-bar)
//   [na + 9, na + 18) F^-1                 (F-bar)
//   [na + 18, na + 27) w P
// The reference gradients of the Gauss point (24) are in the geometry array `geo`, written
// once per context.
constexpr int PK_SYM = 72, PK_FULL = 108;
constexpr int PK = PK_FULL;  // the larger layout, for buffer sizes
__host__ __device__ constexpr int pk_len(bool sym) { return sym ? PK_SYM : PK_FULL; }
__host__ __device__ constexpr int pk_na(bool sym) { return sym ? 45 : 81; }
__host__ __device__ constexpr int pk_sym_index(int r, int c) {  // packed upper triangle
    return r <= c ? c * (c + 1) / 2 + r : r * (r + 1) / 2 + c;
}
// packets of this launch's kind and materials use the packed symmetric layout
inline bool packet_sym(const ElementArgs &a) { return a.kind == 2 || !a.kin; }

void launch_gp_kernel(const ElementArgs &a, int64_t e_begin, int64_t e_count, hipStream_t s);
// first: per element, bit 8a+b set where block (a, b) is the first contribution to its
// slot in colour order (stored, not added: K needs no clearing before the scatter)
// local: the packets of elem_list[li] are at a.packet + 8 li PK (a chunk's buffer) instead
// of a.packet + 8 e PK (one packet per Gauss point of the mesh)
void launch_scatter_kernel(const ElementArgs &a, const int64_t *elem_list, int64_t count,
                           const int32_t *slots, const uint64_t *first, double *Kv, double *R, hipStream_t s,
                           bool local = false);
void build_first_touch(int64_t nslots, int ncolors, const int64_t *order, const int64_t *color_ptr,
                       const int32_t *slots, uint64_t *mask, hipStream_t s);
void launch_dense_ke_kernel(const ElementArgs &a, int64_t nelem, double *Fe, double *Ke,
                            hipStream_t s);
void launch_detF_kernel(const ElementArgs &a, int64_t gp, double *out, hipStream_t s);
// gauss_stress: out = tau / det F (un-barred F) per Gauss point; copy for small strain
void launch_cauchy_kernel(const ElementArgs &a, const double *tau, double *out, hipStream_t s);

}  // namespace pfem
