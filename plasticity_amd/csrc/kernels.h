// kernels.h -- host-side launchers of the device kernels (internal).
#pragma once
#include "common.h"

namespace pfem {

// sparse.hip
void spmv_fine64(const FineMatrix &K, const double *x, double *y, hipStream_t s);
void spmv_fine32(const FineMatrix &K, const float *x, float *y, hipStream_t s);
void spmv_fine64_dot(const FineMatrix &K, const double *x, double *y, double *partial, double *out,
                     hipStream_t s);
void spmv_fine32_dot(const FineMatrix &K, const float *x, float *y, double *partial, double *out,
                     hipStream_t s);
void residual_fine64(const FineMatrix &K, const double *x, const double *b, double *r, hipStream_t s);
size_t reduce_partials_needed(int64_t nb);
// FP16 copy of the diagonally scaled fine K (K.v16, K.sinv) and its products (FP32 vectors)
void half_scaled_copy(FineMatrix &K, hipStream_t s);
// FP32 copy v32 = fl32(K) and its BF16 correction lo16 = bf16(K - v32) in one pass; the
// FP64 product (and the one with the p.y partials) from that split copy
void split_copy(FineMatrix &K, hipStream_t s);
void spmv_fine_split(const FineMatrix &K, const double *x, double *y, hipStream_t s);
void spmv_fine_split_dot(const FineMatrix &K, const double *x, double *y, double *partial, double *out,
                         hipStream_t s);
void spmv_fine_half(const FineMatrix &K, const float *x, float *y, hipStream_t s);
void residual_fine_half(const FineMatrix &K, const float *x, const float *b, float *r, hipStream_t s);
template <class A, class B>
void dot_dev(int64_t n, const A *a, const B *b, double *partial, double *out, hipStream_t s);
// fused flexible-GCR vector work: out[j] = a . v[j] (nv dots in passes of 16; partial
// holds multi_dot_partials() doubles); p -= sum c_j P_j, q -= sum c_j Q_j; and
// x += a p, r -= a q with out = r . r.  Results equal the separate dot/axpby calls.
size_t multi_dot_partials();
void multi_dot(int64_t n, const double *a, const double *const *v, int nv, double *partial, double *out,
               hipStream_t s);
void multi_axpy2(int64_t n, const double *const *P, const double *const *Q, const double *c, int nv, double *p,
                 double *q, hipStream_t s);
void update_dot(int64_t n, double a, const double *p, const double *q, double *x, double *r, double *partial,
                double *out, hipStream_t s);
template <class T>
void axpy_ratio(int64_t n, const double *num, const double *den, double sign, const T *x, T *y, hipStream_t s);
template <class T>
void xpby_ratio(int64_t n, const double *num, const double *den, const T *z, T *p, hipStream_t s);
template <class T>
void cg_update(int64_t n, const double *num, const double *den, const T *p, const T *q, T *x, T *r,
               hipStream_t s);
// cg_update<double> that also writes r32 = (float) r
void cg_update_r32(int64_t n, const double *num, const double *den, const double *p, const double *q, double *x,
                   double *r, float *r32, hipStream_t s);
// out[0] = r'z, out[1] = r'r, out[2] = z'q in one pass (z read as ZT and, when zout is
// not null, written to zout as FP64); partial holds 3 * RED_BLOCKS doubles.  Each
// equals the separate dot product bit for bit.
template <class ZT>
void fcg_dots(int64_t n, const double *r, const ZT *z, double *zout, const double *q, double *partial, double *out,
              hipStream_t s);
// out[0] = r'z, out[1] = r'r, out[2] = z'q of FP32 vectors, FP64 accumulation
void fcg_dots32(int64_t n, const float *r, const float *z, const float *q, double *partial, double *out,
                hipStream_t s);
template <class T>
void axpby(int64_t n, double a, const T *x, double b, T *y, hipStream_t s);
// axpby with b = -(*num / *den) taken from device memory (same arithmetic as axpby)
template <class T>
void axpby_negratio(int64_t n, double a, const T *x, const double *num, const double *den, T *y, hipStream_t s);
void axpy_safe_ratio(int64_t n, const double *num, const double *den, double sign, const double *x, double *y,
                     hipStream_t s);
void cg_update_mixed(int64_t n, const double *num, const double *den, const float *p, const float *q, double *x,
                     float *r, hipStream_t s);
// y += a x, FP32 x into FP64 y
void axpy_f32_to_f64(int64_t n, double a, const float *x, double *y, hipStream_t s);
template <class S, class D>
void convert(int64_t n, const S *src, D *dst, hipStream_t s);
template <class T>
void fill(int64_t n, T v, T *d, hipStream_t s);
void sub_external(int64_t n, double lambda, const double *fext, double *R, hipStream_t s);
// sel: constrained set of this call, a mask of bcflag values (1 fixed, 2 ramped, 3 both)
void impose_dirichlet_dev(FineMatrix &K, int64_t nbc, const int64_t *bcdofs, const uint8_t *flag,
                          const double *val, double lambda, const double *U, double *dg, int64_t nrows,
                          const int32_t *rows, double *R, int sel, hipStream_t s);
void block_diag_inverse3(const FineMatrix &K, double *dinv, hipStream_t s);
template <class T>
void block_jacobi3(int64_t nb, const double *dinv, const T *r, T *z, hipStream_t s);

}  // namespace pfem
