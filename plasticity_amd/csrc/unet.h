// unet.h -- learned preconditioner: kernel launchers and the network object (internal).
#pragma once
#include "common.h"

namespace pfem {

enum NetMode { NET_C3 = 0, NET_C1 = 1, NET_C2S2 = 2, NET_UPC3 = 3, NET_TS2 = 4 };
constexpr int NET_RED_BLOCKS = 512;
constexpr int NET_LAYERS = 12;

struct NetDims {
    int nx, ny, nz;
};
inline int64_t nvox_of(const NetDims &d) { return (int64_t)d.nx * d.ny * d.nz; }

struct ConvArgs {
    int mode, cin, cout, relu;
    NetDims din, dout;
    const void *in;    // input tensor (BF16 or FP32), cin channels on din
    const void *w;     // packed BF16 weights (16x16x16 B fragments, conv_kernel)
    const void *w32;   // packed BF16 weights (16x16x32 B fragments, conv3_tile_kernel) or null
    const void *skip;  // BF16, cout channels on dout, added before the ReLU (or null)
    void *out;         // cout channels on dout
};
struct WgradArgs {
    int mode, cin, cout;
    NetDims din, dout;
    const void *x;  // layer input, cin channels on din
    const void *g;  // gradient at the layer output (after the ReLU mask), cout channels on dout
    float *partial;
    int64_t chunk;  // voxels of dout per block, multiple of 64
};

int mode_taps_host(int mode);
void launch_conv(const ConvArgs &a, bool in_f32, bool out_f32, hipStream_t s);
void launch_wgrad(const WgradArgs &a, int nchunks, bool x_f32, bool g_f32, hipStream_t s);
void launch_wreduce(int64_t nw, int nparts, const float *partial, float *dw, bool accumulate, float scale,
                    hipStream_t s);
void launch_pack(int ntaps, int cin, int cout, const float *W, uint16_t *pk, bool transpose, bool flip,
                 hipStream_t s);
// BF16 elements of the 16x16x32 B-fragment packing of a 3x3x3 layer whose GEMM has
// cinp input and coutp output channels (K = 27 cinp in steps of 32, N = coutp)
int64_t pack32_elems(int cinp, int coutp);
void launch_pack32(int cin, int cout, const float *W, uint16_t *pk, bool transpose, bool flip, hipStream_t s);
void launch_relu_back(int64_t n, float *d, const float *extra, const uint16_t *y, hipStream_t s);
void launch_pool_sum(NetDims fine, NetDims coarse, int C, const float *df, float *dc, hipStream_t s);
void launch_absmax(int64_t n, const double *v, const uint8_t *mask, double *partial, double *out, hipStream_t s);
void launch_input_from_dofs(int64_t nvox, const double *r, const uint8_t *mask, const double *scale, uint16_t *x16,
                            hipStream_t s);
void launch_input_from_pairs(int64_t nvox, const uint16_t *p3, uint16_t *x16, hipStream_t s);
void launch_output_to_dofs(int64_t nvox, const float *o16, const uint8_t *mask, const double *r, double *z,
                           hipStream_t s);
void launch_dofs_to_pair(int64_t n, const double *v, const uint8_t *mask, const double *scale, uint16_t *p3,
                         hipStream_t s);
void launch_cos_loss(int64_t nvox, const float *o16, const uint16_t *z3, double *partial, double *sums,
                     double *loss_acc, float *d16, hipStream_t s);
void launch_output_add_scaled(int64_t nvox, const float *o16, const uint8_t *mask, const double *scale, double *z,
                              hipStream_t s);
void launch_diff_to_pair(int64_t n, const double *a, const double *b, const uint8_t *mask, const double *scale,
                         uint16_t *p3, hipStream_t s);
void launch_mse_loss(int64_t nvox, const float *o16, const uint16_t *t3, double *partial, double *sums,
                     double *loss_acc, float *d16, hipStream_t s);
void launch_adam(int64_t n, float *w, const float *g, float *m, float *v, float lr, float b1, float b2, float eps,
                 int step, hipStream_t s);

// One convolution of the network.
struct NetLayer {
    int mode, cin, cout, lin, lout, relu, skip;  // skip: layer whose output is added, or -1
    int ntaps;
    int64_t woff, nw;                            // slice of the master weight vector
    dbuf<uint16_t> pk_f, pk_b;                   // packed BF16 weights: forward, backward-data
    dbuf<uint16_t> pk32_f, pk32_b;               // the same for the LDS-tiled 3x3x3 kernel (C3, UPC3)
    int nchunks = 1;
    int64_t chunk = 64;
};

}  // namespace pfem

// The network object behind the opaque pfem_net handle.
struct pfem_net {
    hipStream_t stream = nullptr;
    int device = 0;
    pfem::NetDims dims[4];
    int64_t nvox[4];
    int ch[4];
    pfem::NetLayer L[pfem::NET_LAYERS];
    int64_t nparams = 0;
    pfem::dbuf<float> W, G, M, V;                // master weights, gradient, Adam moments
    pfem::dbuf<uint16_t> x16;                    // input, 16 channels on level 0
    pfem::dbuf<uint16_t> act[pfem::NET_LAYERS];  // layer outputs (BF16); the last is unused
    pfem::dbuf<float> out16;                     // last layer output, FP32
    pfem::dbuf<float> gr[pfem::NET_LAYERS];      // gradients at layer outputs (FP32)
    pfem::dbuf<float> tmp;                       // upsample-transpose scratch (fine level)
    pfem::dbuf<float> partial;                   // weight-gradient partials
    pfem::dbuf<double> red, scal;                // reductions; scal[0..2] dots, [3] loss sum, [4..5] scales
    int adam_step = 0;
    bool packed = false;
    int conv_kind = 1;  // 3x3x3 layers: 1 LDS-tiled 16x16x32 kernel, 0 per-voxel-wave 16x16x16 kernel
    int loss_kind = 0;  // training loss: 0 = 1 - cos(N(x), t), 1 = ||N(x) - t||^2 / ||t||^2
    // samples of an error-logging solve awaiting its solution: the BF16 input (pend_in),
    // x_k + z_k (FP64, pend_u) and the input's scale (pend_scale); they enter the pair
    // store only when the solve has converged
    static constexpr int MAX_PEND = 16;
    int npend = 0;
    pfem::dbuf<uint16_t> pend_in;
    pfem::dbuf<double> pend_u, pend_scale;
    // (residual, correction) sample store: capacity pairs of 2 x 3*nvox[0] BF16
    int64_t capacity = 0, npairs = 0, next = 0;
    pfem::dbuf<uint16_t> pairs;
};

namespace pfem {
void net_forward(pfem_net *n, hipStream_t s);                      // x16 -> out16
void net_backward(pfem_net *n, bool accumulate, float scale, hipStream_t s);  // gr[11] = dL/dout16 -> G
void net_pack(pfem_net *n, hipStream_t s);
// z = N(r) for a nodal vector on the net's grid: constrained DOFs (mask) are dropped
// on input and set to 0 on output.
void net_apply(pfem_net *n, const double *r, const uint8_t *mask, double *z, hipStream_t s);
// Append the sample (r, z) (constrained DOFs zeroed, each scaled by its max |.|).
void net_log_pair(pfem_net *n, const double *r, const double *z, const uint8_t *mask, hipStream_t s);
// The correction of a V-cycle output z: z += c N(z / c), c = max |z| over the free DOFs.
void net_correct(pfem_net *n, double *z, const uint8_t *mask, hipStream_t s);
// Error samples of an iterative solve: at iterate x_k with V-cycle output z_k, reserve a
// sample with input z_k / c_k (c_k = max |z_k|); once the solve has converged to x*,
// its target (x* - x_k - z_k) / c_k -- the error the V-cycle left, on the input's scale --
// is written and the samples are kept (keep) or dropped.
void net_log_error_begin(pfem_net *n, const double *z, const double *x, const uint8_t *mask, hipStream_t s);
void net_log_error_finish(pfem_net *n, const double *xstar, const uint8_t *mask, bool keep, hipStream_t s);
}  // namespace pfem
