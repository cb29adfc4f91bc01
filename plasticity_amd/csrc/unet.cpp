// unet.cpp -- the learned preconditioner: network set-up, forward and backward passes,
// Adam training on logged (residual, AMG correction) samples, and the C-ABI.
//
// Network (bias-free, ReLU, so N(c r) = c N(r) for c > 0; channels c0..c3 per level,
// level l+1 has ceil(n/2) nodes per direction):
//   0 in  C3   16(3 used) -> c0   L0        6 d3  C2S2 c2 -> c3   L2 -> L3
//   1 e0  C3   c0 -> c0      L0             7 e3  C3   c3 -> c3   L3
//   2 d1  C2S2 c0 -> c1      L0 -> L1       8 u2  UPC3 c3 -> c2   L3 -> L2, + e2
//   3 e1  C3   c1 -> c1      L1             9 u1  UPC3 c2 -> c1   L2 -> L1, + e1
//   4 d2  C2S2 c1 -> c2      L1 -> L2      10 u0  UPC3 c1 -> c0   L1 -> L0, + e0
//   5 e2  C3   c2 -> c2      L2            11 out C1   c0 -> 16(3 used), linear, FP32
// "+ eK": the encoder output is added before the ReLU (additive skip connection).
#include "unet.h"

#include <algorithm>
#include <cmath>
#include <cstring>

#include "context.h"

using namespace pfem;

namespace pfem {
namespace {

uint64_t splitmix64(uint64_t &s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
double u01(uint64_t &s) { return (double)(splitmix64(s) >> 11) * 0x1.0p-53; }

uint16_t host_f2bf(float f) {
    uint32_t b;
    std::memcpy(&b, &f, 4);
    b += 0x7fffu + ((b >> 16) & 1u);
    return (uint16_t)(b >> 16);
}

const void *layer_input(pfem_net *n, int k) { return k == 0 ? (const void *)n->x16.p : (const void *)n->act[k - 1].p; }

void init_weights(pfem_net *n, uint64_t seed, std::vector<float> &w) {
    w.assign(n->nparams, 0.f);
    uint64_t s = seed ^ 0x5DEECE66Dull;
    for (int k = 0; k < NET_LAYERS; ++k) {
        const NetLayer &l = n->L[k];
        const int cin_used = k == 0 ? 3 : l.cin;
        const int cout_used = k == NET_LAYERS - 1 ? 3 : l.cout;
        const double fan_in = (double)l.ntaps * cin_used;
        const double sd = (k == NET_LAYERS - 1) ? std::sqrt(1.0 / fan_in) : std::sqrt(2.0 / fan_in);  // He
        for (int t = 0; t < l.ntaps; ++t)
            for (int ci = 0; ci < cin_used; ++ci)
                for (int co = 0; co < cout_used; ++co) {
                    const double a = std::max(u01(s), 1e-300), b = u01(s);  // Box-Muller
                    const double g = std::sqrt(-2.0 * std::log(a)) * std::cos(6.283185307179586 * b);
                    w[l.woff + ((int64_t)t * l.cin + ci) * l.cout + co] = (float)(sd * g);
                }
    }
}

template <class F>
int net_guard(F &&f) {
    try {
        return f();
    } catch (const hip_failure &) {
        return PFEM_ERR_HIP;
    } catch (const status_failure &e) {
        return e.status;
    } catch (const std::exception &) {
        return PFEM_ERR_BAD_ARG;
    }
}

}  // namespace

void net_pack(pfem_net *n, hipStream_t s) {
    for (int k = 0; k < NET_LAYERS; ++k) {
        NetLayer &l = n->L[k];
        const float *W = n->W.p + l.woff;
        launch_pack(l.ntaps, l.cin, l.cout, W, l.pk_f.p, false, false, s);
        if (k > 0) launch_pack(l.ntaps, l.cin, l.cout, W, l.pk_b.p, true, l.mode == NET_C3 || l.mode == NET_UPC3, s);
        if (l.pk32_f.p) launch_pack32(l.cin, l.cout, W, l.pk32_f.p, false, false, s);
        if (l.pk32_b.p) launch_pack32(l.cin, l.cout, W, l.pk32_b.p, true, true, s);
    }
    n->packed = true;
}

void net_forward(pfem_net *n, hipStream_t s) {
    for (int k = 0; k < NET_LAYERS; ++k) {
        const NetLayer &l = n->L[k];
        ConvArgs a{};
        a.mode = l.mode;
        a.cin = l.cin;
        a.cout = l.cout;
        a.relu = l.relu;
        a.din = n->dims[l.lin];
        a.dout = n->dims[l.lout];
        a.in = layer_input(n, k);
        a.w = l.pk_f.p;
        a.w32 = n->conv_kind == 1 ? l.pk32_f.p : nullptr;
        a.skip = l.skip >= 0 ? (const void *)n->act[l.skip].p : nullptr;
        const bool last = k == NET_LAYERS - 1;
        a.out = last ? (void *)n->out16.p : (void *)n->act[k].p;
        launch_conv(a, false, last, s);
    }
}

void net_backward(pfem_net *n, bool accumulate, float scale, hipStream_t s) {
    for (int k = NET_LAYERS - 1; k >= 0; --k) {
        NetLayer &l = n->L[k];
        float *g = n->gr[k].p;
        if (k < NET_LAYERS - 1) {
            const float *extra = nullptr;  // gradient arriving through the skip connection
            for (int j = 0; j < NET_LAYERS; ++j)
                if (n->L[j].skip == k) extra = n->gr[j].p;
            launch_relu_back(n->nvox[l.lout] * l.cout, g, extra, n->act[k].p, s);
        }
        WgradArgs w{};
        w.mode = l.mode;
        w.cin = l.cin;
        w.cout = l.cout;
        w.din = n->dims[l.lin];
        w.dout = n->dims[l.lout];
        w.x = layer_input(n, k);
        w.g = g;
        w.partial = n->partial.p;
        w.chunk = l.chunk;
        launch_wgrad(w, l.nchunks, false, true, s);
        launch_wreduce(l.nw, l.nchunks * 4, n->partial.p, n->G.p + l.woff, accumulate, scale, s);
        if (k == 0) break;
        ConvArgs a{};
        a.cin = l.cout;  // backward data: K runs over the layer's output channels
        a.cout = l.cin;
        a.relu = 0;
        a.in = g;
        a.w = l.pk_b.p;
        a.w32 = n->conv_kind == 1 ? l.pk32_b.p : nullptr;
        a.skip = nullptr;
        float *gin = n->gr[k - 1].p;
        switch (l.mode) {
            case NET_C3:
            case NET_C1:
                a.mode = l.mode;
                a.din = a.dout = n->dims[l.lout];
                a.out = gin;
                launch_conv(a, true, true, s);
                break;
            case NET_C2S2:
                a.mode = NET_TS2;
                a.din = n->dims[l.lout];
                a.dout = n->dims[l.lin];
                a.out = gin;
                launch_conv(a, true, true, s);
                break;
            case NET_UPC3:
                a.mode = NET_C3;
                a.din = a.dout = n->dims[l.lout];
                a.out = n->tmp.p;
                launch_conv(a, true, true, s);
                launch_pool_sum(n->dims[l.lout], n->dims[l.lin], l.cin, n->tmp.p, gin, s);
                break;
        }
    }
}

void net_apply(pfem_net *n, const double *r, const uint8_t *mask, double *z, hipStream_t s) {
    if (!n->packed) net_pack(n, s);
    launch_absmax(3 * n->nvox[0], r, mask, n->red.p, n->scal.p + 4, s);
    launch_input_from_dofs(n->nvox[0], r, mask, n->scal.p + 4, n->x16.p, s);
    net_forward(n, s);
    launch_output_to_dofs(n->nvox[0], n->out16.p, mask, nullptr, z, s);
}

void net_log_pair(pfem_net *n, const double *r, const double *z, const uint8_t *mask, hipStream_t s) {
    if (n->capacity <= 0) return;
    const int64_t m = 3 * n->nvox[0];
    uint16_t *slot = n->pairs.p + n->next * 2 * m;
    launch_absmax(m, r, mask, n->red.p, n->scal.p + 4, s);
    launch_dofs_to_pair(m, r, mask, n->scal.p + 4, slot, s);
    launch_absmax(m, z, mask, n->red.p, n->scal.p + 5, s);
    launch_dofs_to_pair(m, z, mask, n->scal.p + 5, slot + m, s);
    n->next = (n->next + 1) % n->capacity;
    n->npairs = std::min(n->npairs + 1, n->capacity);
}

void net_correct(pfem_net *n, double *z, const uint8_t *mask, hipStream_t s) {
    if (!n->packed) net_pack(n, s);
    launch_absmax(3 * n->nvox[0], z, mask, n->red.p, n->scal.p + 4, s);
    launch_input_from_dofs(n->nvox[0], z, mask, n->scal.p + 4, n->x16.p, s);
    net_forward(n, s);
    launch_output_add_scaled(n->nvox[0], n->out16.p, mask, n->scal.p + 4, z, s);
}

void net_log_error_begin(pfem_net *n, const double *z, const double *x, const uint8_t *mask, hipStream_t s) {
    if (n->capacity <= 0 || n->npend >= pfem_net::MAX_PEND || n->npend >= n->capacity) return;
    const int64_t m = 3 * n->nvox[0];
    if (!n->pend_u.p) {
        n->pend_u.alloc((size_t)pfem_net::MAX_PEND * m);
        n->pend_in.alloc((size_t)pfem_net::MAX_PEND * m);
        n->pend_scale.alloc(pfem_net::MAX_PEND);
    }
    const int k = n->npend++;
    double *u = n->pend_u.p + (size_t)k * m, *sc = n->pend_scale.p + k;
    launch_absmax(m, z, mask, n->red.p, sc, s);
    launch_dofs_to_pair(m, z, mask, sc, n->pend_in.p + (size_t)k * m, s);
    PF_HIP(hipMemcpyAsync(u, z, m * sizeof(double), hipMemcpyDeviceToDevice, s));
    axpby<double>(m, 1.0, x, 1.0, u, s);  // u = x_k + z_k
}

void net_log_error_finish(pfem_net *n, const double *xstar, const uint8_t *mask, bool keep, hipStream_t s) {
    const int64_t m = 3 * n->nvox[0];
    if (keep) {
        for (int k = 0; k < n->npend; ++k) {
            uint16_t *slot = n->pairs.p + ((n->next + k) % n->capacity) * 2 * m;
            PF_HIP(hipMemcpyAsync(slot, n->pend_in.p + (size_t)k * m, m * sizeof(uint16_t), hipMemcpyDeviceToDevice,
                                  s));
            launch_diff_to_pair(m, xstar, n->pend_u.p + (size_t)k * m, mask, n->pend_scale.p + k, slot + m, s);
        }
        n->next = (n->next + n->npend) % n->capacity;
        n->npairs = std::min<int64_t>(n->npairs + n->npend, n->capacity);
    }
    n->npend = 0;
}

// One sample: forward from the BF16 3-channel input, the loss (1 - cos(out, target) or
// ||out - target||^2 / ||target||^2) and the backward pass into G.
static void net_sample(pfem_net *n, const uint16_t *in3, const uint16_t *tg3, bool accumulate, float scale,
                       hipStream_t s) {
    launch_input_from_pairs(n->nvox[0], in3, n->x16.p, s);
    net_forward(n, s);
    if (n->loss_kind == 1)
        launch_mse_loss(n->nvox[0], n->out16.p, tg3, n->red.p, n->scal.p, n->scal.p + 3, n->gr[NET_LAYERS - 1].p, s);
    else
        launch_cos_loss(n->nvox[0], n->out16.p, tg3, n->red.p, n->scal.p, n->scal.p + 3, n->gr[NET_LAYERS - 1].p, s);
    net_backward(n, accumulate, scale, s);
}

}  // namespace pfem

extern "C" {

int pfem_net_create(const pfem_net_desc *d, pfem_net **out) {
    if (!d || !out) return PFEM_ERR_BAD_ARG;
    *out = nullptr;
    if (d->nx < 2 || d->ny < 2 || d->nz < 2 || d->capacity < 0) return PFEM_ERR_BAD_ARG;
    for (int l = 0; l < 4; ++l)
        if (d->channels[l] != 16 && d->channels[l] != 32 && d->channels[l] != 64) return PFEM_ERR_BAD_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return PFEM_ERR_NO_DEVICE;
    pfem_net *n = new pfem_net();
    const int rc = net_guard([&]() -> int {
        PF_HIP(hipGetDevice(&n->device));
        PF_HIP(hipStreamCreateWithFlags(&n->stream, hipStreamNonBlocking));
        n->dims[0] = {d->nx, d->ny, d->nz};
        for (int l = 1; l < 4; ++l)
            n->dims[l] = {(n->dims[l - 1].nx + 1) / 2, (n->dims[l - 1].ny + 1) / 2, (n->dims[l - 1].nz + 1) / 2};
        for (int l = 0; l < 4; ++l) {
            n->nvox[l] = nvox_of(n->dims[l]);
            n->ch[l] = d->channels[l];
        }
        const int *c = n->ch;
        // mode, cin, cout, lin, lout, relu, skip
        const int spec[NET_LAYERS][7] = {
            {NET_C3, 16, c[0], 0, 0, 1, -1},      {NET_C3, c[0], c[0], 0, 0, 1, -1},
            {NET_C2S2, c[0], c[1], 0, 1, 1, -1},  {NET_C3, c[1], c[1], 1, 1, 1, -1},
            {NET_C2S2, c[1], c[2], 1, 2, 1, -1},  {NET_C3, c[2], c[2], 2, 2, 1, -1},
            {NET_C2S2, c[2], c[3], 2, 3, 1, -1},  {NET_C3, c[3], c[3], 3, 3, 1, -1},
            {NET_UPC3, c[3], c[2], 3, 2, 1, 5},   {NET_UPC3, c[2], c[1], 2, 1, 1, 3},
            {NET_UPC3, c[1], c[0], 1, 0, 1, 1},   {NET_C1, c[0], 16, 0, 0, 0, -1}};
        int64_t off = 0, max_partial = 0, max_tmp = 0;
        for (int k = 0; k < NET_LAYERS; ++k) {
            NetLayer &l = n->L[k];
            l.mode = spec[k][0];
            l.cin = spec[k][1];
            l.cout = spec[k][2];
            l.lin = spec[k][3];
            l.lout = spec[k][4];
            l.relu = spec[k][5];
            l.skip = spec[k][6];
            l.ntaps = mode_taps_host(l.mode);
            l.woff = off;
            l.nw = (int64_t)l.ntaps * l.cin * l.cout;
            off += l.nw;
            l.pk_f.alloc(l.nw);
            if (k > 0) l.pk_b.alloc(l.nw);
            if (l.mode == NET_C3 || l.mode == NET_UPC3) {
                l.pk32_f.alloc(pack32_elems(l.cin, l.cout));
                if (k > 0) l.pk32_b.alloc(pack32_elems(l.cout, l.cin));
            }
            const int64_t ne = n->nvox[l.lout];
            const int tiles = (l.cin / 16) * (l.cout / 16);
            int64_t target = std::max<int64_t>(1, 2048 / ((int64_t)l.ntaps * tiles));
            int64_t nch = std::min<int64_t>(target, (ne + 63) / 64);
            int64_t chunk = ((ne + nch - 1) / nch + 63) / 64 * 64;
            l.chunk = chunk;
            l.nchunks = (int)((ne + chunk - 1) / chunk);
            max_partial = std::max<int64_t>(max_partial, (int64_t)l.nchunks * 4 * l.nw);
            if (l.mode == NET_UPC3) max_tmp = std::max<int64_t>(max_tmp, n->nvox[l.lout] * l.cin);
            if (k < NET_LAYERS - 1) n->act[k].alloc(n->nvox[l.lout] * l.cout);
            n->gr[k].alloc(n->nvox[l.lout] * l.cout);
        }
        n->nparams = off;
        n->W.alloc(off);
        n->G.alloc(off);
        n->M.alloc(off);
        n->V.alloc(off);
        n->M.zero(n->stream);
        n->V.zero(n->stream);
        n->G.zero(n->stream);
        n->x16.alloc(n->nvox[0] * 16);
        n->out16.alloc(n->nvox[0] * 16);
        n->tmp.alloc(std::max<int64_t>(max_tmp, 1));
        n->partial.alloc(max_partial);
        n->red.alloc(3 * NET_RED_BLOCKS);
        n->scal.alloc(8);
        n->scal.zero(n->stream);
        n->capacity = d->capacity;
        if (n->capacity > 0) n->pairs.alloc((size_t)n->capacity * 2 * 3 * n->nvox[0]);
        std::vector<float> w;
        init_weights(n, d->seed, w);
        n->W.upload(w.data(), w.size(), n->stream);
        net_pack(n, n->stream);
        PF_HIP(hipStreamSynchronize(n->stream));
        return PFEM_OK;
    });
    if (rc != PFEM_OK) {
        pfem_net_destroy(n);
        return rc;
    }
    *out = n;
    return PFEM_OK;
}

void pfem_net_destroy(pfem_net *n) {
    if (!n) return;
    if (n->stream) (void)hipStreamSynchronize(n->stream);
    hipStream_t s = n->stream;
    delete n;  // dbuf destructors free device memory
    if (s) (void)hipStreamDestroy(s);
}

int64_t pfem_net_num_params(const pfem_net *n) { return n ? n->nparams : 0; }

int pfem_net_set_loss(pfem_net *n, int kind) {
    if (!n || (kind != 0 && kind != 1)) return PFEM_ERR_BAD_ARG;
    n->loss_kind = kind;
    return PFEM_OK;
}

int pfem_net_set_conv_kernel(pfem_net *n, int kind) {
    if (!n || (kind != 0 && kind != 1)) return PFEM_ERR_BAD_ARG;
    n->conv_kind = kind;
    return PFEM_OK;
}

int pfem_net_get_params(pfem_net *n, float *w) {
    if (!n || !w) return PFEM_ERR_BAD_ARG;
    return net_guard([&]() -> int {
        n->W.download(w, n->nparams, n->stream);
        PF_HIP(hipStreamSynchronize(n->stream));
        return PFEM_OK;
    });
}

int pfem_net_set_params(pfem_net *n, const float *w) {
    if (!n || !w) return PFEM_ERR_BAD_ARG;
    return net_guard([&]() -> int {
        n->W.upload(w, n->nparams, n->stream);
        net_pack(n, n->stream);
        n->M.zero(n->stream);
        n->V.zero(n->stream);
        n->adam_step = 0;
        PF_HIP(hipStreamSynchronize(n->stream));
        return PFEM_OK;
    });
}

static void upload_bf16(pfem_net *n, const float *h, int64_t count, uint16_t *dst) {
    std::vector<uint16_t> b(count);
    for (int64_t i = 0; i < count; ++i) b[i] = host_f2bf(h[i]);
    PF_HIP(hipMemcpyAsync(dst, b.data(), count * sizeof(uint16_t), hipMemcpyHostToDevice, n->stream));
    PF_HIP(hipStreamSynchronize(n->stream));
}

int pfem_net_forward(pfem_net *n, const float *in, float *out) {
    if (!n || !in || !out) return PFEM_ERR_BAD_ARG;
    return net_guard([&]() -> int {
        const int64_t m = 3 * n->nvox[0];
        dbuf<uint16_t> in3(m);
        upload_bf16(n, in, m, in3.p);
        launch_input_from_pairs(n->nvox[0], in3.p, n->x16.p, n->stream);
        net_forward(n, n->stream);
        std::vector<float> o16(n->nvox[0] * 16);
        n->out16.download(o16.data(), o16.size(), n->stream);
        PF_HIP(hipStreamSynchronize(n->stream));
        for (int64_t v = 0; v < n->nvox[0]; ++v)
            for (int c = 0; c < 3; ++c) out[3 * v + c] = o16[16 * v + c];
        return PFEM_OK;
    });
}

int pfem_net_loss_grad(pfem_net *n, const float *in, const float *target, double *loss, float *grad) {
    if (!n || !in || !target) return PFEM_ERR_BAD_ARG;
    return net_guard([&]() -> int {
        const int64_t m = 3 * n->nvox[0];
        dbuf<uint16_t> s2(2 * m);
        upload_bf16(n, in, m, s2.p);
        upload_bf16(n, target, m, s2.p + m);
        n->scal.zero(n->stream);
        net_sample(n, s2.p, s2.p + m, false, 1.f, n->stream);
        double h[8];
        n->scal.download(h, 8, n->stream);
        if (grad) n->G.download(grad, n->nparams, n->stream);
        PF_HIP(hipStreamSynchronize(n->stream));
        if (loss) *loss = h[3];
        return PFEM_OK;
    });
}

int pfem_net_add_pair(pfem_net *n, const float *r, const float *z) {
    if (!n || !r || !z || n->capacity <= 0) return PFEM_ERR_BAD_ARG;
    return net_guard([&]() -> int {
        const int64_t m = 3 * n->nvox[0];
        uint16_t *slot = n->pairs.p + n->next * 2 * m;
        upload_bf16(n, r, m, slot);
        upload_bf16(n, z, m, slot + m);
        n->next = (n->next + 1) % n->capacity;
        n->npairs = std::min(n->npairs + 1, n->capacity);
        return PFEM_OK;
    });
}

int pfem_net_num_pairs(const pfem_net *n, int64_t *count) {
    if (!n || !count) return PFEM_ERR_BAD_ARG;
    *count = n->npairs;
    return PFEM_OK;
}

int pfem_net_train(pfem_net *n, int steps, int batch, double lr, uint64_t seed, double *mean_loss) {
    if (!n || steps < 0 || batch < 1 || !(lr > 0)) return PFEM_ERR_BAD_ARG;
    if (n->npairs < 1) return PFEM_ERR_BAD_ARG;
    return net_guard([&]() -> int {
        hipStream_t s = n->stream;
        PF_HIP(hipDeviceSynchronize());  // samples may have been logged on a context's stream
        const int64_t m = 3 * n->nvox[0];
        n->scal.zero(s);
        uint64_t rs = seed;
        for (int it = 0; it < steps; ++it) {
            for (int b = 0; b < batch; ++b) {
                const int64_t idx = (int64_t)(splitmix64(rs) % (uint64_t)n->npairs);
                const uint16_t *p = n->pairs.p + idx * 2 * m;
                net_sample(n, p, p + m, b > 0, 1.f / batch, s);
            }
            launch_adam(n->nparams, n->W.p, n->G.p, n->M.p, n->V.p, (float)lr, 0.9f, 0.999f, 1e-8f, ++n->adam_step, s);
            net_pack(n, s);
        }
        double h[8];
        n->scal.download(h, 8, s);
        PF_HIP(hipStreamSynchronize(s));
        if (mean_loss) *mean_loss = steps > 0 ? h[3] / ((double)steps * batch) : 0.0;
        return PFEM_OK;
    });
}

int pfem_attach_net(pfem_ctx *c, pfem_net *n, int mode, int stride) {
    if (!c || mode < 0 || mode > 5 || stride < 1) return PFEM_ERR_BAD_ARG;
    const bool uses_net = mode != 0 && mode != 3;
    if (uses_net && (!n || n->nvox[0] != c->nnodes || n->device != c->device)) return PFEM_ERR_BAD_ARG;
    if ((mode == 4 || mode == 5) && c->opt.precision != PFEM_PREC_FP32_IR) return PFEM_ERR_BAD_ARG;
    if (mode == 4 && n->capacity <= 0) return PFEM_ERR_BAD_ARG;
    if (hipDeviceSynchronize() != hipSuccess) return PFEM_ERR_HIP;
    if (!uses_net) n = nullptr;
    if (n) n->npend = 0;
    c->net = n;
    c->net_mode = mode;
    c->net_stride = stride;
    c->net_calls = 0;
    return PFEM_OK;
}

}  // extern "C"
