// context.h -- the pfem_ctx object (internal).
#pragma once
#include <chrono>
#include <memory>

#include "common.h"
#include "kernels.h"

namespace pfem {
struct Amg;  // amg.cpp
void amg_destroy(Amg *);
}  // namespace pfem

struct pfem_ctx {
    hipStream_t stream = nullptr;
    int device = 0;
    pfem_options opt{};
    int64_t nnodes = 0, nelem = 0, ndof = 0;
    int kind = 1;       // pfem_element_kind
    int fbar = 1;       // kind == PFEM_HEX8_FINITE_FBAR
    int kin = 0;        // some material has Hkin > 0 (finite strain: objective back stress)

    // mesh
    pfem::dbuf<double> X;
    pfem::dbuf<int64_t> conn;
    pfem::dbuf<int32_t> matid;
    pfem::dbuf<pfem::DevMat> mats;
    std::vector<double> h_nodes;
    int built_on_device = 0;       // bit 0: pattern and slots, bit 1: colours
    double t_create_pattern = 0;   // seconds spent on pattern, slots and colours

    // pattern, colours, scatter slots
    pfem::FineMatrix K;
    pfem::dbuf<int32_t> slots;  // 64 per element: slot of block (a, b)
    pfem::dbuf<uint64_t> first_touch;  // per element: blocks that open their slot in colour order
    int ncolors = 0;
    std::vector<int64_t> color_ptr;
    std::vector<int32_t> h_color;
    pfem::dbuf<int64_t> color_order;

    // Gauss state: committed history (epsp, abar, cpinv, beta) and stress (sig_c), the
    // stress of the last assembly (sig: the trial state's sigma), element scratch
    pfem::dbuf<double> epsp, abar, cpinv, sig, sig_c;
    pfem::dbuf<double> beta;   // back stress; allocated for small strain or Hkin > 0 only
    pfem::dbuf<double> packet, g0;
    int64_t packet_elems = 0;  // elements per packet chunk (pipelined assembly), else 0
    pfem::dbuf<double> geo;    // per Gauss point: reference gradients dN/dX (24)
    bool geo_ready = false;    // geo written (by the first full assembly)
    pfem::dbuf<unsigned long long> fail_key;

    // vectors (U_c: U at the last commit, restored by pfem_rollback)
    pfem::dbuf<double> U, R, dU, rhs, Fext0, dg, U_c;
    bool K_valid = true;   // false after a failed assembly (K holds a partial or stale tangent)
    double fext0_norm = 0.0;

    // boundary conditions
    int64_t nbc = 0, nbcrows = 0;
    pfem::dbuf<int64_t> bcdofs;
    pfem::dbuf<uint8_t> bcflag;
    pfem::dbuf<double> bcval;
    pfem::dbuf<int32_t> bcrows;

    // Krylov workspace (FP64 and FP32)
    pfem::dbuf<double> w_r, w_z, w_p, w_q, w_x, w_s, w_t, w_v, w_rh;
    pfem::dbuf<float> f_r, f_z, f_p, f_q, f_x, f_b, f_rh, f_v, f_s, f_t;
    pfem::dbuf<double> partial, scal;
    pfem::dbuf<double> mpartial;      // partials of the fused multi-dot (flexible GCR)
    double *h_scal = nullptr;  // pinned

    // preconditioner
    std::unique_ptr<pfem::Amg, void (*)(pfem::Amg *)> amg{nullptr, pfem::amg_destroy};
    bool pc_ready = false;

    // learned preconditioner (pfem_attach_net)
    pfem_net *net = nullptr;
    int net_mode = 0, net_stride = 1;
    int64_t net_calls = 0;
    int pc_stale = 0;       // AMG rebuild requested while the network preconditions (1 full, 3 keep P)
    bool refresh_P = false;  // AMG policy 6: the last solve fell back, rebuild the prolongators
    // AMG policy 7: stash slot of the current Newton iteration (-1: none); with the
    // learned preconditioner, whether its AMG phase restored (4) or kept (5) coarse
    // levels and that phase's Krylov iterations
    int stash_slot = -1;
    int stash_used = 0, amg_phase_iters = 0;
    pfem::dbuf<double> net_dinv;  // block-diagonal inverse of K for the learned preconditioner
    bool half_fine = false;       // FP32 V-cycle: fine-level products with the FP16 scaled K
    bool split_op = false;        // FP32-IR flexible CG: K p from the FP32 + BF16 split copy of K
    pfem::dbuf<double> fcg_P, fcg_Q; // truncated flexible CG: search directions and K times them

    // statistics
    pfem_stats st{};
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // Krylov loops with the next product launched before the stopping test: two
    // (start, end) pairs timing the products of consecutive iterations, and the event
    // after each iteration's scalar copy to the host
    hipEvent_t ev_kry[4] = {}, ev_scal = nullptr;
    // pipelined assembly (assembly_chunk > 0): the scatter's stream and the events that
    // order it against the Gauss-point kernels (start, gp[2], scatter[2], join)
    hipStream_t stream2 = nullptr;
    hipEvent_t ev_pipe[6] = {};

    double sync_read(const double *dptr);  // D2H of one device scalar
    void sync() { PF_HIP(hipStreamSynchronize(stream)); }
};

namespace pfem {
inline double now() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}
int assemble_impl(pfem_ctx *c, int commit, pfem_fail *fail, bool state_only = false);
int linear_solve_impl(pfem_ctx *c, double rtol, int rebuild, int *niter, int *used, double *relres,
                      bool *discarded = nullptr);
// Dirichlet imposition on the device K and R for the constrained set `sel` (bit 0: the
// fixed set, bit 1: the ramped set) at load factor lambda, against the device U.
void impose_set(pfem_ctx *c, int sel, double lambda);
}  // namespace pfem
