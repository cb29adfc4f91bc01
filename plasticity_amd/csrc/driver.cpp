// driver.cpp -- assembly driver, Krylov solves and the Newton loop.
//
//  assemble_impl      assemble! (Assembly.jl:238-257, :361-385): zero K and R, one
//                     per-GP launch over all elements, then the colour-ordered
//                     scatter launches (colours serially, as Assembly.jl:378-383).
//  linear_solve_impl  _linear_solve! (Solver.jl:235-290): PCG with the AMG V-cycle
//                     (stopping on the preconditioned residual sqrt(r'z) relative to
//                     its initial value, atol = 0), refresh-and-retry, then a
//                     BiCGStab fallback with the same preconditioner in place of the
//                     reference's direct solve (a 3M-DOF LU is not an option on the
//                     device); or, in FP32-IR mode, FP64 flexible PCG whose
//                     preconditioner is the FP32 V-cycle (each iteration refines the
//                     FP64 solution with an FP32 correction), then FP32 Krylov inside
//                     FP64 iterative refinement, then the FP64 ladder.
//  pfem_newton        newton! (Solver.jl:317-389) with the Eisenstat-Walker forcing
//                     term (:159-171) and the reference-scaled convergence test.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <type_traits>

#include "amg.h"
#include "context.h"
#include "unet.h"

using namespace pfem;

namespace pfem {

static ElementArgs element_args(pfem_ctx *c, int commit) {
    ElementArgs a{};
    a.nelem = c->nelem;
    a.X = c->X.p;
    a.conn = c->conn.p;
    a.U = c->U.p;
    a.matid = c->matid.p;
    a.mats = c->mats.p;
    a.kind = c->kind;
    a.fbar = c->fbar;
    a.kin = c->kin;
    a.commit = commit;
    a.epsp = c->epsp.p;
    a.beta = c->beta.p;
    a.abar = c->abar.p;
    a.cpinv = c->cpinv.p;
    a.sig = c->sig.p;
    a.packet = c->packet.p;
    a.g0 = c->g0.p;
    a.fail_key = c->fail_key.p;
    a.geo = c->geo.p;
    return a;
}

int assemble_impl(pfem_ctx *c, int commit, pfem_fail *fail, bool state_only) {
    hipStream_t s = c->stream;
    const double t0 = now();
    ElementArgs a = element_args(c, commit);
    // a commit writes the committed stress; a trial assembly the trial state's sigma
    if (commit) a.sig = c->sig_c.p;
    if (state_only) a.packet = nullptr;  // stress and history only: no tangent, no K or R
    else c->K_valid = false;             // until the scatter has run
    a.geo_write = !state_only && !c->geo_ready;  // the reference gradients, once per context
    PF_HIP(hipMemsetAsync(c->fail_key.p, 0xff, sizeof(unsigned long long), s));
    // Assembly.jl:361-362 zeroes K and R: R here, K through the scatter's first-touch
    // stores
    if (!state_only) c->R.zero(s);
    const bool pipelined = c->packet_elems > 0 && !state_only;
    PF_HIP(hipEventRecord(c->ev0, s));
    if (pipelined) {
        // Chunks of at most packet_elems elements of one colour, in colour order.  Chunk k:
        // Gauss-point kernel on `s` into buffer k % 2 (after the scatter of chunk k - 2 has
        // read it), then its scatter on `stream2` (after that Gauss-point kernel).  The
        // scatters run in colour order on one stream, so first-touch stores still precede
        // the adds into each block.
        hipStream_t s2 = c->stream2;
        hipEvent_t ev_start = c->ev_pipe[0], ev_join = c->ev_pipe[5];
        hipEvent_t ev_gp[2] = {c->ev_pipe[1], c->ev_pipe[2]}, ev_sc[2] = {c->ev_pipe[3], c->ev_pipe[4]};
        PF_HIP(hipEventRecord(ev_start, s));  // everything before (R cleared, solves reading K)
        PF_HIP(hipStreamWaitEvent(s2, ev_start, 0));
        const int64_t chunk = c->packet_elems;
        a.elist = c->color_order.p;
        int64_t k = 0;
        for (int col = 0; col < c->ncolors; ++col)
            for (int64_t b = c->color_ptr[col]; b < c->color_ptr[col + 1]; b += chunk, ++k) {
                const int64_t n = std::min(chunk, c->color_ptr[col + 1] - b);
                const int buf = (int)(k & 1);
                if (k >= 2) PF_HIP(hipStreamWaitEvent(s, ev_sc[buf], 0));
                a.packet = c->packet.p + (size_t)buf * chunk * 8 * pk_len(packet_sym(a));
                launch_gp_kernel(a, b, n, s);
                PF_HIP(hipEventRecord(ev_gp[buf], s));
                PF_HIP(hipStreamWaitEvent(s2, ev_gp[buf], 0));
                launch_scatter_kernel(a, c->color_order.p + b, n, c->slots.p, c->first_touch.p, c->K.v64.p, c->R.p,
                                      s2, true);
                PF_HIP(hipEventRecord(ev_sc[buf], s2));
            }
        PF_HIP(hipEventRecord(ev_join, s2));
        PF_HIP(hipStreamWaitEvent(s, ev_join, 0));
    } else {
        launch_gp_kernel(a, 0, c->nelem, s);
    }
    PF_HIP(hipEventRecord(c->ev1, s));
    unsigned long long key = 0;
    PF_HIP(hipMemcpyAsync(&key, c->fail_key.p, sizeof key, hipMemcpyDeviceToHost, s));
    c->sync();
    if (a.geo_write) c->geo_ready = true;  // every Gauss point writes it, inverted or not
    float ms_gp = 0.f;
    PF_HIP(hipEventElapsedTime(&ms_gp, c->ev0, c->ev1));
    // per full assembly (n_assemble); pipelined: the span of both kernels together
    if (!state_only) c->st.t_gp_kernel += 1e-3 * ms_gp;
    if (key != ~0ull) {  // ElementInversionError(e, g, J)
        if (pipelined) c->R.zero(s);  // the scatters ran: R is returned cleared as without the pipeline
        launch_detF_kernel(a, (int64_t)key, c->scal.p, s);
        double J = c->sync_read(c->scal.p);
        if (fail) { fail->e = (int64_t)(key / 8); fail->g = (int32_t)(key % 8); fail->J = J; }
        return PFEM_ERR_INVERTED;
    }
    if (commit) {  // the committed displacement, restored by pfem_rollback
        PF_HIP(hipMemcpyAsync(c->U_c.p, c->U.p, c->ndof * sizeof(double), hipMemcpyDeviceToDevice, s));
        if (!state_only)  // this assembly's stress is also the trial stress
            PF_HIP(hipMemcpyAsync(c->sig.p, c->sig_c.p, c->sig.bytes(), hipMemcpyDeviceToDevice, s));
    }
    if (state_only) {
        c->sync();
        return PFEM_OK;
    }
    if (!pipelined) {
        PF_HIP(hipEventRecord(c->ev0, s));
        for (int col = 0; col < c->ncolors; ++col)
            launch_scatter_kernel(a, c->color_order.p + c->color_ptr[col],
                                  c->color_ptr[col + 1] - c->color_ptr[col], c->slots.p, c->first_touch.p,
                                  c->K.v64.p, c->R.p, s);
        PF_HIP(hipEventRecord(c->ev1, s));
        c->sync();
        float ms_sc = 0.f;
        PF_HIP(hipEventElapsedTime(&ms_sc, c->ev0, c->ev1));
        c->st.t_scatter_kernel += 1e-3 * ms_sc;
    }
    c->st.t_assemble += now() - t0;
    c->st.n_assemble++;
    c->K_valid = true;
    return PFEM_OK;
}

void impose_set(pfem_ctx *c, int sel, double lambda) {
    impose_dirichlet_dev(c->K, c->nbc, c->bcdofs.p, c->bcflag.p, c->bcval.p, lambda, c->U.p, c->dg.p, c->nbcrows,
                         c->bcrows.p, c->R.p, sel, c->stream);
}

// ---------------------------------------------------------------- preconditioner
static void pc_setup(pfem_ctx *c, bool keep_P = false) {
    const double t0 = now();
    if (!c->amg) c->amg.reset(amg_create(c));
    amg_numeric_setup(c, c->amg.get(), keep_P && c->pc_ready);
    c->pc_ready = true;
    c->st.t_setup += now() - t0;
    c->st.n_setup++;
}

// Precision traits: fine operator, V-cycle and workspace for FP64 / FP32 vectors.
template <class T> struct Prec;
template <> struct Prec<double> {
    static void spmv(pfem_ctx *c, const double *x, double *y) { spmv_fine64(c->K, x, y, c->stream); }
    static void spmv_dot(pfem_ctx *c, const double *x, double *y, double *out) {
        spmv_fine64_dot(c->K, x, y, c->partial.p, out, c->stream);
    }
    static void pc(pfem_ctx *c, const double *r, double *z) { amg_apply64(c, c->amg.get(), r, z); }
    static double *w(pfem_ctx *c, int k) {
        pfem::dbuf<double> *b[] = {&c->w_r, &c->w_z, &c->w_p, &c->w_q, &c->w_rh, &c->w_v, &c->w_s, &c->w_t};
        return b[k]->p;
    }
};
template <> struct Prec<float> {
    static void spmv(pfem_ctx *c, const float *x, float *y) { spmv_fine32(c->K, x, y, c->stream); }
    static void spmv_dot(pfem_ctx *c, const float *x, float *y, double *out) {
        spmv_fine32_dot(c->K, x, y, c->partial.p, out, c->stream);
    }
    static void pc(pfem_ctx *c, const float *r, float *z) { amg_apply32(c, c->amg.get(), r, z); }
    static float *w(pfem_ctx *c, int k) {
        pfem::dbuf<float> *b[] = {&c->f_r, &c->f_z, &c->f_p, &c->f_q, &c->f_rh, &c->f_v, &c->f_s, &c->f_t};
        return b[k]->p;
    }
};

// PCG (Krylov.jl cg! semantics as used at Solver.jl:258: x0 = 0, stop when
// sqrt(r'z) <= rtol * sqrt(r0'z0)).  Dot products accumulate in FP64.  Returns
// 0 solved, 1 not converged (itmax), 2 breakdown (p'Kp <= 0 or r'z < 0: the
// operator or the preconditioner is not positive definite).
template <class T>
static int pcg(pfem_ctx *c, const T *b, T *x, double rtol, int itmax, int *iters) {
    hipStream_t s = c->stream;
    const int64_t n = c->ndof;
    T *r = Prec<T>::w(c, 0), *z = Prec<T>::w(c, 1), *p = Prec<T>::w(c, 2), *q = Prec<T>::w(c, 3);
    double *sc = c->scal.p;
    fill<T>(n, T(0), x, s);
    PF_HIP(hipMemcpyAsync(r, b, n * sizeof(T), hipMemcpyDeviceToDevice, s));
    Prec<T>::pc(c, r, z);
    dot_dev<T, T>(n, r, z, c->partial.p, sc + 0, s);
    double rz = c->sync_read(sc + 0);
    *iters = 0;
    if (rz < 0) return 2;
    const double rn0 = std::sqrt(rz);
    if (rn0 == 0.0) return 0;
    const double eps = rtol * rn0;
    PF_HIP(hipMemcpyAsync(p, z, n * sizeof(T), hipMemcpyDeviceToDevice, s));
    // q = K p with sc[1] = p'q for iteration k, timed by the event pair of k's parity
    auto product = [&](int k) {
        PF_HIP(hipEventRecord(c->ev_kry[2 * (k & 1)], s));
        Prec<T>::spmv_dot(c, p, q, sc + 1);
        PF_HIP(hipEventRecord(c->ev_kry[2 * (k & 1) + 1], s));
    };
    // p = z + (rzn/rz) p, then the next product
    auto next_direction = [&](int k) {
        xpby_ratio<T>(n, sc + 2, sc + 0, z, p, s);
        PF_HIP(hipMemcpyAsync(sc + 0, sc + 2, sizeof(double), hipMemcpyDeviceToDevice, s));
        product(k);
    };
    product(1);
    double nrm_prev = rn0, nrm_prev2 = 0.0;
    for (int it = 1; it <= itmax; ++it) {
        // p'q is read with r'z, one host round trip per iteration: after a breakdown
        // (p'q <= 0) this iteration's update is discarded with the solve (every caller's
        // next solver restarts from x = 0)
        cg_update<T>(n, sc + 0, sc + 1, p, q, x, r, s);  // x += a p, r -= a q
        Prec<T>::pc(c, r, z);
        dot_dev<T, T>(n, r, z, c->partial.p, sc + 2, s);
        PF_HIP(hipMemcpyAsync(c->h_scal + 1, sc + 1, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
        // Unless this iteration is expected to meet the stopping test (the last rate
        // carried forward), the next direction and its product are queued before the
        // host waits for the scalars: x and r stay as they are if it stops here.
        const bool ahead = !c->opt.krylov_sync && it < itmax &&
                           !(nrm_prev2 > 0.0 && nrm_prev * (nrm_prev / nrm_prev2) <= eps);
        if (ahead) {
            PF_HIP(hipEventRecord(c->ev_scal, s));
            next_direction(it + 1);
            PF_HIP(hipEventSynchronize(c->ev_scal));
        } else {
            c->sync();
        }
        const double pq = c->h_scal[1], rzn = c->h_scal[2];
        if (sizeof(T) == 8) {
            float ms = 0.f;
            PF_HIP(hipEventElapsedTime(&ms, c->ev_kry[2 * (it & 1)], c->ev_kry[2 * (it & 1) + 1]));
            c->st.t_pcg_spmv += 1e-3 * ms;
            c->st.n_pcg_spmv++;
        }
        *iters = it;
        if (!(pq > 0)) {
            if (c->opt.verbose > 1) fprintf(stderr, "      pcg breakdown p'Kp = %.3e at it %d\n", pq, it);
            return 2;
        }
        if (!(rzn >= 0)) {
            if (c->opt.verbose > 1) fprintf(stderr, "      pcg breakdown r'z = %.3e at it %d\n", rzn, it);
            return 2;
        }
        if (c->opt.verbose > 2) fprintf(stderr, "      pcg %3d  sqrt(rz) = %.3e  (target %.3e)\n", it, std::sqrt(rzn), eps);
        if (std::sqrt(rzn) <= eps) return 0;
        if (!ahead && it < itmax) next_direction(it + 1);
        nrm_prev2 = nrm_prev;
        nrm_prev = std::sqrt(rzn);
        rz = rzn;
    }
    return 1;
}

// Right-preconditioned BiCGStab (van der Vorst, SIAM J. Sci. Stat. Comput. 13, 1992)
// with the same V-cycle; stops on ||b - Kx|| <= rtol ||b||.  Handles the indefinite
// tangents on which PCG breaks down.
template <class T>
static int bicgstab(pfem_ctx *c, const T *b, T *x, double rtol, int itmax, int *iters) {
    hipStream_t s = c->stream;
    const int64_t n = c->ndof;
    T *r = Prec<T>::w(c, 0), *ph = Prec<T>::w(c, 1), *p = Prec<T>::w(c, 2), *sh = Prec<T>::w(c, 3),
      *rh = Prec<T>::w(c, 4), *v = Prec<T>::w(c, 5), *sv = Prec<T>::w(c, 6), *t = Prec<T>::w(c, 7);
    double *sc = c->scal.p;
    auto dot = [&](const T *a, const T *bb) {
        dot_dev<T, T>(n, a, bb, c->partial.p, sc + 8, s);
        return c->sync_read(sc + 8);
    };
    fill<T>(n, T(0), x, s);
    PF_HIP(hipMemcpyAsync(r, b, n * sizeof(T), hipMemcpyDeviceToDevice, s));
    PF_HIP(hipMemcpyAsync(rh, b, n * sizeof(T), hipMemcpyDeviceToDevice, s));
    fill<T>(n, T(0), p, s);
    fill<T>(n, T(0), v, s);
    const double bn = std::sqrt(dot(b, b));
    *iters = 0;
    if (bn == 0.0) return 0;
    double rho = 1, alpha = 1, omega = 1;
    for (int it = 1; it <= itmax; ++it) {
        *iters = it;
        const double rho_n = dot(rh, r);
        if (rho_n == 0.0) return 2;
        const double beta = (rho_n / rho) * (alpha / omega);
        axpby<T>(n, -omega, v, 1.0, p, s);  // p = p - w v
        axpby<T>(n, 1.0, r, beta, p, s);    // p = r + beta p
        Prec<T>::pc(c, p, ph);
        Prec<T>::spmv(c, ph, v);
        const double rv = dot(rh, v);
        if (rv == 0.0) return 2;
        alpha = rho_n / rv;
        PF_HIP(hipMemcpyAsync(sv, r, n * sizeof(T), hipMemcpyDeviceToDevice, s));
        axpby<T>(n, -alpha, v, 1.0, sv, s);  // s = r - a v
        axpby<T>(n, alpha, ph, 1.0, x, s);   // x += a ph
        const double sn = std::sqrt(dot(sv, sv));
        if (sn <= rtol * bn) return 0;
        Prec<T>::pc(c, sv, sh);
        Prec<T>::spmv(c, sh, t);
        const double tt = dot(t, t);
        if (tt == 0.0) return 2;
        omega = dot(t, sv) / tt;
        axpby<T>(n, omega, sh, 1.0, x, s);  // x += w sh
        PF_HIP(hipMemcpyAsync(r, sv, n * sizeof(T), hipMemcpyDeviceToDevice, s));
        axpby<T>(n, -omega, t, 1.0, r, s);  // r = s - w t
        const double rn = std::sqrt(dot(r, r));
        if (c->opt.verbose > 2) fprintf(stderr, "      bicgstab %3d  |r|/|b| = %.3e\n", it, rn / bn);
        if (rn <= rtol * bn) return 0;
        if (omega == 0.0) return 2;
        rho = rho_n;
    }
    return 1;
}

// Flexible PCG (FCG(1), Notay, SIAM J. Sci. Comput. 22, 2000) in FP64 with any
// preconditioner z = M(r): beta = -(z_new' q)/(p' q), so a preconditioner that varies
// from one application to the next (an FP32 V-cycle, a network) keeps the iteration
// convergent.  Stop test as pcg() (sqrt(r'z) relative to its initial value) when
// `pc_norm`, else ||r|| <= rtol ||b|| (preconditioners that need not be SPD).
// 0 solved, 1 itmax, 2 breakdown; stalls (no residual decrease over `stall` iterations)
// return 3 so that the caller can fall back.
struct Amg32Pc;
template <class PC>
static int fpcg(pfem_ctx *c, const double *b, double *x, double rtol, int itmax, bool pc_norm, int stall,
                int *iters, PC &&pcfun) {
    hipStream_t s = c->stream;
    const int64_t n = c->ndof;
    double *r = c->w_r.p, *z = c->w_z.p, *p = c->w_p.p, *q = c->w_q.p, *sc = c->scal.p;
    // the FP32 V-cycle: the update writes its FP32 input, the dot pass widens its output
    constexpr bool amg32 = std::is_same<std::decay_t<PC>, Amg32Pc>::value;
    if (c->mpartial.n < multi_dot_partials()) c->mpartial.alloc(multi_dot_partials());
    fill<double>(n, 0.0, x, s);
    PF_HIP(hipMemcpyAsync(r, b, n * sizeof(double), hipMemcpyDeviceToDevice, s));
    pcfun(r, z);
    dot_dev<double, double>(n, r, z, c->partial.p, sc + 0, s);
    dot_dev<double, double>(n, r, r, c->partial.p, sc + 3, s);
    PF_HIP(hipMemcpyAsync(c->h_scal, sc, 4 * sizeof(double), hipMemcpyDeviceToHost, s));
    c->sync();
    double rz = c->h_scal[0], rr = c->h_scal[3];
    *iters = 0;
    if (rr == 0.0) return 0;
    if (pc_norm && !(rz > 0)) return 2;
    const double eps = pc_norm ? rtol * std::sqrt(rz) : rtol * std::sqrt(rr);
    double best = std::sqrt(rr);
    int best_it = 0;
    PF_HIP(hipMemcpyAsync(p, z, n * sizeof(double), hipMemcpyDeviceToDevice, s));
    // q = K p with sc[1] = p'q for iteration k, timed by the event pair of k's parity
    // (FP32-IR with split_operator: from the FP32 + BF16 split copy of K)
    auto product = [&](int k) {
        PF_HIP(hipEventRecord(c->ev_kry[2 * (k & 1)], s));
        if (c->split_op) spmv_fine_split_dot(c->K, p, q, c->partial.p, sc + 1, s);
        else Prec<double>::spmv_dot(c, p, q, sc + 1);
        PF_HIP(hipEventRecord(c->ev_kry[2 * (k & 1) + 1], s));
    };
    product(1);
    double nrm_prev = pc_norm ? std::sqrt(rz) : std::sqrt(rr), nrm_prev2 = 0.0;
    for (int it = 1; it <= itmax; ++it) {
        *iters = it;
        // x += a p, r -= a q; z = M(r); sc[2] = r'z, sc[3] = r'r, sc[4] = z'q
        if constexpr (amg32) {
            cg_update_r32(n, sc + 0, sc + 1, p, q, x, r, c->f_r.p, s);
            amg_apply32(c, c->amg.get(), c->f_r.p, c->f_z.p);
            fcg_dots<float>(n, r, c->f_z.p, z, q, c->mpartial.p, sc + 2, s);
        } else {
            cg_update<double>(n, sc + 0, sc + 1, p, q, x, r, s);
            pcfun(r, z);
            fcg_dots<double>(n, r, z, nullptr, q, c->mpartial.p, sc + 2, s);
        }
        PF_HIP(hipMemcpyAsync(c->h_scal, sc, 5 * sizeof(double), hipMemcpyDeviceToHost, s));
        // Unless this iteration is expected to meet the stopping test (the last rate
        // carried forward), the next direction p = z + beta p, beta = -(z'q)/(p'q), and
        // its product are queued before the host waits for the scalars: x and r stay
        // as they are if the iteration stops here.
        const bool ahead = !c->opt.krylov_sync && it < itmax &&
                           !(nrm_prev2 > 0.0 && nrm_prev * (nrm_prev / nrm_prev2) <= eps);
        if (ahead) {
            PF_HIP(hipEventRecord(c->ev_scal, s));
            axpby_negratio<double>(n, 1.0, z, sc + 4, sc + 1, p, s);
            PF_HIP(hipMemcpyAsync(sc + 0, sc + 2, sizeof(double), hipMemcpyDeviceToDevice, s));
            product(it + 1);
            PF_HIP(hipEventSynchronize(c->ev_scal));
        } else {
            c->sync();
        }
        const double pq = c->h_scal[1], rzn = c->h_scal[2], rrn = c->h_scal[3];
        float ms = 0.f;
        PF_HIP(hipEventElapsedTime(&ms, c->ev_kry[2 * (it & 1)], c->ev_kry[2 * (it & 1) + 1]));
        c->st.t_pcg_spmv += 1e-3 * ms;
        c->st.n_pcg_spmv++;
        if (!(pq > 0) || (pc_norm && !(rzn >= 0))) {
            if (c->opt.verbose > 1) fprintf(stderr, "      fpcg breakdown p'Kp = %.3e r'z = %.3e at it %d\n", pq, rzn, it);
            return 2;
        }
        const double nrm = pc_norm ? std::sqrt(rzn) : std::sqrt(rrn);
        if (c->opt.verbose > 2) fprintf(stderr, "      fpcg %3d  %.3e  (target %.3e)\n", it, nrm, eps);
        if (nrm <= eps) return 0;
        if (std::sqrt(rrn) < 0.999 * best) { best = std::sqrt(rrn); best_it = it; }
        else if (stall > 0 && it - best_it >= stall) return 3;
        if (!ahead && it < itmax) {
            // p = z + beta p, beta = -(z'q)/(p'q)
            const double beta = -c->h_scal[4] / pq;
            axpby<double>(n, 1.0, z, beta, p, s);
            PF_HIP(hipMemcpyAsync(sc + 0, sc + 2, sizeof(double), hipMemcpyDeviceToDevice, s));
            product(it + 1);
        }
        nrm_prev2 = nrm_prev;
        nrm_prev = nrm;
        rz = rzn;
    }
    return 1;
}

// Flexible truncated GCR(m) (Eisenstat, Elman & Schultz, SIAM J. Numer. Anal. 20,
// 1983; flexible form: Vuik, J. Comput. Appl. Math. 61, 1995): with z = M(r) and
// q = K z, the pair (p, q) is orthogonalised against the last m pairs in the q'q inner
// product, then alpha = r'q / q'q minimises ||r|| along p.  It needs neither a
// symmetric K (the F-bar tangent is not symmetric) nor a linear preconditioner (a
// network), and costs one SpMV per iteration.  Stops on ||r|| <= rtol ||b||;
// 0 solved, 1 itmax, 2 breakdown (q = 0), 3 stall: no 0.1% gain of the best residual
// over `stall` iterations, or (slow_window > 0) the last slow_window iterations
// together reduced ||r|| by less than slow_rate^slow_window.
template <class PC>
static int gcr_m(pfem_ctx *c, const double *b, double *x, double rtol, int itmax, int m, int stall, int *iters,
                 PC &&pcfun, int slow_window = 0, double slow_rate = 1.0) {
    hipStream_t s = c->stream;
    const int64_t n = c->ndof;
    const int slots = m + 1;
    if (c->fcg_P.n < (size_t)slots * n) {
        c->fcg_P.alloc((size_t)slots * n);
        c->fcg_Q.alloc((size_t)slots * n);
    }
    double *r = c->w_r.p, *sc = c->scal.p;
    fill<double>(n, 0.0, x, s);
    PF_HIP(hipMemcpyAsync(r, b, n * sizeof(double), hipMemcpyDeviceToDevice, s));
    dot_dev<double, double>(n, r, r, c->partial.p, sc + 3, s);
    const double bn = std::sqrt(c->sync_read(sc + 3));
    *iters = 0;
    if (bn == 0.0) return 0;
    const double eps = rtol * bn;
    std::vector<double> qq(slots, 0.0), beta(m, 0.0), hist(1, bn);
    std::vector<const double *> vp(m, nullptr), vq(m, nullptr);
    if (c->mpartial.n < multi_dot_partials()) c->mpartial.alloc(multi_dot_partials());
    double best = bn;
    int best_it = 0;
    for (int it = 1; it <= itmax; ++it) {
        *iters = it;
        const int k = (it - 1) % slots;
        const int nold = std::min(it - 1, m);
        double *pk = c->fcg_P.p + (size_t)k * n, *qk = c->fcg_Q.p + (size_t)k * n;
        pcfun(r, pk);
        PF_HIP(hipEventRecord(c->ev0, s));
        spmv_fine64(c->K, pk, qk, s);  // q = K z
        PF_HIP(hipEventRecord(c->ev1, s));
        // classical Gram-Schmidt against the kept directions: all (K z)' (K p_j) in one
        // pass, then p -= sum beta_j p_j and K p -= sum beta_j K p_j in one pass
        for (int j = 1; j <= nold; ++j) {
            const int sj = (k - j + slots) % slots;
            vq[j - 1] = c->fcg_Q.p + (size_t)sj * n;
            vp[j - 1] = c->fcg_P.p + (size_t)sj * n;
        }
        if (nold > 0) {
            multi_dot(n, qk, vq.data(), nold, c->mpartial.p, sc + 17, s);
            PF_HIP(hipMemcpyAsync(c->h_scal + 17, sc + 17, nold * sizeof(double), hipMemcpyDeviceToHost, s));
            c->sync();
            for (int j = 1; j <= nold; ++j) beta[j - 1] = c->h_scal[16 + j] / qq[(k - j + slots) % slots];
            multi_axpy2(n, vp.data(), vq.data(), beta.data(), nold, pk, qk, s);
        }
        const double *qr[2] = {qk, r};
        multi_dot(n, qk, qr, 2, c->mpartial.p, sc + 1, s);  // sc[1] = q'q, sc[2] = q'r
        PF_HIP(hipMemcpyAsync(c->h_scal, sc, 4 * sizeof(double), hipMemcpyDeviceToHost, s));
        c->sync();
        float ms = 0.f;
        PF_HIP(hipEventElapsedTime(&ms, c->ev0, c->ev1));
        c->st.t_pcg_spmv += 1e-3 * ms;
        c->st.n_pcg_spmv++;
        const double qqk = c->h_scal[1], rq = c->h_scal[2];
        if (!(qqk > 0)) return 2;
        qq[k] = qqk;
        const double alpha = rq / qqk;
        update_dot(n, alpha, pk, qk, x, r, c->partial.p, sc + 3, s);  // x += alpha p, r -= alpha K p
        const double rn = std::sqrt(c->sync_read(sc + 3));
        if (c->opt.verbose > 2) fprintf(stderr, "      gcr(%d) %3d  %.3e  (target %.3e)\n", m, it, rn, eps);
        if (rn <= eps) return 0;
        hist.push_back(rn);
        if (slow_window > 0 && it >= slow_window && rn > std::pow(slow_rate, slow_window) * hist[it - slow_window])
            return 3;
        if (rn < 0.999 * best) { best = rn; best_it = it; }
        else if (stall > 0 && it - best_it >= stall) return 3;
    }
    return 1;
}

// The FP32 V-cycle as a preconditioner of FP64 vectors: r -> FP32, V-cycle on the FP32
// copies of every level, z -> FP64.
struct Amg32Pc {
    pfem_ctx *c;
    void operator()(const double *r, double *z) const {
        convert<double, float>(c->ndof, r, c->f_r.p, c->stream);
        amg_apply32(c, c->amg.get(), c->f_r.p, c->f_z.p);
        convert<float, double>(c->ndof, c->f_z.p, z, c->stream);
    }
};

struct Amg64Pc {
    pfem_ctx *c;
    void operator()(const double *r, double *z) const { amg_apply64(c, c->amg.get(), r, z); }
};
// The FP32 V-cycle followed by the network's correction of its output (attach mode 5):
// z = V(r), then z += c N(z / c) with c = max |z| -- the network sees the V-cycle's
// approximation of the error and adds what it predicts the cycle left out.
struct Amg32NetPc {
    pfem_ctx *c;
    void operator()(const double *r, double *z) const {
        Amg32Pc{c}(r, z);
        net_correct(c->net, z, c->bcflag.p, c->stream);
    }
};
// The FP32 V-cycle, with every `net_stride`-th application reserving an error sample
// (attach mode 4): input z_k = V(r_k) at the iterate x_k; the target x* - x_k - z_k is
// written when the solve has converged (net_log_error_finish).
struct Amg32ErrLogPc {
    pfem_ctx *c;
    const double *x;
    void operator()(const double *r, double *z) const {
        Amg32Pc{c}(r, z);
        if (c->net_calls++ % c->net_stride == 0) net_log_error_begin(c->net, z, x, c->bcflag.p, c->stream);
    }
};
// The learned preconditioner, a two-level method whose coarse correction is the
// network: n = N(r) (constrained DOFs 0), scaled by g = (Kn)'r / (Kn)'(Kn) (0 if
// Kn = 0), then one block-Jacobi sweep on what is left:
//   z = g n + D^-1 (r - g K n).
// Without a network (attach mode 3) it is block-Jacobi alone, z = D^-1 r: the same
// method without its learned part.
struct LearnedPc {
    pfem_ctx *c;
    void operator()(const double *r, double *z) const {
        hipStream_t s = c->stream;
        const int64_t n = c->ndof;
        double *sc = c->scal.p;
        if (!c->net) {
            block_jacobi3<double>(c->nnodes, c->net_dinv.p, r, z, s);
            return;
        }
        // g minimises ||r - g K n|| (g = (Kn)'r / (Kn)'(Kn)), so the correction never
        // increases the residual the Jacobi sweep then works on, whatever the network
        // returns; the Galerkin choice r'n / n'Kn blew up when n'Kn was small.
        double *nv = c->w_t.p, *q = c->w_v.p, *r1 = c->w_rh.p;
        net_apply(c->net, r, c->bcflag.p, nv, s);
        spmv_fine64(c->K, nv, q, s);                                 // q = K n
        dot_dev<double, double>(n, q, q, c->partial.p, sc + 10, s);  // sc[10] = q'q
        dot_dev<double, double>(n, q, r, c->partial.p, sc + 11, s);  // sc[11] = q'r
        PF_HIP(hipMemcpyAsync(r1, r, n * sizeof(double), hipMemcpyDeviceToDevice, s));
        axpy_safe_ratio(n, sc + 11, sc + 10, -1.0, q, r1, s);  // r1 = r - g K n
        block_jacobi3<double>(c->nnodes, c->net_dinv.p, r1, z, s);
        axpy_safe_ratio(n, sc + 11, sc + 10, 1.0, nv, z, s);  // z += g n
    }
};
// Any preconditioner, with every `net_stride`-th application logged as a training
// sample (r, z) into the attached network's store.
template <class Inner>
struct LoggingPc {
    pfem_ctx *c;
    Inner inner;
    void operator()(const double *r, double *z) {
        inner(r, z);
        if (c->net_calls++ % c->net_stride == 0) net_log_pair(c->net, r, z, c->bcflag.p, c->stream);
    }
};

// FP64 iterative refinement around FP32 Krylov: r = b - K x in FP64, inner FP32
// solve (PCG, or BiCGStab once PCG has broken down on this system or when
// `indefinite` is passed in) of K d = r to a modest relative accuracy, x += d in FP64,
// until ||b - K x|| <= rtol ||b||.
static int ir_solve(pfem_ctx *c, const double *b, double *x, double rtol, int itmax, bool indefinite, int *iters,
                    int *used) {
    hipStream_t s = c->stream;
    const int64_t n = c->ndof;
    double *r = c->w_s.p, *sc = c->scal.p;
    fill<double>(n, 0.0, x, s);
    dot_dev<double, double>(n, b, b, c->partial.p, sc + 4, s);
    const double bn = std::sqrt(c->sync_read(sc + 4));
    *iters = 0;
    if (bn == 0.0) return 0;
    PF_HIP(hipMemcpyAsync(r, b, n * sizeof(double), hipMemcpyDeviceToDevice, s));
    double rn = bn;
    const double inner_tol = 1e-4;
    for (int outer = 0; outer < 40; ++outer) {
        convert<double, float>(n, r, c->f_b.p, s);
        const double tol = std::max(inner_tol, 0.5 * rtol * bn / rn);
        int it = 0, rc = 2;
        if (!indefinite) {
            rc = pcg<float>(c, c->f_b.p, c->f_x.p, tol, std::max(1, itmax - *iters), &it);
            *iters += it;
            if (rc == 2) indefinite = true;
        }
        if (indefinite) {
            rc = bicgstab<float>(c, c->f_b.p, c->f_x.p, tol, std::max(1, itmax - *iters), &it);
            *iters += it;
            if (rc == 2) return 2;
        }
        convert<float, double>(n, c->f_x.p, c->w_t.p, s);
        axpby<double>(n, 1.0, c->w_t.p, 1.0, x, s);
        residual_fine64(c->K, x, b, r, s);  // FP64 refinement residual
        dot_dev<double, double>(n, r, r, c->partial.p, sc + 4, s);
        rn = std::sqrt(c->sync_read(sc + 4));
        if (c->opt.verbose > 1)
            fprintf(stderr, "      IR %2d  inner %3d (%s)  |r|/|b| = %.3e\n", outer, it, indefinite ? "bicgstab" : "pcg",
                    rn / bn);
        if (rn <= rtol * bn) break;
        if (*iters >= itmax) return 1;
    }
    *used = indefinite ? 4 : 3;
    return rn <= rtol * bn ? 0 : 1;
}

// FP32-IR as iterative refinement: the Krylov iteration itself runs in FP32 -- flexible
// CG (FCG(1), as fpcg) with the FP32 copy of K, FP32 vectors and the FP32 V-cycle -- while
// the solution accumulates in FP64 and every refinement sweep starts from the FP64
// residual b - K x.  A sweep stops when its preconditioned residual norm has fallen by
// IR_TAU (an FP32 recursion tracks the true residual over about five decades), when it
// meets the target, or when the recursion loses definiteness near its rounding level
// (p'Kp <= 0 or r'z < 0: the iteration's update is undone or kept, and the next sweep
// restarts from the FP64 residual).  Convergence is only ever accepted at the start of a
// sweep, on the FP64 residual: sqrt(r'z) <= rtol sqrt(r0'z0) with r0 = b, the test of
// pcg().  0 solved; 1 itmax; 2 breakdown before any progress (x = 0); 4 stopped short
// with progress in x (a breakdown at the first iteration of a later sweep, or the sweep
// cap): *rn_left = sqrt(r'z) of the FP64 residual at x and *eps the absolute target, so
// a caller can finish the solve from x.
static int ir32_fcg(pfem_ctx *c, const double *b, double *x, double rtol, int itmax, int *iters, double *rn_left,
                    double *eps_out) {
    constexpr double IR_TAU = 1e-5;
    constexpr int MAX_SWEEPS = 16;
    hipStream_t s = c->stream;
    const int64_t n = c->ndof;
    float *r = c->f_r.p, *z = c->f_z.p, *p = c->f_p.p, *q = c->f_q.p;
    double *r64 = c->w_r.p, *sc = c->scal.p;
    if (c->mpartial.n < multi_dot_partials()) c->mpartial.alloc(multi_dot_partials());
    fill<double>(n, 0.0, x, s);
    *iters = 0;
    double eps = 0.0;
    for (int sweep = 0;; ++sweep) {
        if (sweep == 0) {
            convert<double, float>(n, b, r, s);
        } else {  // FP64 refinement residual
            residual_fine64(c->K, x, b, r64, s);
            convert<double, float>(n, r64, r, s);
        }
        amg_apply32(c, c->amg.get(), r, z);
        dot_dev<float, float>(n, r, z, c->partial.p, sc + 0, s);
        const double rz = c->sync_read(sc + 0);
        if (!(rz >= 0)) return sweep == 0 ? 2 : 4;
        const double rn = std::sqrt(rz);
        *rn_left = rn;
        if (sweep == 0) {
            if (rn == 0.0) return 0;
            eps = rtol * rn;
            *eps_out = eps;
        } else if (rn <= eps) {
            return 0;
        }
        if (sweep == MAX_SWEEPS) return 4;
        const double stop = std::max(eps, IR_TAU * rn);
        PF_HIP(hipMemcpyAsync(p, z, n * sizeof(float), hipMemcpyDeviceToDevice, s));
        for (int k = 0;; ++k) {
            if (*iters >= itmax) return 1;
            ++*iters;
            PF_HIP(hipEventRecord(c->ev0, s));
            spmv_fine32_dot(c->K, p, q, c->partial.p, sc + 1, s);  // q = K p, sc[1] = p'q
            PF_HIP(hipEventRecord(c->ev1, s));
            cg_update_mixed(n, sc + 0, sc + 1, p, q, x, r, s);     // x += a p (FP64), r -= a q
            amg_apply32(c, c->amg.get(), r, z);
            fcg_dots32(n, r, z, q, c->mpartial.p, sc + 2, s);      // sc[2] = r'z, sc[3] = r'r, sc[4] = z'q
            PF_HIP(hipMemcpyAsync(c->h_scal, sc, 5 * sizeof(double), hipMemcpyDeviceToHost, s));
            c->sync();
            float ms = 0.f;
            PF_HIP(hipEventElapsedTime(&ms, c->ev0, c->ev1));
            c->st.t_krylov_spmv32 += 1e-3 * ms;
            c->st.n_krylov_spmv32++;
            const double pq = c->h_scal[1], rzn = c->h_scal[2];
            if (!(pq > 0)) {
                if (c->opt.verbose > 1) fprintf(stderr, "      ir32 p'Kp = %.3e at sweep it %d\n", pq, k + 1);
                // undo this iteration's x += a p, a = r'z / p'q (when a is finite)
                if (std::isfinite(pq) && pq != 0.0) axpy_f32_to_f64(n, -c->h_scal[0] / pq, p, x, s);
                if (k == 0) return sweep == 0 ? 2 : 4;
                break;  // the next sweep restarts from the FP64 residual
            }
            if (!(rzn >= 0)) {
                // the FP32 recursion has lost definiteness (near its rounding level, or an
                // indefinite cycle): x keeps this iteration's update, a sweep that had made
                // progress ends here and the next one restarts from the FP64 residual
                if (c->opt.verbose > 1) fprintf(stderr, "      ir32 r'z = %.3e at sweep it %d\n", rzn, k + 1);
                if (k == 0) return sweep == 0 ? 2 : 4;
                break;
            }
            const double nrm = std::sqrt(rzn);
            if (c->opt.verbose > 2) fprintf(stderr, "      ir32 sweep %d it %3d  %.3e  (stop %.3e)\n", sweep, *iters, nrm, stop);
            if (nrm <= stop) break;  // checked on the FP64 residual at the next sweep's start
            const double beta = -c->h_scal[4] / pq;  // p = z + beta p
            axpby<float>(n, 1.0, z, beta, p, s);
            PF_HIP(hipMemcpyAsync(sc + 0, sc + 2, sizeof(double), hipMemcpyDeviceToDevice, s));
        }
        if (c->opt.verbose > 1) fprintf(stderr, "      ir32 sweep %d: start %.3e, %d its so far\n", sweep, rn, *iters);
    }
}

// FP32 workspaces and the FP32 copy of the imposed fine K.
static void prepare_fp32(pfem_ctx *c) {
    if (!c->f_r.p)
        for (auto *b : {&c->f_r, &c->f_z, &c->f_p, &c->f_q, &c->f_x, &c->f_b, &c->f_rh, &c->f_v, &c->f_s, &c->f_t})
            b->alloc(c->ndof);
    c->split_op = c->opt.split_operator == 1;
    if (c->split_op) {
        split_copy(c->K, c->stream);  // writes the FP32 copy too
    } else {
        if (!c->K.v32.p) c->K.v32.alloc(9 * c->K.nslots);
        convert<double, float>(9 * c->K.nslots, c->K.v64.p, c->K.v32.p, c->stream);
    }
    c->half_fine = c->opt.fp16_fine == 1;  // FP16 scaled fine level inside the V-cycle's smoother
    if (c->half_fine) half_scaled_copy(c->K, c->stream);
}

// GCR(m) with the learned preconditioner; when it stalls (no 0.1% gain in the best
// residual over 40 iterations, or less than a 0.9^10 reduction over the last 10),
// breaks down or hits itmax, AMG solves for the remaining residual b - K x and the
// correction is added (the progress is kept).
static int learned_solve(pfem_ctx *c, double rtol, int itmax, int *total, int *which) {
    hipStream_t s = c->stream;
    int it = 0;
    *which = 5;
    if (!c->net_dinv.p) c->net_dinv.alloc(9 * c->nnodes);
    block_diag_inverse3(c->K, c->net_dinv.p, s);
    const int m = c->opt.gcr_directions > 0 ? std::min(c->opt.gcr_directions, 40) : 16;
    // hand over to AMG once the last 10 iterations together gained less than 0.9^10
    int rc = gcr_m(c, c->rhs.p, c->dU.p, rtol, itmax, m, 40, &it, LearnedPc{c}, 10, 0.9);
    *total += it;
    if (c->opt.verbose > 1) fprintf(stderr, "      learned gcr: rc %d after %d its\n", rc, it);
    if (rc == 0) return 0;
    *which = 6;
    if (c->pc_stale == 2 && c->pc_ready) {  // the fine-level smoother only
        const double t0 = now();
        amg_refresh_fine(c, c->amg.get());
        c->st.t_setup += now() - t0;
        c->pc_stale = 0;
    } else if ((c->pc_stale == 4 || c->pc_stale == 5) && c->pc_ready) {  // policy 7, as linear_solve_impl
        double t0 = now();
        const bool restored = c->pc_stale == 4 && amg_restore_coarse(c, c->amg.get(), c->stash_slot);
        c->st.t_setup += now() - t0;
        if (!restored) {
            pc_setup(c, true);
            t0 = now();
            amg_stash_coarse(c, c->amg.get(), c->stash_slot);
            c->st.t_setup += now() - t0;
        }
        c->stash_used = restored ? 4 : 5;
        c->pc_stale = 0;
    } else if (c->pc_stale || !c->pc_ready) {
        pc_setup(c, c->pc_stale == 3);
        c->pc_stale = 0;
    }
    // remaining residual b2 = b - K x and the AMG correction x2 live in the GCR direction
    // store (free now): PCG and BiCGStab use every w_* workspace themselves
    double *b2 = c->fcg_P.p, *x2 = c->fcg_P.p + c->ndof;
    residual_fine64(c->K, c->dU.p, c->rhs.p, b2, s);
    dot_dev<double, double>(c->ndof, b2, b2, c->partial.p, c->scal.p + 6, s);
    dot_dev<double, double>(c->ndof, c->rhs.p, c->rhs.p, c->partial.p, c->scal.p + 7, s);
    PF_HIP(hipMemcpyAsync(c->h_scal, c->scal.p, 8 * sizeof(double), hipMemcpyDeviceToHost, s));
    c->sync();
    const double rn = std::sqrt(c->h_scal[6]), bn = std::sqrt(c->h_scal[7]);
    if (!(rn > rtol * bn)) return 0;
    const double rt = std::min(0.5, rtol * bn / rn);
    if (c->opt.precision == PFEM_PREC_FP32_IR) {
        rc = fpcg(c, b2, x2, rt, itmax, true, 0, &it, Amg32Pc{c});
    } else {
        rc = pcg<double>(c, b2, x2, rt, itmax, &it);
    }
    *total += it;
    c->amg_phase_iters = it;
    if (rc != 0) {
        int it2 = 0;
        rc = bicgstab<double>(c, b2, x2, rt, 2 * itmax, &it2);
        *total += it2;
        c->amg_phase_iters += it2;
    }
    axpby<double>(c->ndof, 1.0, x2, 1.0, c->dU.p, s);
    return rc;
}

// rebuild: 0 reuse the preconditioner, 1 rebuild it, 2 refresh only the fine-level
// smoother from the current K (interpolation and coarse operators kept), 3 rebuild it
// keeping the prolongators of the last full set-up
int linear_solve_impl(pfem_ctx *c, double rtol, int rebuild, int *niter, int *used, double *relres,
                      bool *discarded) {
    const bool learned = c->net_mode == 2 || c->net_mode == 3;
    c->stash_used = 0;
    if (learned) {
        if (c->opt.precision == PFEM_PREC_FP32_IR) prepare_fp32(c);
        // the AMG hierarchy is rebuilt only if it is needed (a full rebuild request wins
        // over a keep-the-prolongators one)
        if (rebuild) c->pc_stale = (c->pc_stale == 1 || rebuild == 1) ? 1 : rebuild;
    } else if (rebuild == 1 || !c->pc_ready) {
        pc_setup(c);
    } else if (rebuild == 2) {
        const double t0 = now();
        amg_refresh_fine(c, c->amg.get());
        c->st.t_setup += now() - t0;
    } else if (rebuild == 3) {
        pc_setup(c, true);
    } else if (rebuild == 4) {  // policy 7: the kept coarse levels, the fine smoother from this K
        const double t0 = now();
        const bool restored = amg_restore_coarse(c, c->amg.get(), c->stash_slot);
        c->st.t_setup += now() - t0;
        if (!restored) rebuild = 5;
    }
    if (rebuild == 5) {  // policy 7: set-up keeping P, and its coarse levels kept aside
        pc_setup(c, true);
        const double t0 = now();
        amg_stash_coarse(c, c->amg.get(), c->stash_slot);
        c->st.t_setup += now() - t0;
    }
    const double t1 = now();
    int it = 0, total = 0, which = 0, rc;
    const int itmax = c->opt.cg_itmax > 0 ? c->opt.cg_itmax : 200;
    hipStream_t s = c->stream;
    if (learned) {
        rc = learned_solve(c, rtol, itmax, &total, &which);
    } else if (c->net && c->net_mode == 1 && c->opt.precision != PFEM_PREC_FP32_IR) {
        rc = fpcg(c, c->rhs.p, c->dU.p, rtol, itmax, true, 0, &it, LoggingPc<Amg64Pc>{c, Amg64Pc{c}});
        total += it;
        if (rc != 0) {
            which = 2;
            rc = bicgstab<double>(c, c->rhs.p, c->dU.p, rtol, 2 * itmax, &it);
            total += it;
        }
    } else if (c->opt.precision == PFEM_PREC_FP32_IR) {
        prepare_fp32(c);
        which = 3;
        if (c->opt.fp32_krylov && !c->net) {
            which = 9;
            double rn_left = 0.0, eps_abs = 0.0;
            rc = ir32_fcg(c, c->rhs.p, c->dU.p, rtol, itmax, &it, &rn_left, &eps_abs);
            if (rc == 2) {
                // the FP32 recursion broke down before any progress: the FP64 flexible CG
                // with the FP32 cycle solves from scratch (a breakdown of the FP32
                // recursion says nothing about the hierarchy, so policy 6 keeps it)
                total += it;
                which = 3;
                rc = fpcg(c, c->rhs.p, c->dU.p, rtol, itmax, true, 0, &it, Amg32Pc{c});
            } else if (rc == 4) {
                // the FP32 recursion stopped short of the target (a breakdown at its rounding
                // level): FP64 flexible CG with the FP32 cycle solves for the remaining
                // residual, to the same absolute target, and its correction is added
                total += it;
                which = 11;
                if (c->fcg_P.n < (size_t)2 * c->ndof) c->fcg_P.alloc((size_t)2 * c->ndof);
                double *b2 = c->fcg_P.p, *x2 = c->fcg_P.p + c->ndof;
                residual_fine64(c->K, c->dU.p, c->rhs.p, b2, s);
                const double rt = std::min(0.5, eps_abs / rn_left);
                rc = fpcg(c, b2, x2, rt, itmax, true, 0, &it, Amg32Pc{c});
                axpby<double>(c->ndof, 1.0, x2, 1.0, c->dU.p, s);
            }
        } else if (c->net && c->net_mode == 1) {
            rc = fpcg(c, c->rhs.p, c->dU.p, rtol, itmax, true, 0, &it, LoggingPc<Amg32Pc>{c, Amg32Pc{c}});
        } else if (c->net && c->net_mode == 4) {
            rc = fpcg(c, c->rhs.p, c->dU.p, rtol, itmax, true, 0, &it, Amg32ErrLogPc{c, c->dU.p});
            net_log_error_finish(c->net, c->dU.p, c->bcflag.p, rc == 0, s);
        } else if (c->net && c->net_mode == 5) {
            which = 10;
            rc = fpcg(c, c->rhs.p, c->dU.p, rtol, itmax, true, 0, &it, Amg32NetPc{c});
            if (rc != 0) {  // the corrected cycle failed: the plain FP32 cycle from scratch
                total += it;
                which = 3;
                rc = fpcg(c, c->rhs.p, c->dU.p, rtol, itmax, true, 0, &it, Amg32Pc{c});
            }
        } else {
            rc = fpcg(c, c->rhs.p, c->dU.p, rtol, itmax, true, 0, &it, Amg32Pc{c});
        }
        total += it;
        // after a breakdown or a stall: flexible GCR(16) with the same FP32 cycle (it needs
        // no symmetry, minimises ||r|| and costs one product and one cycle per
        // iteration); krylov_fallback = 1: refinement with BiCGStab instead
        if (rc != 0 && c->opt.krylov_fallback == 0) {
            which = 7;
            rc = gcr_m(c, c->rhs.p, c->dU.p, rtol, 4 * itmax, 16, 40, &it, Amg32Pc{c});
            if (rc == 3) rc = 1;  // stalled: not converged
            total += it;
        } else if (rc != 0) {  // indefinite (2) or stalled (1): refinement sweeps, BiCGStab after a breakdown
            rc = ir_solve(c, c->rhs.p, c->dU.p, rtol, 4 * itmax, rc == 2, &it, &which);
            total += it;
        }
        if (rc != 0) {  // FP32 refinement failed: the FP64 ladder takes over
            which = 2;
            rc = pcg<double>(c, c->rhs.p, c->dU.p, rtol, itmax, &it);
            total += it;
            if (rc != 0) {
                rc = bicgstab<double>(c, c->rhs.p, c->dU.p, rtol, 2 * itmax, &it);
                total += it;
            }
        }
    } else {
        rc = pcg<double>(c, c->rhs.p, c->dU.p, rtol, itmax, &it);
        total += it;
        if (rc == 1) {  // Solver.jl:270-273: refresh the preconditioner, retry once
            pc_setup(c);
            which = 1;
            rc = pcg<double>(c, c->rhs.p, c->dU.p, rtol, itmax, &it);
            total += it;
        }
        if (rc != 0 && c->opt.krylov_fallback == 0) {  // flexible GCR(16) with the FP64 cycle
            which = 8;
            rc = gcr_m(c, c->rhs.p, c->dU.p, rtol, 2 * itmax, 16, 40, &it, Amg64Pc{c});
            if (rc == 3) rc = 1;  // stalled: not converged
            total += it;
        } else if (rc != 0) {  // in place of the direct fallback (Solver.jl:262, :284)
            which = 2;
            rc = bicgstab<double>(c, c->rhs.p, c->dU.p, rtol, 2 * itmax, &it);
            total += it;
        }
    }
    // true residual ||b - K dU|| / ||b||
    residual_fine64(c->K, c->dU.p, c->rhs.p, c->w_s.p, s);
    dot_dev<double, double>(c->ndof, c->w_s.p, c->w_s.p, c->partial.p, c->scal.p + 5, s);
    dot_dev<double, double>(c->ndof, c->rhs.p, c->rhs.p, c->partial.p, c->scal.p + 6, s);
    PF_HIP(hipMemcpyAsync(c->h_scal, c->scal.p, 8 * sizeof(double), hipMemcpyDeviceToHost, s));
    c->sync();
    const double rr = std::sqrt(c->h_scal[5]), bb = std::sqrt(c->h_scal[6]);
    if (relres) *relres = bb > 0 ? rr / bb : 0.0;
    const bool drop = !(rr <= bb);
    if (drop) {  // a failed solve that made things worse is not applied
        fill<double>(c->ndof, 0.0, c->dU.p, s);
        c->sync();
    }
    if (discarded) *discarded = drop;
    if (niter) *niter = total;
    // policy 7: the solve right after a kept set-up sets the mark; one on the restored
    // levels that needs clearly more iterations makes the next first iteration rebuild
    // (with the learned preconditioner: the AMG phase's kept levels and iterations)
    const int stash_code = learned ? c->stash_used : rebuild;
    const int stash_its = learned ? c->amg_phase_iters : total;
    if (stash_code == 5) {
        AmgStash &st = amg_stash_slot(c->amg.get(), c->stash_slot);
        st.iters = stash_its;
        st.stale = false;
    } else if (stash_code == 4) {
        AmgStash &st = amg_stash_slot(c->amg.get(), c->stash_slot);
        if (stash_its > st.iters + std::max(1, st.iters / 16)) st.stale = true;
    }
    if (used) *used = which;
    c->st.t_solve += now() - t1;
    c->st.n_solve++;
    c->st.n_krylov += total;
    if (rc == 1) return PFEM_ERR_NOT_CONVERGED;
    if (rc == 2) return PFEM_ERR_NON_SPD;
    return PFEM_OK;
}

}  // namespace pfem

extern "C" {

// AMG policy 7 at Newton iteration `it` (not the second): the coarse levels kept from the
// same iteration of an earlier load step (4) unless that slot is empty or stale, then a
// set-up keeping P whose coarse levels are kept (5).  The second iteration, the first on
// the new increment's plastic tangent, always gets its own set-up: coarse levels kept
// from another load step's second iteration made the Krylov solve break down.
static int policy7_code(pfem_ctx *c, int it) {
    constexpr int SLOTS = 8;
    c->stash_slot = std::min(it, SLOTS) - 1;
    Amg *a = c->amg.get();
    if (!a || (int)a->stash.size() <= c->stash_slot) return 5;
    const AmgStash &st = a->stash[c->stash_slot];
    return (st.ok && !st.stale) ? 4 : 5;
}

// _forcing_term (Solver.jl:159-171); it is 1-based
static double forcing_term(const pfem_options &o, int it, double rk, double rkm1, double eta_prev) {
    if (it == 1 || rkm1 <= 0) return o.eta_max;
    double eta = o.ew_gamma * std::pow(rk / rkm1, o.ew_alpha);
    const double safe = o.ew_gamma * std::pow(eta_prev, o.ew_alpha);
    if (safe > 0.1) eta = std::max(eta, safe);
    return std::min(std::max(eta, o.eta_min), o.eta_max);
}

int pfem_newton(pfem_ctx *c, double lambda, double tol, int maxiter, pfem_newton_result *res, pfem_fail *fail) {
    if (!c || !res || maxiter < 1 || maxiter > 64) return PFEM_ERR_BAD_ARG;
    std::memset(res, 0, sizeof *res);
    try {
        hipStream_t s = c->stream;
        const double T0 = now();
        const double fext_norm = std::fabs(lambda) * c->fext0_norm;  // Solver.jl:336
        double ref = 1.0, rkm1 = 0.0, eta_prev = c->opt.eta_max, rbest = 0.0;
        bool first = true;
        // K, R and ||R|| at the current U (assemble, R -= lambda Fext, impose_dirichlet! for
        // the ramped set, then for the fixed set)
        auto evaluate = [&](double *rn) -> int {
            double t = now();
            int rc = assemble_impl(c, 0, fail);  // Solver.jl:347
            if (rc) return rc;
            res->t_assemble += now() - t;
            t = now();
            sub_external(c->ndof, lambda, c->Fext0.p, c->R.p, s);  // :349
            impose_set(c, 2, lambda);                              // :354 dir_ramp
            impose_set(c, 1, lambda);                              // :355 dir_fix
            dot_dev<double, double>(c->ndof, c->R.p, c->R.p, c->partial.p, c->scal.p + 7, s);
            *rn = std::sqrt(c->sync_read(c->scal.p + 7));  // :358
            res->t_bc += now() - t;
            return PFEM_OK;
        };
        // Optional backtracking on ||R|| (line_search = most halvings; 0 = the reference's
        // full Newton step).  The residual at the accepted point is the next iteration's.
        const int ls_max = std::max(0, std::min(c->opt.line_search, 20));
        double rnorm = 0.0;
        bool have_residual = false, force_rebuild = false, last_discarded = false;
        for (int it = 1; it <= maxiter; ++it) {
            if (!have_residual) {
                const int rc0 = evaluate(&rnorm);
                if (rc0) {
                    res->t_total = now() - T0;  // the iterations before an inversion stay reported
                    return rc0;
                }
            }
            have_residual = false;
            int rc = PFEM_OK;
            res->residuals[it - 1] = rnorm;
            res->iters = it;
            if (it == 1) ref = std::max(std::max(rnorm, fext_norm), 1.0);  // :360-362
            if (c->opt.verbose) fprintf(stderr, "    iter %d  |R| = %.6e\n", it, rnorm);
            if (rnorm <= tol * ref) {  // :364
                res->converged = 1;
                break;
            }
            // optional early stop of a diverging load step (not in the reference, which runs
            // maxiter iterations): from the third iteration on, ||R|| above newton_diverge
            // times the smallest residual so far
            if (c->opt.newton_diverge > 0 && it >= 3 && !(rnorm <= c->opt.newton_diverge * rbest)) {
                res->stopped = 1;
                break;
            }
            rbest = it == 1 ? rnorm : std::min(rbest, rnorm);
            const double eta = forcing_term(c->opt, it, rnorm, rkm1, eta_prev);  // :371
            res->eta[it - 1] = eta;
            // :374 rebuilds once per load step; the other policies (pfem_amg_rebuild) rebuild
            // from later tangents too, keep the prolongators (3 = rebuild all but P), or
            // refresh only the fine-level smoother (2)
            const int policy = c->opt.amg_rebuild;
            const bool keep_P = c->pc_ready && !c->refresh_P;
            int rebuild = policy == 5                                            ? (c->pc_ready ? 3 : 1)
                          : policy == 6 ? (keep_P ? 3 : 1)
                          : policy == 7 ? (!c->pc_ready ? 1 : it == 2 ? 3 : policy7_code(c, it))
                          : (first || policy == 1 || (policy == 2 && it == 2)) ? 1
                          : policy == 3                                          ? 2
                          : policy == 4                                          ? 3
                                                                                 : 0;
            // a discarded correction is retried once with a preconditioner built from scratch
            if (force_rebuild) rebuild = 1;
            force_rebuild = false;
            first = false;
            axpby<double>(c->ndof, -1.0, c->R.p, 0.0, c->rhs.p, s);  // :378
            const double t = now();
            const double setup0 = c->st.t_setup;
            int niter = 0, used = 0;
            double rel = 0;
            bool dropped = false;
            rc = linear_solve_impl(c, eta, rebuild, &niter, &used, &rel, &dropped);  // :379
            res->t_setup += c->st.t_setup - setup0;
            res->t_solve += now() - t - (c->st.t_setup - setup0);
            res->krylov_iters[it - 1] = niter;
            res->solver_used[it - 1] = used;
            res->solves++;
            // policy 6: a solve that needed a fallback (PCG retried after a refresh, the
            // BiCGStab ladder, or BiCGStab refinement) rebuilds the prolongators next time
            // (not under fp32_krylov: its fallbacks follow breakdowns of the FP32 recursion,
            // and prolongators rebuilt from a plastic tangent then failed the next load step)
            const bool fell_back = used == 1 || used == 2 || used == 4 || used == 7 || used == 8;
            if (policy == 6) c->refresh_P = !c->opt.fp32_krylov && fell_back;
            // policy 7 keeps the prolongators (a rebuild from a plastic tangent made every
            // later first-iteration solve several times slower) and drops the kept coarse
            // levels instead
            if (policy == 7 && fell_back && c->amg)
                for (AmgStash &st : c->amg->stash) st.ok = false;
            if (c->opt.verbose)
                fprintf(stderr, "    iter %d  eta = %.2e  krylov = %d (solver %d, true relres %.2e)%s\n", it, eta,
                        niter, used, rel, dropped ? " discarded" : "");
            if (rc == PFEM_ERR_NOT_CONVERGED || rc == PFEM_ERR_NON_SPD) {
                // keep the (inexact) correction, as the reference keeps the direct fallback's
            } else if (rc) {
                return rc;
            }
            if (dropped) {
                res->discarded++;
                if (last_discarded) {  // no usable correction twice in a row: give up the step
                    res->stopped = 2;
                    break;
                }
                force_rebuild = true;
            }
            last_discarded = dropped;
            axpby<double>(c->ndof, 1.0, c->dU.p, 1.0, c->U.p, s);  // :383 (also at it == maxiter)
            if (ls_max > 0 && it < maxiter) {
                double alpha = 1.0, rn_new = 0.0;
                for (int h = 0;; ++h) {
                    const int rc1 = evaluate(&rn_new);
                    if (rc1 != PFEM_OK && rc1 != PFEM_ERR_INVERTED) return rc1;
                    if (rc1 == PFEM_OK && rn_new <= (1.0 - 1e-4 * alpha) * rnorm) break;
                    if (h == ls_max) {
                        if (rc1 != PFEM_OK) return rc1;
                        break;  // no decrease found: keep the shortest step
                    }
                    axpby<double>(c->ndof, -0.5 * alpha, c->dU.p, 1.0, c->U.p, s);  // U = U0 + alpha/2 dU
                    alpha *= 0.5;
                    if (c->opt.verbose) fprintf(stderr, "    line search: step %.4f\n", alpha);
                }
                rkm1 = rnorm;
                eta_prev = eta;
                rnorm = rn_new;
                have_residual = true;
                continue;
            }
            rkm1 = rnorm;
            eta_prev = eta;
        }
        c->sync();
        res->t_total = now() - T0;
        return PFEM_OK;
    } catch (const hip_failure &e) {
        if (fail) fail->hip = (int32_t)e.code;
        return PFEM_ERR_HIP;
    } catch (const status_failure &e) {
        return e.status;
    } catch (const std::exception &) {
        return PFEM_ERR_BAD_ARG;
    }
}

int pfem_commit(pfem_ctx *c, pfem_fail *fail) {
    if (!c) return PFEM_ERR_BAD_ARG;
    try {
        // Solver.jl:470-471 re-assembles with commit = true and discards K and R; only the
        // history update is needed, so the tangent, K and R are skipped (they keep the
        // last Newton iteration's values)
        return assemble_impl(c, 1, fail, true);
    } catch (const hip_failure &e) {
        if (fail) fail->hip = (int32_t)e.code;
        return PFEM_ERR_HIP;
    }
}

int pfem_element_batch(int32_t kind, const pfem_material *mat, int64_t nelem, const double *Xe, const double *ue,
                       double *epsp, double *beta, double *abar, double *cpinv, int commit, double *Fe, double *Ke,
                       double *sig, pfem_fail *fail) {
    if (kind != PFEM_HEX8_FINITE && kind != PFEM_HEX8_FINITE_FBAR && kind != PFEM_HEX8_SMALL)
        return PFEM_ERR_BAD_ARG;
    if (!mat || nelem <= 0 || !Xe || !ue || !epsp || !abar || (!cpinv && kind != PFEM_HEX8_SMALL) || !Fe || !Ke ||
        !sig)
        return PFEM_ERR_BAD_ARG;
    DevMat dm;
    int rc = make_devmat(*mat, dm);
    if (rc) return rc;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return PFEM_ERR_NO_DEVICE;
    try {
        hipStream_t s;
        PF_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        struct SG { hipStream_t s; ~SG() { (void)hipStreamDestroy(s); } } sg{s};
        const int64_t ngp = 8 * nelem;
        dbuf<double> X(24 * nelem), U(24 * nelem), ep(6 * ngp), bt(6 * ngp), ab(ngp), cp(6 * ngp), sg6(6 * ngp),
            pk((size_t)ngp * PK), geo((size_t)ngp * 24), g0(24 * nelem), fe(24 * nelem), ke(576 * nelem);
        dbuf<int64_t> cn(8 * nelem);
        dbuf<DevMat> md(1);
        dbuf<unsigned long long> fk(1);
        std::vector<int64_t> hc(8 * nelem);
        for (int64_t k = 0; k < 8 * nelem; ++k) hc[k] = k;
        X.upload(Xe, 24 * nelem, s); U.upload(ue, 24 * nelem, s); cn.upload(hc.data(), 8 * nelem, s);
        ep.upload(epsp, 6 * ngp, s); ab.upload(abar, ngp, s);
        if (cpinv) cp.upload(cpinv, 6 * ngp, s); else cp.zero(s);
        if (beta) bt.upload(beta, 6 * ngp, s); else bt.zero(s);
        md.upload(&dm, 1, s);
        sg6.zero(s);
        PF_HIP(hipMemsetAsync(fk.p, 0xff, sizeof(unsigned long long), s));
        ElementArgs a{};
        a.nelem = nelem; a.X = X.p; a.conn = cn.p; a.U = U.p; a.matid = nullptr; a.mats = md.p;
        a.kind = kind; a.fbar = kind == PFEM_HEX8_FINITE_FBAR; a.kin = dm.Hkin > 0; a.commit = commit;
        a.epsp = ep.p; a.beta = bt.p; a.abar = ab.p; a.cpinv = cp.p;
        a.sig = sg6.p; a.packet = pk.p; a.g0 = g0.p; a.fail_key = fk.p;
        a.geo = geo.p; a.geo_write = 1;
        launch_gp_kernel(a, 0, nelem, s);
        unsigned long long key = 0;
        PF_HIP(hipMemcpyAsync(&key, fk.p, sizeof key, hipMemcpyDeviceToHost, s));
        PF_HIP(hipStreamSynchronize(s));
        if (key != ~0ull) {
            dbuf<double> J(1);
            launch_detF_kernel(a, (int64_t)key, J.p, s);
            double hJ = 0;
            J.download(&hJ, 1, s);
            PF_HIP(hipStreamSynchronize(s));
            if (fail) { fail->e = (int64_t)(key / 8); fail->g = (int32_t)(key % 8); fail->J = hJ; }
            return PFEM_ERR_INVERTED;
        }
        launch_dense_ke_kernel(a, nelem, fe.p, ke.p, s);
        fe.download(Fe, 24 * nelem, s);
        ke.download(Ke, 576 * nelem, s);
        sg6.download(sig, 6 * ngp, s);
        if (commit) {
            ep.download(epsp, 6 * ngp, s);
            ab.download(abar, ngp, s);
            if (cpinv) cp.download(cpinv, 6 * ngp, s);
            if (beta) bt.download(beta, 6 * ngp, s);
        }
        PF_HIP(hipStreamSynchronize(s));
        return PFEM_OK;
    } catch (const hip_failure &e) {
        if (fail) fail->hip = (int32_t)e.code;
        return PFEM_ERR_HIP;
    }
}

int pfem_dense_inverse(int64_t n, const double *A, double *inv) {
    if (n <= 0 || !A || !inv) return PFEM_ERR_BAD_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return PFEM_ERR_NO_DEVICE;
    try {
        hipStream_t s;
        PF_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        struct SG { hipStream_t s; ~SG() { (void)hipStreamDestroy(s); } } sg{s};
        dbuf<double> W((size_t)2 * n * n), prow(2 * n), fcol(n), out((size_t)n * n), dw(gj_scratch_doubles(n));
        dbuf<int> iw(gj_scratch_ints(n));
        PF_HIP(hipMemcpy2DAsync(W.p, 2 * n * sizeof(double), A, n * sizeof(double), n * sizeof(double), n,
                                hipMemcpyHostToDevice, s));
        gauss_jordan_inverse(n, W.p, prow.p, fcol.p, iw.p, dw.p, out.p, s);
        out.download(inv, (size_t)n * n, s);
        PF_HIP(hipStreamSynchronize(s));
        return PFEM_OK;
    } catch (const hip_failure &) {
        return PFEM_ERR_HIP;
    }
}

int pfem_bench_spmv(pfem_ctx *c, int prec, int reps, double *sec) {
    if (!c || !sec || reps < 1 || prec < 0 || prec > 2) return PFEM_ERR_BAD_ARG;
    try {
        hipStream_t s = c->stream;
        const int64_t n = c->ndof;
        if (prec == 2) {
            split_copy(c->K, s);
            fill<double>(n, 1.0, c->w_p.p, s);
            spmv_fine_split(c->K, c->w_p.p, c->w_q.p, s);
            PF_HIP(hipEventRecord(c->ev0, s));
            for (int i = 0; i < reps; ++i) spmv_fine_split(c->K, c->w_p.p, c->w_q.p, s);
            PF_HIP(hipEventRecord(c->ev1, s));
        } else if (prec == 1) {
            if (!c->K.v32.p) {
                c->K.v32.alloc(9 * c->K.nslots);
                convert<double, float>(9 * c->K.nslots, c->K.v64.p, c->K.v32.p, s);
            }
            if (!c->f_p.p) { c->f_p.alloc(n); c->f_q.alloc(n); }
            fill<float>(n, 1.0f, c->f_p.p, s);
            spmv_fine32(c->K, c->f_p.p, c->f_q.p, s);
            PF_HIP(hipEventRecord(c->ev0, s));
            for (int i = 0; i < reps; ++i) spmv_fine32(c->K, c->f_p.p, c->f_q.p, s);
            PF_HIP(hipEventRecord(c->ev1, s));
        } else {
            fill<double>(n, 1.0, c->w_p.p, s);
            spmv_fine64(c->K, c->w_p.p, c->w_q.p, s);
            PF_HIP(hipEventRecord(c->ev0, s));
            for (int i = 0; i < reps; ++i) spmv_fine64(c->K, c->w_p.p, c->w_q.p, s);
            PF_HIP(hipEventRecord(c->ev1, s));
        }
        c->sync();
        float ms = 0;
        PF_HIP(hipEventElapsedTime(&ms, c->ev0, c->ev1));
        *sec = 1e-3 * ms / reps;
        c->st.t_spmv = *sec;
        return PFEM_OK;
    } catch (const hip_failure &e) {
        return PFEM_ERR_HIP;
    }
}

int pfem_bench_vcycle(pfem_ctx *c, int prec, int reps, double *sec_cycle, double *bytes_cycle, double *sec_step,
                      double *bytes_step) {
    if (!c || !sec_cycle || !bytes_cycle || !sec_step || !bytes_step || reps < 1 || (prec != 0 && prec != 1))
        return PFEM_ERR_BAD_ARG;
    if (!c->amg || !c->pc_ready || (prec == 1 && !c->amg->have32)) return PFEM_ERR_BAD_ARG;
    try {
        amg_bench(c, c->amg.get(), prec, reps, sec_cycle, bytes_cycle, sec_step, bytes_step);
        return PFEM_OK;
    } catch (const hip_failure &e) {
        return PFEM_ERR_HIP;
    }
}

}  // extern "C"
