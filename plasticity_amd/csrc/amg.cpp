// amg.cpp -- smoothed-aggregation AMG: host symbolic set-up, device numeric set-up
// orchestration and the V-cycle (see amg.h for the method and the reference seam).
#include "amg.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <functional>
#include <type_traits>

#include "context.h"

namespace pfem {
namespace {

// CSR view of a level's block graph.
struct Graph {
    int64_t nb = 0;
    std::vector<int64_t> ptr;
    std::vector<int32_t> idx;
};

Graph fine_graph(const FineMatrix &K) {
    Graph g;
    g.nb = K.nb;
    g.ptr.assign(K.nb + 1, 0);
    for (int64_t i = 0; i < K.nb; ++i) g.ptr[i + 1] = g.ptr[i] + K.h_rowlen[i];
    g.idx.resize(g.ptr[K.nb]);
    for (int64_t i = 0; i < K.nb; ++i) {
        const int64_t base = K.h_slice_ptr[i / SELL_C] + i % SELL_C;
        for (int k = 0; k < K.h_rowlen[i]; ++k) g.idx[g.ptr[i] + k] = K.h_colind[base + (int64_t)k * SELL_C];
    }
    return g;
}

// Standard three-phase aggregation (Vanek, Mandel & Brezina, Computing 56, 1996) on
// the block graph with every structural connection strong (theta = 0).
int64_t aggregate(const Graph &g, const std::vector<uint8_t> &active, std::vector<int32_t> &agg) {
    const int64_t nb = g.nb;
    agg.assign(nb, -1);
    int32_t na = 0;
    for (int64_t i = 0; i < nb; ++i) {  // phase 1: whole free neighbourhoods
        if (!active[i] || agg[i] >= 0) continue;
        bool free_nbhd = true;
        int nact = 0;
        for (int64_t k = g.ptr[i]; k < g.ptr[i + 1]; ++k) {
            const int32_t j = g.idx[k];
            if (j == i || !active[j]) continue;
            ++nact;
            if (agg[j] >= 0) { free_nbhd = false; break; }
        }
        if (!free_nbhd || nact == 0) continue;
        agg[i] = na;
        for (int64_t k = g.ptr[i]; k < g.ptr[i + 1]; ++k) {
            const int32_t j = g.idx[k];
            if (active[j]) agg[j] = na;
        }
        ++na;
    }
    std::vector<int32_t> p1(agg);
    for (int64_t i = 0; i < nb; ++i) {  // phase 2: join a neighbouring phase-1 aggregate
        if (!active[i] || agg[i] >= 0) continue;
        for (int64_t k = g.ptr[i]; k < g.ptr[i + 1]; ++k) {
            const int32_t j = g.idx[k];
            if (j != i && p1[j] >= 0) { agg[i] = p1[j]; break; }
        }
    }
    for (int64_t i = 0; i < nb; ++i) {  // phase 3: leftovers with their free neighbours
        if (!active[i] || agg[i] >= 0) continue;
        agg[i] = na;
        for (int64_t k = g.ptr[i]; k < g.ptr[i + 1]; ++k) {
            const int32_t j = g.idx[k];
            if (active[j] && agg[j] < 0) agg[j] = na;
        }
        ++na;
    }
    return na;
}

// Per-aggregate thin QR (modified Gram-Schmidt, two passes) of the near-null-space
// rows B (bs x 6 per node).  Pt receives Q's rows per node, Bc the 6x6 R factor per
// aggregate.  A column whose norm collapses below 1e-10 of its original is dropped.
void tentative(int bs, int64_t nb, int64_t nagg, const std::vector<int32_t> &agg, const std::vector<double> &B,
               std::vector<double> &Pt, std::vector<double> &Bc) {
    Pt.assign((size_t)nb * bs * 6, 0.0);
    Bc.assign((size_t)nagg * 36, 0.0);
    std::vector<int64_t> ptr(nagg + 1, 0);
    for (int64_t i = 0; i < nb; ++i)
        if (agg[i] >= 0) ptr[agg[i] + 1]++;
    for (int64_t a = 0; a < nagg; ++a) ptr[a + 1] += ptr[a];
    std::vector<int64_t> members(ptr[nagg]), f(ptr.begin(), ptr.end() - 1);
    for (int64_t i = 0; i < nb; ++i)
        if (agg[i] >= 0) members[f[agg[i]]++] = i;
    std::vector<double> M, Qm;
    for (int64_t a = 0; a < nagg; ++a) {
        const int64_t m0 = ptr[a], cnt = ptr[a + 1] - m0;
        const int64_t rows = cnt * bs;
        M.assign(rows * 6, 0.0);  // column-major rows x 6
        for (int64_t t = 0; t < cnt; ++t)
            for (int r = 0; r < bs; ++r)
                for (int col = 0; col < 6; ++col)
                    M[col * rows + t * bs + r] = B[((size_t)members[m0 + t] * bs + r) * 6 + col];
        Qm.assign(rows * 6, 0.0);
        double Rf[36] = {0};
        for (int col = 0; col < 6; ++col) {
            double *v = &M[col * rows];
            double n0 = 0.0;
            for (int64_t r = 0; r < rows; ++r) n0 += v[r] * v[r];
            n0 = std::sqrt(n0);
            for (int pass = 0; pass < 2; ++pass)
                for (int j = 0; j < col; ++j) {
                    const double *q = &Qm[j * rows];
                    double d = 0.0;
                    for (int64_t r = 0; r < rows; ++r) d += q[r] * v[r];
                    Rf[6 * j + col] += d;
                    for (int64_t r = 0; r < rows; ++r) v[r] -= d * q[r];
                }
            double nn = 0.0;
            for (int64_t r = 0; r < rows; ++r) nn += v[r] * v[r];
            nn = std::sqrt(nn);
            if (n0 == 0.0 || nn <= 1e-10 * n0) continue;  // dropped: zero column
            Rf[6 * col + col] = nn;
            for (int64_t r = 0; r < rows; ++r) Qm[col * rows + r] = v[r] / nn;
        }
        for (int64_t t = 0; t < cnt; ++t)
            for (int r = 0; r < bs; ++r)
                for (int col = 0; col < 6; ++col)
                    Pt[((size_t)members[m0 + t] * bs + r) * 6 + col] = Qm[col * rows + t * bs + r];
        std::memcpy(&Bc[36 * a], Rf, sizeof Rf);
    }
}

template <class T>
void up(dbuf<T> &d, const std::vector<T> &h, hipStream_t s) {
    d.alloc(std::max<size_t>(1, h.size()));
    d.upload(h.data(), h.size(), s);
}

void sorted_union_rows(int64_t nrows, const std::function<void(int64_t, std::vector<int32_t> &)> &gather,
                       std::vector<int32_t> &ptr, std::vector<int32_t> &idx) {
    ptr.assign(nrows + 1, 0);
    idx.clear();
    std::vector<int32_t> tmp;
    for (int64_t i = 0; i < nrows; ++i) {
        tmp.clear();
        gather(i, tmp);
        std::sort(tmp.begin(), tmp.end());
        tmp.erase(std::unique(tmp.begin(), tmp.end()), tmp.end());
        idx.insert(idx.end(), tmp.begin(), tmp.end());
        if (idx.size() >= (size_t)INT32_MAX) throw status_failure(PFEM_ERR_BAD_ARG, "AMG pattern too large");
        ptr[i + 1] = (int32_t)idx.size();
    }
}

}  // namespace

Amg *amg_create(pfem_ctx *c) {
    hipStream_t s = c->stream;
    auto *a = new Amg();
    a->smoother = c->opt.smoother;
    // V-cycle by default; amg_wcycle selects the W-cycle (a second coarse cycle at
    // every level below the finest).  Under FP32-IR the W-cycle needed 20 instead of 28
    // Krylov iterations per solve when the prolongators were rebuilt at every Newton
    // iteration, but with them kept (rebuild policies 4-6) the V-cycle is faster overall.
    a->wcycle = c->opt.amg_wcycle != 0;
    a->degree = c->opt.smoother_degree > 0 ? c->opt.smoother_degree : (a->smoother == PFEM_SMOOTH_JACOBI ? 1 : 2);
    a->coarse_degree = c->opt.coarse_degree > 0 ? c->opt.coarse_degree : a->degree;
    const int64_t coarse_max = c->opt.coarse_max > 0 ? c->opt.coarse_max : 1500;
    const int max_levels = c->opt.max_levels > 1 ? c->opt.max_levels : 10;

    // level 0 graph, activity and near-null space (Solver.jl:117-129, from undeformed
    // coordinates); rows of constrained dofs are zeroed (their equations are identities)
    std::vector<uint8_t> flag(c->ndof);
    c->bcflag.download(flag.data(), c->ndof, s);
    c->sync();
    Graph g = fine_graph(c->K);
    std::vector<uint8_t> active(c->nnodes, 1);
    std::vector<double> B((size_t)c->nnodes * 18, 0.0);
    for (int64_t n = 0; n < c->nnodes; ++n) {
        const double x = c->h_nodes[3 * n], y = c->h_nodes[3 * n + 1], z = c->h_nodes[3 * n + 2];
        const double rb[3][6] = {{1, 0, 0, 0, z, -y}, {0, 1, 0, -z, 0, x}, {0, 0, 1, y, -x, 0}};
        int nfix = 0;
        for (int r = 0; r < 3; ++r) {
            const bool fixed = flag[3 * n + r] != 0;
            nfix += fixed;
            for (int col = 0; col < 6; ++col) B[(3 * n + r) * 6 + col] = fixed ? 0.0 : rb[r][col];
        }
        active[n] = nfix < 3;
    }
    int bs = 3;
    const int64_t fine_nnz = c->K.nnzb * 9;
    int64_t total_nnz = fine_nnz;
    for (int lev = 0;; ++lev) {
        a->L.emplace_back();
        AmgLevel &L = a->L.back();
        L.bs = bs;
        L.nb = g.nb;
        L.n = (int64_t)bs * g.nb;
        if (lev > 0) {  // coarse operator storage, pattern = previous level's R A P
            L.nnzb = (int64_t)g.idx.size();
            total_nnz += L.nnzb * bs * bs;
            L.h_rowptr.resize(g.nb + 1);
            for (int64_t i = 0; i <= g.nb; ++i) L.h_rowptr[i] = (int32_t)g.ptr[i];
            L.h_colind = g.idx;
            std::vector<int32_t> dpos(g.nb, 0);
            for (int64_t i = 0; i < g.nb; ++i)
                for (int64_t k = g.ptr[i]; k < g.ptr[i + 1]; ++k)
                    if (g.idx[k] == i) dpos[i] = (int32_t)k;
            up(L.rowptr, L.h_rowptr, s);
            up(L.colind, L.h_colind, s);
            up(L.diagpos, dpos, s);
            L.A.alloc(std::max<int64_t>(1, L.nnzb * 36));
        }
        L.x.alloc(L.n); L.b.alloc(L.n); L.r.alloc(L.n); L.d.alloc(L.n); L.t.alloc(L.n);
        L.wr.alloc(L.n); L.we.alloc(L.n);
        L.dinv.alloc((size_t)L.nb * bs * bs);
        if (L.n <= coarse_max || lev + 1 >= max_levels) break;
        std::vector<int32_t> agg;
        const int64_t nagg = aggregate(g, active, agg);
        if (nagg == 0 || nagg * 6 >= L.n) break;  // no useful coarsening: this is the coarsest
        L.nagg = nagg;
        std::vector<double> Pt, Bc;
        tentative(bs, L.nb, nagg, agg, B, Pt, Bc);
        // P pattern: aggregates of the node and its neighbours
        std::vector<int32_t> Pptr, Pidx;
        sorted_union_rows(L.nb, [&](int64_t i, std::vector<int32_t> &o) {
            if (!active[i]) return;
            for (int64_t k = g.ptr[i]; k < g.ptr[i + 1]; ++k)
                if (agg[g.idx[k]] >= 0) o.push_back(agg[g.idx[k]]);
        }, Pptr, Pidx);
        // Q = A P pattern
        std::vector<int32_t> Qptr, Qidx;
        sorted_union_rows(L.nb, [&](int64_t i, std::vector<int32_t> &o) {
            if (!active[i]) return;
            for (int64_t k = g.ptr[i]; k < g.ptr[i + 1]; ++k) {
                const int32_t j = g.idx[k];
                for (int32_t m = Pptr[j]; m < Pptr[j + 1]; ++m) o.push_back(Pidx[m]);
            }
        }, Qptr, Qidx);
        // R = P^T pattern (+ source block of each R block)
        std::vector<int32_t> Rptr(nagg + 1, 0), Ridx(Pidx.size()), Rsrc(Pidx.size());
        for (int32_t J : Pidx) Rptr[J + 1]++;
        for (int64_t J = 0; J < nagg; ++J) Rptr[J + 1] += Rptr[J];
        {
            std::vector<int32_t> fpos(Rptr.begin(), Rptr.end() - 1);
            for (int64_t i = 0; i < L.nb; ++i)
                for (int32_t m = Pptr[i]; m < Pptr[i + 1]; ++m) {
                    const int32_t J = Pidx[m];
                    Ridx[fpos[J]] = (int32_t)i;
                    Rsrc[fpos[J]++] = m;
                }
        }
        // coarse pattern C = R Q
        std::vector<int32_t> Cptr, Cidx;
        sorted_union_rows(nagg, [&](int64_t I, std::vector<int32_t> &o) {
            for (int32_t m = Rptr[I]; m < Rptr[I + 1]; ++m) {
                const int32_t i = Ridx[m];
                for (int32_t q = Qptr[i]; q < Qptr[i + 1]; ++q) o.push_back(Qidx[q]);
            }
        }, Cptr, Cidx);
        std::vector<int32_t> Qrow(Qidx.size()), Crow(Cidx.size());
        for (int64_t i = 0; i < L.nb; ++i)
            for (int32_t q = Qptr[i]; q < Qptr[i + 1]; ++q) Qrow[q] = (int32_t)i;
        for (int64_t I = 0; I < nagg; ++I)
            for (int32_t q = Cptr[I]; q < Cptr[I + 1]; ++q) Crow[q] = (int32_t)I;
        L.nP = (int64_t)Pidx.size();
        L.nQ = (int64_t)Qidx.size();
        L.nR = (int64_t)Ridx.size();
        up(L.agg, agg, s);
        up(L.Pt, Pt, s);
        std::vector<int32_t> Prow(Pidx.size());
        for (int64_t i = 0; i < L.nb; ++i)
            for (int32_t m = Pptr[i]; m < Pptr[i + 1]; ++m) Prow[m] = (int32_t)i;
        up(L.P_rowptr, Pptr, s); up(L.P_colind, Pidx, s); up(L.P_row, Prow, s);
        // chunks of the streamed prolongation (fine level; prolong_rowwise: one thread
        // per row instead): greedy runs of whole rows of at most P_STREAM_CAP blocks
        if (bs == 3 && !c->opt.prolong_rowwise) {
            std::vector<int32_t> ch{0};
            bool fits = true;
            for (int64_t i = 0; i < L.nb; ++i) {
                const int32_t len = Pptr[i + 1] - Pptr[i];
                if (len > P_STREAM_CAP) { fits = false; break; }
                if (Pptr[i + 1] - Pptr[ch.back()] > P_STREAM_CAP) ch.push_back((int32_t)i);
            }
            if (fits) {
                ch.push_back((int32_t)L.nb);
                L.nPchunk = (int64_t)ch.size() - 1;
                up(L.P_chunk, ch, s);
            }
        }
        L.P.alloc(std::max<size_t>(1, Pidx.size() * bs * 6));
        up(L.R_rowptr, Rptr, s); up(L.R_colind, Ridx, s); up(L.R_src, Rsrc, s);
        L.R.alloc(std::max<size_t>(1, Ridx.size() * bs * 6));
        up(L.Q_rowptr, Qptr, s); up(L.Q_colind, Qidx, s); up(L.Q_row, Qrow, s);
        L.Q.alloc(std::max<size_t>(1, Qidx.size() * bs * 6));
        up(L.C_row, Crow, s);
        c->sync();
        if (c->opt.verbose)
            fprintf(stderr, "[amg] level %d: %lld block rows -> %lld aggregates; blocks P %lld, Q = AP %lld, "
                            "coarse %lld\n", lev, (long long)L.nb, (long long)nagg, (long long)L.nP, (long long)L.nQ,
                    (long long)Cidx.size());
        // the next level's graph is the coarse pattern
        Graph gn;
        gn.nb = nagg;
        gn.ptr.assign(Cptr.begin(), Cptr.end());
        gn.idx = std::move(Cidx);
        g = std::move(gn);
        active.assign(nagg, 1);
        B = std::move(Bc);
        bs = 6;
    }
    // power-iteration start vectors (fixed pseudo-random sequence)
    for (auto &L : a->L) {
        std::vector<double> v(L.n);
        uint64_t st = 0x243F6A8885A308D3ull;
        for (int64_t i = 0; i < L.n; ++i) {
            st = st * 6364136223846793005ull + 1442695040888963407ull;
            v[i] = (double)(st >> 11) * 0x1.0p-53 - 0.5;
        }
        L.pw.alloc(L.n);
        L.pw.upload(v.data(), L.n, s);
    }
    a->nc = a->L.back().n;
    a->coarse_inv.alloc((size_t)a->nc * a->nc);
    a->op_complexity = (double)total_nnz / (double)fine_nnz;
    c->st.amg_levels = (int32_t)a->L.size();
    for (size_t l = 0; l < a->L.size() && l < 12; ++l) {
        c->st.amg_ndof[l] = a->L[l].n;
        c->st.amg_nnzb[l] = l == 0 ? c->K.nnzb : a->L[l].nnzb;
    }
    c->st.amg_operator_complexity = a->op_complexity;
    c->sync();
    return a;
}

namespace {
void graph_reset(Amg *a) {
    for (int k = 0; k < 2; ++k)
        if (a->gexec[k]) {
            (void)hipGraphExecDestroy(a->gexec[k]);
            a->gexec[k] = nullptr;
        }
}
}  // namespace

void amg_destroy(Amg *a) {
    if (!a) return;
    graph_reset(a);
    for (hipEvent_t e : a->ev_level)
        if (e) (void)hipEventDestroy(e);
    if (a->ev_rho) (void)hipEventDestroy(a->ev_rho);
    if (a->s_rho) (void)hipStreamDestroy(a->s_rho);
    delete a;
}

// ---------------------------------------------------------------- numeric set-up
namespace {

// rho(D^-1 A) by power iteration, launched on stream s with the reduction scratch
// `partial` and ||D^-1 A x||^2 left in *nrm (device).  The iterate is kept between
// set-ups, so later rebuilds (a slowly changing tangent) start from the previous
// dominant direction and need few iterations; normalisation happens on the device.
void power_iteration(pfem_ctx *c, Amg *a, int l, int iters, hipStream_t s, double *partial, double *nrm) {
    AmgLevel &L = a->L[l];
    double *x = L.pw.p, *y = L.r.p;
    dot_dev<double, double>(L.n, x, x, partial, nrm, s);
    for (int it = 0; it < iters; ++it) {
        scale_by_rnorm(L.n, nrm, x, s);  // ||x|| = 1
        if (l == 0) sell_spmv<3, double, double>(c->K, c->K.v64.p, x, L.t.p, s);
        else bsr_spmv<6, 6, double, double>(L.nb, L.rowptr.p, L.colind.p, L.A.p, x, L.t.p, 0.0, s);
        if (L.bs == 3) block_diag_apply<3, double, double>(L.nb, L.dinv.p, L.t.p, y, 1.0, s);
        else block_diag_apply<6, double, double>(L.nb, L.dinv.p, L.t.p, y, 1.0, s);
        dot_dev<double, double>(L.n, y, y, partial, nrm, s);  // ||D^-1 A x||^2
        PF_HIP(hipMemcpyAsync(x, y, L.n * sizeof(double), hipMemcpyDeviceToDevice, s));
    }
    L.pw_warm = true;
}
double rho_of(double nrm2) {
    const double lam = std::sqrt(nrm2);
    return lam > 0 ? lam : 1.0;
}
double spectral_radius(pfem_ctx *c, Amg *a, int l, int iters) {
    double *nrm = c->scal.p + 10;
    power_iteration(c, a, l, iters, c->stream, c->partial.p, nrm);
    return rho_of(c->sync_read(nrm));
}
int rho_iters(pfem_ctx *c, const AmgLevel &L) {
    // power iteration converges from below, slowly where the top of the spectrum is
    // crowded (12 steps gave 0.91 rho on a notched bar): 25 steps cold, 4 warm
    const int cold = c->opt.rho_iters_cold > 0 ? c->opt.rho_iters_cold : 25;
    const int warm = c->opt.rho_iters_warm > 0 ? c->opt.rho_iters_warm : 4;
    return L.pw_warm ? warm : cold;
}

// Dense inverse of the coarsest operator, computed on the device.
void coarse_inverse(pfem_ctx *c, Amg *a) {
    AmgLevel &L = a->L.back();
    const int64_t n = L.n;
    if (!a->gj_work.p) {
        a->gj_work.alloc((size_t)2 * n * n);
        a->gj_prow.alloc(2 * n);
        a->gj_fcol.alloc(n);
        a->gj_iw.alloc(gj_scratch_ints(n));
        a->gj_dw.alloc(gj_scratch_doubles(n));
    }
    if (a->L.size() == 1)
        dense_coarse_inverse(L.nb, 3, true, c->K.slice_ptr.p, c->K.rowlen.p, nullptr, c->K.colind.p, c->K.v64.p,
                             a->gj_work.p, a->gj_prow.p, a->gj_fcol.p, a->gj_iw.p, a->gj_dw.p, a->coarse_inv.p,
                             c->stream);
    else
        dense_coarse_inverse(L.nb, 6, false, nullptr, nullptr, L.rowptr.p, L.colind.p, L.A.p, a->gj_work.p,
                             a->gj_prow.p, a->gj_fcol.p, a->gj_iw.p, a->gj_dw.p, a->coarse_inv.p, c->stream);
}

}  // namespace

void amg_numeric_setup(pfem_ctx *c, Amg *a, bool keep_P) {
    graph_reset(a);  // smoother coefficients change with rho: captured V-cycles are stale
    if (!keep_P)  // new prolongators: stashed coarse levels no longer fit
        for (AmgStash &st : a->stash) st.ok = false;
    hipStream_t s = c->stream;
    const bool want32 = c->opt.precision == PFEM_PREC_FP32_IR;
    const size_t nlev = a->L.size();
    // With the prolongators kept, rho only feeds the smoothers: each level's power
    // iteration runs on s_rho as soon as its operator and block-diagonal inverse exist,
    // beside the Galerkin products of the next level on the context's stream (the same
    // kernels in the same order per stream, so the hierarchy is bit for bit the one the
    // sequential set-up builds).  A rebuilt P needs rho for its smoothing weight first.
    const bool overlap = keep_P && nlev > 1;
    if (overlap && !a->s_rho) {
        PF_HIP(hipStreamCreateWithFlags(&a->s_rho, hipStreamNonBlocking));
        PF_HIP(hipEventCreateWithFlags(&a->ev_rho, hipEventDisableTiming));
        a->ev_level.assign(nlev, nullptr);
        for (auto &e : a->ev_level) PF_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        a->rho_partial.alloc(c->partial.n);
        a->rho_scal.alloc(nlev);
    }
    double t_rho = 0.0;
    const double t_begin = now();
    for (size_t l = 0; l < nlev; ++l) {
        AmgLevel &L = a->L[l];
        const bool fine = (l == 0);
        const double *Av = fine ? c->K.v64.p : L.A.p;
        if (L.bs == 3) amg_block_inverse<3>(L.nb, fine ? c->K.diagpos.p : L.diagpos.p, Av, L.dinv.p, s);
        else amg_block_inverse<6>(L.nb, L.diagpos.p, Av, L.dinv.p, s);
        if (l + 1 == nlev) break;  // coarsest: dense solve
        if (overlap) {
            PF_HIP(hipEventRecord(a->ev_level[l], s));
            PF_HIP(hipStreamWaitEvent(a->s_rho, a->ev_level[l], 0));
            power_iteration(c, a, (int)l, rho_iters(c, L), a->s_rho, a->rho_partial.p, a->rho_scal.p + l);
        } else {
            c->sync();
            const double tr = now();
            L.rho = spectral_radius(c, a, (int)l, rho_iters(c, L));
            t_rho += now() - tr;
        }
        AmgLevel &N = a->L[l + 1];
        if (L.bs == 3) {
            if (!keep_P) {
                const double omega = (4.0 / 3.0) / L.rho;
                amg_smooth_P<3, true>(L.nP, L.P_row.p, L.P_colind.p, c->K.slice_ptr.p, c->K.rowlen.p,
                                      nullptr, c->K.colind.p, Av, L.agg.p, L.Pt.p, L.dinv.p, omega, L.P.p, s);
                amg_transpose_P<3>(L.nR, L.R_src.p, L.P.p, L.R.p, s);
            }
            amg_AP<3, true>(L.nQ, L.Q_row.p, L.Q_colind.p, c->K.slice_ptr.p, c->K.rowlen.p, nullptr,
                            c->K.colind.p, Av, L.P_rowptr.p, L.P_colind.p, L.P.p, L.Q.p, s);
            amg_RAP<3>(N.nnzb, L.C_row.p, N.colind.p, L.R_rowptr.p, L.R_colind.p, L.R.p, L.Q_rowptr.p, L.Q_colind.p,
                       L.Q.p, N.A.p, s);
        } else {
            if (!keep_P) {
                const double omega = (4.0 / 3.0) / L.rho;
                amg_smooth_P<6, false>(L.nP, L.P_row.p, L.P_colind.p, nullptr, nullptr, L.rowptr.p,
                                       L.colind.p, Av, L.agg.p, L.Pt.p, L.dinv.p, omega, L.P.p, s);
                amg_transpose_P<6>(L.nR, L.R_src.p, L.P.p, L.R.p, s);
            }
            amg_AP<6, false>(L.nQ, L.Q_row.p, L.Q_colind.p, nullptr, nullptr, L.rowptr.p,
                             L.colind.p, Av, L.P_rowptr.p, L.P_colind.p, L.P.p, L.Q.p, s);
            amg_RAP<6>(N.nnzb, L.C_row.p, N.colind.p, L.R_rowptr.p, L.R_colind.p, L.R.p, L.Q_rowptr.p, L.Q_colind.p,
                       L.Q.p, N.A.p, s);
        }
    }
    c->sync();
    const double t_coarse0 = now();
    coarse_inverse(c, a);
    c->sync();
    c->st.t_setup_coarse += now() - t_coarse0;
    if (overlap) {  // the power iterations ran beside; collect their results
        const double tr = now();
        PF_HIP(hipEventRecord(a->ev_rho, a->s_rho));
        PF_HIP(hipStreamWaitEvent(s, a->ev_rho, 0));
        std::vector<double> nrm(nlev, 0.0);
        PF_HIP(hipMemcpyAsync(nrm.data(), a->rho_scal.p, (nlev - 1) * sizeof(double), hipMemcpyDeviceToHost,
                              a->s_rho));
        PF_HIP(hipStreamSynchronize(a->s_rho));
        for (size_t l = 0; l + 1 < nlev; ++l) a->L[l].rho = rho_of(nrm[l]);
        t_rho += now() - tr;
    }
    c->st.t_setup_rho += t_rho;
    c->st.t_setup_galerkin += t_coarse0 - t_begin - (overlap ? 0.0 : t_rho);
    if (want32) {
        for (size_t l = 0; l < a->L.size(); ++l) {
            AmgLevel &L = a->L[l];
            if (!L.x32.p) {
                L.x32.alloc(L.n); L.b32.alloc(L.n); L.r32.alloc(L.n); L.d32.alloc(L.n); L.t32.alloc(L.n);
                L.wr32.alloc(L.n); L.we32.alloc(L.n);
            }
            if (!L.dinv32.p) L.dinv32.alloc(L.dinv.n);
            convert<double, float>((int64_t)L.dinv.n, L.dinv.p, L.dinv32.p, s);
            if (l > 0) {
                if (!L.A32.p) L.A32.alloc(L.A.n);
                convert<double, float>((int64_t)L.A.n, L.A.p, L.A32.p, s);
            }
            if (l + 1 < a->L.size() && !(keep_P && L.P32.p)) {
                if (!L.P32.p) { L.P32.alloc(L.P.n); L.R32.alloc(L.R.n); }
                convert<double, float>((int64_t)L.P.n, L.P.p, L.P32.p, s);
                convert<double, float>((int64_t)L.R.n, L.R.p, L.R32.p, s);
            }
        }
        a->have32 = true;
    }
    c->sync();
}

namespace {
template <class T>
void copy_dev(dbuf<T> &dst, const dbuf<T> &src, hipStream_t s) {
    if (dst.n != src.n) dst.alloc(src.n);
    if (src.n) PF_HIP(hipMemcpyAsync(dst.p, src.p, src.n * sizeof(T), hipMemcpyDeviceToDevice, s));
}
}  // namespace

AmgStash &amg_stash_slot(Amg *a, int slot) {
    if ((int)a->stash.size() <= slot) a->stash.resize(slot + 1);
    return a->stash[slot];
}

void amg_stash_coarse(pfem_ctx *c, Amg *a, int slot) {
    hipStream_t s = c->stream;
    AmgStash &st = amg_stash_slot(a, slot);
    st.lv.resize(a->L.size());
    for (size_t l = 1; l < a->L.size(); ++l) {
        AmgLevel &L = a->L[l];
        AmgStash::Level &S = st.lv[l];
        copy_dev(S.A, L.A, s);
        copy_dev(S.dinv, L.dinv, s);
        if (a->have32) {
            copy_dev(S.A32, L.A32, s);
            copy_dev(S.dinv32, L.dinv32, s);
        }
        S.rho = L.rho;
    }
    copy_dev(st.coarse_inv, a->coarse_inv, s);
    st.ok = true;
    st.have32 = a->have32;
}

bool amg_restore_coarse(pfem_ctx *c, Amg *a, int slot) {
    if ((int)a->stash.size() <= slot) return false;
    AmgStash &st = a->stash[slot];
    if (!st.ok || (a->have32 && !st.have32) || st.lv.size() != a->L.size()) return false;
    hipStream_t s = c->stream;
    for (size_t l = 1; l < a->L.size(); ++l) {
        AmgLevel &L = a->L[l];
        AmgStash::Level &S = st.lv[l];
        copy_dev(L.A, S.A, s);
        copy_dev(L.dinv, S.dinv, s);
        if (a->have32) {
            copy_dev(L.A32, S.A32, s);
            copy_dev(L.dinv32, S.dinv32, s);
        }
        L.rho = S.rho;
    }
    copy_dev(a->coarse_inv, st.coarse_inv, s);
    amg_refresh_fine(c, a);  // resets the captured cycles too
    return true;
}

// Fine-level refresh: the level-0 smoother follows the current K (block-diagonal inverse,
// warm-started rho), the interpolation and the coarse operators stay as last built.
void amg_refresh_fine(pfem_ctx *c, Amg *a) {
    graph_reset(a);
    hipStream_t s = c->stream;
    AmgLevel &L = a->L[0];
    amg_block_inverse<3>(L.nb, c->K.diagpos.p, c->K.v64.p, L.dinv.p, s);
    if (a->L.size() > 1) {
        const double tr = now();
        const int warm = c->opt.rho_iters_warm > 0 ? c->opt.rho_iters_warm : 4;
        L.rho = spectral_radius(c, a, 0, warm);
        c->st.t_setup_rho += now() - tr;
    } else {
        coarse_inverse(c, a);
    }
    if (a->have32 && L.dinv32.p) convert<double, float>((int64_t)L.dinv.n, L.dinv.p, L.dinv32.p, s);
    c->sync();
}

// ---------------------------------------------------------------- V-cycle
namespace {

template <class T> struct Vec;
template <> struct Vec<double> {
    static double *x(AmgLevel &L) { return L.x.p; }
    static double *b(AmgLevel &L) { return L.b.p; }
    static double *r(AmgLevel &L) { return L.r.p; }
    static double *d(AmgLevel &L) { return L.d.p; }
    static double *t(AmgLevel &L) { return L.t.p; }
    static double *wr(AmgLevel &L) { return L.wr.p; }
    static double *we(AmgLevel &L) { return L.we.p; }
    static const double *A(AmgLevel &L) { return L.A.p; }
    static const double *P(AmgLevel &L) { return L.P.p; }
    static const double *R(AmgLevel &L) { return L.R.p; }
    static const double *Dinv(AmgLevel &L) { return L.dinv.p; }
    static const double *K(pfem_ctx *c) { return c->K.v64.p; }
};
template <> struct Vec<float> {
    static float *x(AmgLevel &L) { return L.x32.p; }
    static float *b(AmgLevel &L) { return L.b32.p; }
    static float *r(AmgLevel &L) { return L.r32.p; }
    static float *d(AmgLevel &L) { return L.d32.p; }
    static float *t(AmgLevel &L) { return L.t32.p; }
    static float *wr(AmgLevel &L) { return L.wr32.p; }
    static float *we(AmgLevel &L) { return L.we32.p; }
    static const float *A(AmgLevel &L) { return L.A32.p; }
    static const float *P(AmgLevel &L) { return L.P32.p; }
    static const float *R(AmgLevel &L) { return L.R32.p; }
    static const float *Dinv(AmgLevel &L) { return L.dinv32.p; }
    static const float *K(pfem_ctx *c) { return c->K.v32.p; }
};

// The fine-level products inside the FP32 V-cycle's smoother use the FP16 scaled copy
// when it is present; the residual restricted to the coarse level (vcycle) stays FP32.
template <class T>
void level_matvec(pfem_ctx *c, AmgLevel &L, int l, const T *x, T *y) {
    if (l == 0) {
        if constexpr (std::is_same<T, float>::value)
            if (c->half_fine) return spmv_fine_half(c->K, x, y, c->stream);
        sell_spmv<3, T, T>(c->K, Vec<T>::K(c), x, y, c->stream);
    } else {
        bsr_spmv<6, 6, T, T>(L.nb, L.rowptr.p, L.colind.p, Vec<T>::A(L), x, y, 0.0, c->stream);
    }
}
template <class T>
void level_residual(pfem_ctx *c, AmgLevel &L, int l, const T *x, const T *b, T *r, bool in_smoother = true) {
    if (l == 0) {
        if constexpr (std::is_same<T, float>::value)
            if (c->half_fine && in_smoother) return residual_fine_half(c->K, x, b, r, c->stream);
        sell_residual<3, T, T>(c->K, Vec<T>::K(c), x, b, r, c->stream);
    } else {
        bsr_residual<6, T, T>(L.nb, L.rowptr.p, L.colind.p, Vec<T>::A(L), x, b, r, c->stream);
    }
}
// One Chebyshev step after the first: t = A d, r -= t, d = c1 d + c2 Dinv r, x += d.  On
// the fine level one fused launch (sell_cheb) unless the FP16 copy serves the smoother's
// products or smoother_unfused is set; `last` (neither r nor d read again before the
// next smoother overwrites them) lets the fused launch skip their stores.
template <class T>
void cheb_update(pfem_ctx *c, AmgLevel &L, int l, T *&d, T *&t, T *r, T *x, double c1, double c2, bool last) {
    bool fused = l == 0 && !c->opt.smoother_unfused;
    if constexpr (std::is_same<T, float>::value) fused = fused && !c->half_fine;
    if (fused) {
        sell_cheb<T, T>(c->K, Vec<T>::K(c), d, Vec<T>::Dinv(L), r, last ? nullptr : t, x, c1, c2, c->stream);
        if (!last) std::swap(d, t);  // the new direction is in t; the old one is free
        return;
    }
    if (l > 0 && L.bs == 6 && !c->opt.smoother_unfused) {  // coarse level: the same fusion
        bsr_cheb<6, T, T>(L.nb, L.rowptr.p, L.colind.p, Vec<T>::A(L), d, Vec<T>::Dinv(L), r, last ? nullptr : t, x,
                          c1, c2, c->stream);
        if (!last) std::swap(d, t);
        return;
    }
    level_matvec<T>(c, L, l, d, t);
    if (L.bs == 3) cheb_step<3, T, T>(L.nb, Vec<T>::Dinv(L), t, r, d, x, c1, c2, c->stream);
    else cheb_step<6, T, T>(L.nb, Vec<T>::Dinv(L), t, r, d, x, c1, c2, c->stream);
}
template <class T>
void diag_apply(AmgLevel &L, const T *r, T *z, double scale, hipStream_t s) {
    if (L.bs == 3) block_diag_apply<3, T, T>(L.nb, Vec<T>::Dinv(L), r, z, scale, s);
    else block_diag_apply<6, T, T>(L.nb, Vec<T>::Dinv(L), r, z, scale, s);
}
// d = scale Dinv v and, from a zero guess (v = b), r = b and x = d; else (v = r) x += d
template <class T>
void diag_first_step(AmgLevel &L, const T *v, T *r, T *d, T *x, double scale, bool zero_init, hipStream_t s) {
    if (L.bs == 3) diag_first<3, T, T>(L.nb, Vec<T>::Dinv(L), v, r, d, x, scale, zero_init, s);
    else diag_first<6, T, T>(L.nb, Vec<T>::Dinv(L), v, r, d, x, scale, zero_init, s);
}

// Polynomial smoother on level l: x <- x + p(D^-1 A) D^-1 (b - A x).
// Chebyshev on [lmax/30, lmax] with lmax = 1.1 rho (Adams, Brezina, Hu & Tuminaro,
// J. Comput. Phys. 188, 2003), fourth-kind Chebyshev on (0, lmax], or damped block
// Jacobi with w = 4/(3 rho).
template <class T>
void smooth(pfem_ctx *c, Amg *a, int l, const T *b, T *x, bool zero_init) {
    AmgLevel &L = a->L[l];
    hipStream_t s = c->stream;
    const int deg = a->degree_at(l);
    T *r = Vec<T>::r(L), *d = Vec<T>::d(L), *t = Vec<T>::t(L);
    if (a->smoother == PFEM_SMOOTH_JACOBI) {
        const double w = (4.0 / 3.0) / L.rho;
        for (int it = 0; it < deg; ++it) {
            if (zero_init && it == 0) {
                diag_apply<T>(L, b, x, w, s);
                continue;
            }
            level_residual<T>(c, L, l, x, b, r);
            diag_first_step<T>(L, r, nullptr, d, x, w, false, s);
        }
        return;
    }
    if (a->smoother == PFEM_SMOOTH_CHEBYSHEV4) {
        // fourth-kind Chebyshev (Lottes, Numer. Linear Algebra Appl. 30, 2023, Alg. 1):
        // d_0 = 4/(3 lmax) D^-1 r_0, d_i = (2i-1)/(2i+3) d_{i-1} + (8i+4)/((2i+3) lmax) D^-1 r_i
        const double lam = 1.1 * L.rho;
        if (!zero_init) level_residual<T>(c, L, l, x, b, r);
        diag_first_step<T>(L, zero_init ? b : r, zero_init ? r : nullptr, d, x, 4.0 / (3.0 * lam), zero_init, s);
        for (int i = 1; i < deg; ++i) {
            const double c1 = (2.0 * i - 1.0) / (2.0 * i + 3.0), c2 = (8.0 * i + 4.0) / ((2.0 * i + 3.0) * lam);
            cheb_update<T>(c, L, l, d, t, r, x, c1, c2, i + 1 == deg);
        }
        return;
    }
    const double lmax = 1.1 * L.rho, lmin = lmax / 30.0;
    const double theta = 0.5 * (lmax + lmin), delta = 0.5 * (lmax - lmin);
    const double sigma = theta / delta;
    double rho_old = 1.0 / sigma;
    if (!zero_init) level_residual<T>(c, L, l, x, b, r);
    diag_first_step<T>(L, zero_init ? b : r, zero_init ? r : nullptr, d, x, 1.0 / theta, zero_init, s);
    for (int k = 2; k <= deg; ++k) {
        const double rho_new = 1.0 / (2.0 * sigma - rho_old);
        const double c1 = rho_new * rho_old, c2 = 2.0 * rho_new / delta;
        cheb_update<T>(c, L, l, d, t, r, x, c1, c2, k == deg);
        rho_old = rho_new;
    }
}

template <class T>
void vcycle(pfem_ctx *c, Amg *a, int l, const T *b, T *x) {
    AmgLevel &L = a->L[l];
    hipStream_t s = c->stream;
    if ((size_t)l + 1 == a->L.size()) {  // coarsest: FP64 dense solve
        dense_gemv<T>(a->nc, a->coarse_inv.p, b, x, s);
        return;
    }
    AmgLevel &N = a->L[l + 1];
    smooth<T>(c, a, l, b, x, true);
    level_residual<T>(c, L, l, x, b, Vec<T>::r(L), false);
    if (L.bs == 3) bsr_spmv_wave<6, 3, T, T>(L.nagg, L.R_rowptr.p, L.R_colind.p, Vec<T>::R(L), Vec<T>::r(L), Vec<T>::b(N), s);
    else bsr_spmv_wave<6, 6, T, T>(L.nagg, L.R_rowptr.p, L.R_colind.p, Vec<T>::R(L), Vec<T>::r(L), Vec<T>::b(N), s);
    vcycle<T>(c, a, l + 1, Vec<T>::b(N), Vec<T>::x(N));
    if (a->wcycle && (size_t)l + 2 < a->L.size()) {  // W-cycle: a second coarse cycle on its residual
        level_residual<T>(c, N, l + 1, Vec<T>::x(N), Vec<T>::b(N), Vec<T>::wr(N), false);
        vcycle<T>(c, a, l + 1, Vec<T>::wr(N), Vec<T>::we(N));
        axpby<T>(N.n, 1.0, Vec<T>::we(N), 1.0, Vec<T>::x(N), s);
    }
    if (L.bs == 3 && L.nPchunk > 0)
        bsr_spmv_stream<3, 6, T, T>(L.nPchunk, L.P_chunk.p, L.P_rowptr.p, L.P_colind.p, Vec<T>::P(L), Vec<T>::x(N), x,
                                    1.0, s);
    else if (L.bs == 3) bsr_spmv<3, 6, T, T>(L.nb, L.P_rowptr.p, L.P_colind.p, Vec<T>::P(L), Vec<T>::x(N), x, 1.0, s);
    else bsr_spmv<6, 6, T, T>(L.nb, L.P_rowptr.p, L.P_colind.p, Vec<T>::P(L), Vec<T>::x(N), x, 1.0, s);
    smooth<T>(c, a, l, b, x, false);
}

}  // namespace

namespace {
// The V-cycle is a fixed sequence of ~50 launches (most of them small, on the coarse
// levels) between rebuilds: capture it once on the context's stream and replay it as
// one graph launch.  vcycle_eager runs it eagerly instead.  It records kernels and
// device-to-device copies only (no memset nodes).
template <class T>
void apply_vcycle(pfem_ctx *c, Amg *a, int k, const T *r, T *z) {
    if (c->opt.vcycle_eager) {
        vcycle<T>(c, a, 0, r, z);
        return;
    }
    if (a->gexec[k] && (a->g_in[k] != r || a->g_out[k] != z)) {
        (void)hipGraphExecDestroy(a->gexec[k]);
        a->gexec[k] = nullptr;
    }
    if (!a->gexec[k]) {
        hipGraph_t g = nullptr;
        PF_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed));
        try {
            vcycle<T>(c, a, 0, r, z);
        } catch (...) {
            (void)hipStreamEndCapture(c->stream, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        PF_HIP(hipStreamEndCapture(c->stream, &g));
        const hipError_t e = hipGraphInstantiate(&a->gexec[k], g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        PF_HIP(e);
        a->g_in[k] = r;
        a->g_out[k] = z;
    }
    PF_HIP(hipGraphLaunch(a->gexec[k], c->stream));
}
}  // namespace

void amg_apply64(pfem_ctx *c, Amg *a, const double *r, double *z) { apply_vcycle<double>(c, a, 0, r, z); }
void amg_apply32(pfem_ctx *c, Amg *a, const float *r, float *z) { apply_vcycle<float>(c, a, 1, r, z); }

// ---------------------------------------------------------------- measurement
// Algorithmic bytes of the cycle's launches, each counted with its operands read once
// and its outputs written once (s = bytes per value; integer indices 4 bytes).  The
// FP16 smoother copy of the fine K (fp16_fine) is not counted.
namespace {
double op_bytes(pfem_ctx *c, const AmgLevel &L, int l, double s) {
    return l == 0 ? (double)c->K.nnzb * (9.0 * s + 4.0) + 4.0 * (L.nb + 1)
                  : (double)L.nnzb * (36.0 * s + 4.0) + 4.0 * (L.nb + 1);
}
double smooth_bytes(pfem_ctx *c, Amg *a, int l, double s, bool zero_init) {
    const AmgLevel &L = a->L[l];
    const double A = op_bytes(c, L, l, s), v = (double)L.n * s, D = (double)L.nb * L.bs * L.bs * s;
    const bool fused = l == 0 ? !c->opt.smoother_unfused && !(s == 4.0 && c->half_fine)
                              : L.bs == 6 && !c->opt.smoother_unfused;
    const int deg = a->degree_at(l);
    if (a->smoother == PFEM_SMOOTH_JACOBI) {
        double b = 0;
        for (int it = 0; it < deg; ++it)
            b += (zero_init && it == 0) ? 2 * v + D : (A + 3 * v) + (4 * v + D);
        return b;
    }
    double b = zero_init ? 4 * v + D : (A + 3 * v) + (4 * v + D);  // first step
    for (int k = 2; k <= deg; ++k)
        b += fused ? A + (k == deg ? 4 : 6) * v + D : (A + 2 * v) + (7 * v + D);
    return b;
}
double cycle_bytes(pfem_ctx *c, Amg *a, int l, double s) {
    const AmgLevel &L = a->L[l];
    if ((size_t)l + 1 == a->L.size()) return (double)a->nc * a->nc * 8.0 + 2.0 * a->nc * s;
    const AmgLevel &N = a->L[l + 1];
    const double v = (double)L.n * s, vN = (double)N.n * s;
    double b = smooth_bytes(c, a, l, s, true) + op_bytes(c, L, l, s) + 3 * v;        // pre-smooth, residual
    b += (double)L.nR * (6.0 * L.bs * s + 4.0) + 4.0 * (L.nagg + 1) + v + vN;          // restriction
    b += cycle_bytes(c, a, l + 1, s);
    if (a->wcycle && (size_t)l + 2 < a->L.size())
        b += op_bytes(c, N, l + 1, s) + 3 * vN + cycle_bytes(c, a, l + 1, s) + 3 * vN;
    b += (double)L.nP * (6.0 * L.bs * s + 4.0) + 4.0 * (L.nb + 1) + vN + 2 * v;        // prolongation
    return b + smooth_bytes(c, a, l, s, false);                                         // post-smooth
}
}  // namespace

void amg_bench(pfem_ctx *c, Amg *a, int prec, int reps, double *sec_cycle, double *bytes_cycle, double *sec_step,
               double *bytes_step) {
    hipStream_t s = c->stream;
    AmgLevel &L = a->L[0];
    const int64_t n = L.n;
    const double vs = prec == 1 ? 4.0 : 8.0;
    float ms_cycle = 0, ms_step = 0;
    // the cycle maps the level-0 W-cycle vectors (used on coarser levels only) to each other
    auto run = [&](auto *in, auto *out, auto *d, auto *r, auto *x, auto *t, const auto *Kv, const auto *dinv) {
        using T = std::remove_const_t<std::remove_pointer_t<decltype(in)>>;
        fill<T>(n, T(1), in, s);
        fill<T>(n, T(1), d, s);
        fill<T>(n, T(1), r, s);
        fill<T>(n, T(0), x, s);
        apply_vcycle<T>(c, a, prec == 1 ? 1 : 0, in, out);  // capture (graph) and warm caches
        PF_HIP(hipEventRecord(c->ev0, s));
        for (int i = 0; i < reps; ++i) apply_vcycle<T>(c, a, prec == 1 ? 1 : 0, in, out);
        PF_HIP(hipEventRecord(c->ev1, s));
        c->sync();
        PF_HIP(hipEventElapsedTime(&ms_cycle, c->ev0, c->ev1));
        // one fused fine-level Chebyshev step (product + update, r and d stored)
        sell_cheb<T, T>(c->K, Kv, d, dinv, r, t, x, 0.5, 1e-3, s);
        PF_HIP(hipEventRecord(c->ev0, s));
        for (int i = 0; i < reps; ++i) sell_cheb<T, T>(c->K, Kv, d, dinv, r, t, x, 0.5, 1e-3, s);
        PF_HIP(hipEventRecord(c->ev1, s));
        c->sync();
        PF_HIP(hipEventElapsedTime(&ms_step, c->ev0, c->ev1));
    };
    if (prec == 1)
        run(L.wr32.p, L.we32.p, L.d32.p, L.r32.p, L.x32.p, L.t32.p, (const float *)c->K.v32.p,
            (const float *)L.dinv32.p);
    else
        run(L.wr.p, L.we.p, L.d.p, L.r.p, L.x.p, L.t.p, (const double *)c->K.v64.p, (const double *)L.dinv.p);
    *sec_cycle = 1e-3 * ms_cycle / reps;
    *sec_step = 1e-3 * ms_step / reps;
    *bytes_cycle = cycle_bytes(c, a, 0, vs);
    *bytes_step = op_bytes(c, L, 0, vs) + 6.0 * n * vs + (double)L.nb * 9 * vs;
}

}  // namespace pfem
