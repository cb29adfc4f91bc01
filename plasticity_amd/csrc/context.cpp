// context.cpp -- pfem_ctx set-up and the C-ABI entry points (include/pfem.h).
//
// Model set-up (Model.jl:63-75): mesh upload, node adjacency and the 3x3 block
// pattern (Assembly.jl:62-162, stored as SELL-64 BSR), greedy element colouring
// (first fit in element order; equals the parity colours of Assembly.jl:172-188
// on box meshes), and the per-element scatter slot map that replaces the
// reference's 576 binary searches per element (Assembly.jl:210-224, :401-409).
#include "context.h"
#include "build.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <unordered_map>

using namespace pfem;

namespace pfem {

int make_devmat(const pfem_material &m, DevMat &d) {
    // J2Material guards (Materials.jl:90-103)
    if (!(m.E > 0) || !(-1 < m.nu && m.nu < 0.5) || !(m.sy0 > 0) || !(m.Hiso >= 0) || !(m.Hkin >= 0) ||
        !(m.ssat >= m.sy0) || !(m.delta >= 0))
        return PFEM_ERR_BAD_ARG;
    if (m.delta > 0 && m.ssat > m.sy0 && m.Hkin > 0) return PFEM_ERR_BAD_ARG;
    d.E = m.E; d.nu = m.nu; d.sy0 = m.sy0; d.Hiso = m.Hiso; d.Hkin = m.Hkin; d.ssat = m.ssat;
    d.delta = m.delta;
    d.G = m.E / (2 * (1 + m.nu));                       // Materials.jl:104
    d.K = m.E / (3 * (1 - 2 * m.nu));                   // :105
    d.lam = m.E * m.nu / ((1 + m.nu) * (1 - 2 * m.nu)); // :106
    return PFEM_OK;
}

}  // namespace pfem

double pfem_ctx::sync_read(const double *dptr) {
    PF_HIP(hipMemcpyAsync(h_scal, dptr, sizeof(double), hipMemcpyDeviceToHost, stream));
    sync();
    return *h_scal;
}

namespace {

thread_local std::string g_last_error;

template <class F>
int guarded(pfem_fail *fail, F &&f) {
    try {
        return f();
    } catch (const hip_failure &e) {
        g_last_error = e.what();
        if (fail) fail->hip = (int32_t)e.code;
        return PFEM_ERR_HIP;
    } catch (const status_failure &e) {
        g_last_error = e.what();
        return e.status;
    } catch (const std::bad_alloc &) {
        g_last_error = "host allocation failed";
        return PFEM_ERR_BAD_ARG;
    } catch (const std::exception &e) {
        g_last_error = e.what();
        return PFEM_ERR_BAD_ARG;
    }
}

// Node adjacency: sorted unique neighbour lists (Assembly.jl:62-82), CSR.
void node_adjacency(int64_t nn, int64_t ne, const int64_t *conn, std::vector<int64_t> &ptr,
                    std::vector<int32_t> &idx) {
    std::vector<int64_t> cnt(nn + 1, 0);
    for (int64_t e = 0; e < ne; ++e)
        for (int a = 0; a < 8; ++a) cnt[conn[8 * e + a] + 1] += 8;
    for (int64_t n = 0; n < nn; ++n) cnt[n + 1] += cnt[n];
    std::vector<int32_t> tmp(cnt[nn]);
    std::vector<int64_t> fill(cnt.begin(), cnt.end() - 1);
    for (int64_t e = 0; e < ne; ++e)
        for (int a = 0; a < 8; ++a) {
            const int64_t na = conn[8 * e + a];
            for (int b = 0; b < 8; ++b) tmp[fill[na]++] = (int32_t)conn[8 * e + b];
        }
    ptr.assign(nn + 1, 0);
    for (int64_t n = 0; n < nn; ++n) {
        auto b = tmp.begin() + cnt[n], en = tmp.begin() + cnt[n + 1];
        std::sort(b, en);
        ptr[n + 1] = ptr[n] + (std::unique(b, en) - b);
    }
    idx.resize(ptr[nn]);
    for (int64_t n = 0; n < nn; ++n)
        std::copy(tmp.begin() + cnt[n], tmp.begin() + cnt[n] + (ptr[n + 1] - ptr[n]), idx.begin() + ptr[n]);
}

}  // namespace

extern "C" {

const char *pfem_status_string(int s) {
    switch (s) {
        case PFEM_OK: return "ok";
        case PFEM_ERR_INVERTED: return "inverted element (det F <= 0)";
        case PFEM_ERR_NOT_CONVERGED: return "Krylov solve did not converge";
        case PFEM_ERR_NON_SPD: return "Krylov breakdown (non-SPD operator or preconditioner)";
        case PFEM_ERR_UNSUPPORTED: return "unsupported material/element branch";
        case PFEM_ERR_HIP: return g_last_error.empty() ? "HIP error" : g_last_error.c_str();
        case PFEM_ERR_BAD_ARG: return g_last_error.empty() ? "bad argument" : g_last_error.c_str();
        case PFEM_ERR_NO_DEVICE: return "no GPU device";
        default: return "unknown status";
    }
}

void pfem_default_options(pfem_options *o) {
    std::memset(o, 0, sizeof *o);
    o->precision = PFEM_PREC_FP64;
    o->smoother = PFEM_SMOOTH_CHEBYSHEV;
    o->smoother_degree = 0;
    o->cg_itmax = 200;      // Solver.jl:104
    o->eta_min = 1e-8;      // :105
    o->eta_max = 0.1;
    o->ew_gamma = 0.9;      // :106
    o->ew_alpha = 1.5;
    o->coarse_max = 1500;
    o->max_levels = 10;
    o->verbose = 0;
    o->amg_rebuild = PFEM_AMG_KEEP_P_UNTIL_FALLBACK;
    o->split_operator = 1;  // FP32-IR: K p from the FP32 + BF16 split copy of K
}

int pfem_create(const pfem_mesh_desc *md, const pfem_options *opts, pfem_ctx **out) {
    if (!md || !out || md->nnodes <= 0 || md->nelem <= 0 || !md->nodes || !md->conn) return PFEM_ERR_BAD_ARG;
    if (md->nmaterials < 1 || !md->materials) return PFEM_ERR_BAD_ARG;
    if (md->element_kind != PFEM_HEX8_FINITE && md->element_kind != PFEM_HEX8_FINITE_FBAR &&
        md->element_kind != PFEM_HEX8_SMALL)
        return PFEM_ERR_BAD_ARG;
    *out = nullptr;
    std::vector<DevMat> dm(md->nmaterials);
    for (int i = 0; i < md->nmaterials; ++i) {
        int rc = make_devmat(md->materials[i], dm[i]);
        if (rc) return rc;
    }
    for (int64_t k = 0; k < 8 * md->nelem; ++k)
        if (md->conn[k] < 0 || md->conn[k] >= md->nnodes) return PFEM_ERR_BAD_ARG;
    // eight distinct nodes per element: the scatter's slots (one per node pair) and its
    // first-touch stores assume that no two lanes of an element share a K block
    for (int64_t e = 0; e < md->nelem; ++e)
        for (int a = 1; a < 8; ++a)
            for (int b = 0; b < a; ++b)
                if (md->conn[8 * e + a] == md->conn[8 * e + b]) {
                    g_last_error = "element " + std::to_string(e) + " repeats node " + std::to_string(md->conn[8 * e + a]);
                    return PFEM_ERR_BAD_ARG;
                }
    if (md->material_id)
        for (int64_t e = 0; e < md->nelem; ++e)
            if (md->material_id[e] < 0 || md->material_id[e] >= md->nmaterials) return PFEM_ERR_BAD_ARG;
    int ndev = 0;  // argument checks first: they need no device
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return PFEM_ERR_NO_DEVICE;
    auto *c = new pfem_ctx();
    int rc = guarded(nullptr, [&]() -> int {
        if (opts) c->opt = *opts; else pfem_default_options(&c->opt);
        PF_HIP(hipGetDevice(&c->device));
        PF_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        PF_HIP(hipEventCreate(&c->ev0));
        PF_HIP(hipEventCreate(&c->ev1));
        for (auto &ev : c->ev_kry) PF_HIP(hipEventCreate(&ev));
        PF_HIP(hipEventCreateWithFlags(&c->ev_scal, hipEventDisableTiming));
        if (c->opt.assembly_chunk > 0) {
            PF_HIP(hipStreamCreateWithFlags(&c->stream2, hipStreamNonBlocking));
            for (auto &ev : c->ev_pipe) PF_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
        }
        PF_HIP(hipHostMalloc((void **)&c->h_scal, 64 * sizeof(double), hipHostMallocDefault));
        hipStream_t s = c->stream;
        const int64_t nn = md->nnodes, ne = md->nelem;
        c->nnodes = nn; c->nelem = ne; c->ndof = 3 * nn;
        c->kind = md->element_kind;
        c->fbar = md->element_kind == PFEM_HEX8_FINITE_FBAR;
        for (const DevMat &m : dm) c->kin |= m.Hkin > 0;
        c->h_nodes.assign(md->nodes, md->nodes + 3 * nn);
        c->X.alloc(3 * nn); c->X.upload(md->nodes, 3 * nn, s);
        c->conn.alloc(8 * ne); c->conn.upload(md->conn, 8 * ne, s);
        if (md->material_id) { c->matid.alloc(ne); c->matid.upload(md->material_id, ne, s); }
        c->mats.alloc(dm.size()); c->mats.upload(dm.data(), dm.size(), s);

        // ---- pattern (Assembly.jl:62-162) as SELL-64 BSR, scatter slots (on the device
        //      unless host_build is set or a node is too dense for the device build)
        const double tb0 = now();
        const bool host_build = c->opt.host_build != 0;
        FineMatrix &K = c->K;
        DevicePattern P;
        const int brc = host_build ? BUILD_TOO_DENSE : build_pattern_device(nn, ne, c->conn.p, P, s);
        if (brc == BUILD_OK) {
            K.nb = nn;
            K.nslices = P.nslices;
            K.nslots = P.nslots;
            K.slice_ptr = std::move(P.slice_ptr);
            K.rowlen = std::move(P.rowlen);
            K.colind = std::move(P.colind);
            K.diagpos = std::move(P.diagpos);
            c->slots = std::move(P.slots);
            // host copies of the pattern (AMG symbolic phase, exports, Dirichlet rows)
            K.h_slice_ptr.resize(K.nslices + 1);
            K.h_rowlen.resize(nn);
            K.h_colind.resize(K.nslots);
            K.slice_ptr.download(K.h_slice_ptr.data(), K.nslices + 1, s);
            K.rowlen.download(K.h_rowlen.data(), nn, s);
            K.colind.download(K.h_colind.data(), K.nslots, s);
            c->sync();
            K.nnzb = 0;
            for (int64_t i = 0; i < nn; ++i) K.nnzb += K.h_rowlen[i];
            c->built_on_device |= 1;
        } else {
            std::vector<int64_t> aptr;
            std::vector<int32_t> aidx;
            node_adjacency(nn, ne, md->conn, aptr, aidx);
            K.nb = nn;
            K.nnzb = aptr[nn];
            K.nslices = (nn + SELL_C - 1) / SELL_C;
            K.h_slice_ptr.assign(K.nslices + 1, 0);
            K.h_rowlen.resize(nn);
            for (int64_t i = 0; i < nn; ++i) K.h_rowlen[i] = (int32_t)(aptr[i + 1] - aptr[i]);
            for (int64_t sl = 0; sl < K.nslices; ++sl) {
                int32_t w = 0;
                for (int64_t i = sl * SELL_C; i < std::min<int64_t>(nn, (sl + 1) * SELL_C); ++i)
                    w = std::max(w, K.h_rowlen[i]);
                K.h_slice_ptr[sl + 1] = K.h_slice_ptr[sl] + (int64_t)SELL_C * w;
            }
            K.nslots = K.h_slice_ptr[K.nslices];
            K.h_colind.assign(K.nslots, 0);
            std::vector<int32_t> diag(nn, -1);
            for (int64_t sl = 0; sl < K.nslices; ++sl) {
                const int64_t w = (K.h_slice_ptr[sl + 1] - K.h_slice_ptr[sl]) / SELL_C;
                for (int64_t l = 0; l < SELL_C; ++l) {
                    const int64_t i = sl * SELL_C + l;
                    for (int64_t k = 0; k < w; ++k) {
                        const int64_t pos = K.h_slice_ptr[sl] + k * SELL_C + l;
                        if (i < nn && k < K.h_rowlen[i]) {
                            K.h_colind[pos] = aidx[aptr[i] + k];
                            if (aidx[aptr[i] + k] == i) diag[i] = (int32_t)pos;
                        } else {
                            K.h_colind[pos] = (int32_t)std::min<int64_t>(i, nn - 1);  // padding: value 0
                        }
                    }
                }
            }
            if (K.nslots >= (int64_t)1 << 31) throw status_failure(PFEM_ERR_BAD_ARG, "pattern exceeds int32 slots");
            K.slice_ptr.alloc(K.nslices + 1); K.slice_ptr.upload(K.h_slice_ptr.data(), K.nslices + 1, s);
            K.rowlen.alloc(nn); K.rowlen.upload(K.h_rowlen.data(), nn, s);
            K.colind.alloc(K.nslots); K.colind.upload(K.h_colind.data(), K.nslots, s);
            K.diagpos.alloc(nn); K.diagpos.upload(diag.data(), nn, s);

            // ---- scatter slots: block (a, b) of element e -> slot in row conn[a]
            std::vector<int32_t> hs(64 * ne);
            for (int64_t e = 0; e < ne; ++e)
                for (int a = 0; a < 8; ++a) {
                    const int64_t na = md->conn[8 * e + a];
                    const int32_t *lo = aidx.data() + aptr[na], *hi = aidx.data() + aptr[na + 1];
                    const int64_t sb = K.h_slice_ptr[na / SELL_C] + na % SELL_C;
                    for (int b = 0; b < 8; ++b) {
                        const int32_t nbn = (int32_t)md->conn[8 * e + b];
                        const int64_t k = std::lower_bound(lo, hi, nbn) - lo;
                        hs[64 * e + 8 * a + b] = (int32_t)(sb + k * SELL_C);
                    }
                }
            c->slots.alloc(64 * ne); c->slots.upload(hs.data(), 64 * ne, s);
        }
        K.v64.alloc(9 * K.nslots); K.v64.zero(s);

        // ---- colours: parity classes of element_colors (Assembly.jl:172-188) computed on
        //      the device for box numbering, else greedy first-fit in element order (equal
        //      to the parity colours on box meshes)
        int64_t dims[3];
        if (!host_build && box_dims_from_conn(nn, ne, md->conn, c->conn.p, dims, s)) {
            c->color_order.alloc(ne);
            dbuf<int32_t> col(ne);
            c->ncolors = parity_colours_device(dims, col.p, c->color_order.p, c->color_ptr, s);
            c->h_color.resize(ne);
            col.download(c->h_color.data(), ne, s);
            c->sync();
            c->built_on_device |= 2;
        } else {
            std::vector<uint64_t> mask(nn, 0);
            c->h_color.resize(ne);
            int nc = 0;
            for (int64_t e = 0; e < ne; ++e) {
                uint64_t forb = 0;
                for (int a = 0; a < 8; ++a) forb |= mask[md->conn[8 * e + a]];
                int col = 0;
                while (col < 64 && ((forb >> col) & 1ull)) ++col;
                if (col == 64) throw status_failure(PFEM_ERR_BAD_ARG, "more than 64 colours");
                c->h_color[e] = col;
                nc = std::max(nc, col + 1);
                for (int a = 0; a < 8; ++a) mask[md->conn[8 * e + a]] |= 1ull << col;
            }
            c->ncolors = nc;
            c->color_ptr.assign(nc + 1, 0);
            for (int64_t e = 0; e < ne; ++e) c->color_ptr[c->h_color[e] + 1]++;
            for (int k = 0; k < nc; ++k) c->color_ptr[k + 1] += c->color_ptr[k];
            std::vector<int64_t> order(ne), f(c->color_ptr.begin(), c->color_ptr.end() - 1);
            for (int64_t e = 0; e < ne; ++e) order[f[c->h_color[e]]++] = e;
            c->color_order.alloc(ne); c->color_order.upload(order.data(), ne, s);
    }
        c->first_touch.alloc(ne);
        build_first_touch(K.nslots, c->ncolors, c->color_order.p, c->color_ptr.data(), c->slots.p, c->first_touch.p, s);
        c->t_create_pattern = now() - tb0;

        // ---- state (GaussState Mesh.jl:132-138) and scratch
        const int64_t ngp = 8 * ne;
        c->epsp.alloc(6 * ngp); c->abar.alloc(ngp); c->cpinv.alloc(6 * ngp); c->sig.alloc(6 * ngp);
        c->sig_c.alloc(6 * ngp);
        if (c->kin || c->kind == PFEM_HEX8_SMALL) c->beta.alloc(6 * ngp);
        // packets: one per Gauss point of the mesh, or two chunk-sized buffers when the
        // assembly is pipelined (the buffer is also pfem_gauss_stress's 6 x ngp scratch)
        int64_t max_colour = 0;
        for (int col = 0; col < c->ncolors; ++col)
            max_colour = std::max<int64_t>(max_colour, c->color_ptr[col + 1] - c->color_ptr[col]);
        c->packet_elems = c->opt.assembly_chunk > 0 ? std::min<int64_t>(c->opt.assembly_chunk, max_colour) : 0;
        const size_t plen = pk_len(c->kind == PFEM_HEX8_SMALL || !c->kin);
        c->packet.alloc(std::max<size_t>(c->packet_elems ? (size_t)2 * c->packet_elems * 8 * plen : (size_t)ngp * plen,
                                         (size_t)6 * ngp));
        c->geo.alloc((size_t)24 * ngp);
        c->g0.alloc(24 * ne);
        c->fail_key.alloc(1);
        const int64_t nd = c->ndof;
        c->U.alloc(nd); c->R.alloc(nd); c->dU.alloc(nd); c->rhs.alloc(nd); c->Fext0.alloc(nd); c->dg.alloc(nd);
        c->U_c.alloc(nd);
        c->Fext0.zero(s);
        c->bcflag.alloc(nd); c->bcflag.zero(s); c->bcval.alloc(nd); c->bcval.zero(s);
        for (auto *b : {&c->w_r, &c->w_z, &c->w_p, &c->w_q, &c->w_x, &c->w_s, &c->w_t, &c->w_v, &c->w_rh})
            b->alloc(nd);
        c->partial.alloc(reduce_partials_needed(nn) + 16);
        c->scal.alloc(64);
        c->st.nnodes = nn; c->st.nelem = ne; c->st.ndof = nd; c->st.nnzb = K.nnzb; c->st.ncolors = c->ncolors;
        c->sync();
        return PFEM_OK;
    });
    if (rc) {
        pfem_destroy(c);
        return rc;
    }
    rc = pfem_reset(c);
    if (rc) { pfem_destroy(c); return rc; }
    *out = c;
    return PFEM_OK;
}

void pfem_destroy(pfem_ctx *c) {
    if (!c) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->stream2) (void)hipStreamSynchronize(c->stream2);
    c->amg.reset();
    if (c->h_scal) (void)hipHostFree(c->h_scal);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    for (hipEvent_t ev : c->ev_kry)
        if (ev) (void)hipEventDestroy(ev);
    if (c->ev_scal) (void)hipEventDestroy(c->ev_scal);
    for (auto ev : c->ev_pipe)
        if (ev) (void)hipEventDestroy(ev);
    hipStream_t s = c->stream, s2 = c->stream2;
    delete c;  // buffers freed by their destructors
    if (s) (void)hipStreamDestroy(s);
    if (s2) (void)hipStreamDestroy(s2);
}

int pfem_reset(pfem_ctx *c) {
    if (!c) return PFEM_ERR_BAD_ARG;
    return guarded(nullptr, [&]() -> int {
        hipStream_t s = c->stream;
        const int64_t ngp = 8 * c->nelem;
        c->U.zero(s); c->dU.zero(s); c->R.zero(s); c->U_c.zero(s);
        c->epsp.zero(s); c->abar.zero(s); c->sig.zero(s); c->sig_c.zero(s);
        if (c->beta.p) c->beta.zero(s);
        fill_identity6(ngp, c->cpinv.p, s);
        c->sync();
        return PFEM_OK;
    });
}

int pfem_set_dirichlet(pfem_ctx *c, int64_t n, const int64_t *dofs, const double *values, const uint8_t *ramp) {
    if (!c || n < 0 || (n > 0 && (!dofs || !values))) return PFEM_ERR_BAD_ARG;
    for (int64_t i = 0; i < n; ++i)
        if (dofs[i] < 0 || dofs[i] >= c->ndof) return PFEM_ERR_BAD_ARG;
    return guarded(nullptr, [&]() -> int {
        // _freeze_bcs (Solver.jl:192-218): last assignment wins per dof
        std::vector<uint8_t> flag(c->ndof, 0);
        std::vector<double> val(c->ndof, 0.0);
        for (int64_t i = 0; i < n; ++i) {
            flag[dofs[i]] = (ramp && ramp[i]) ? 2 : 1;
            val[dofs[i]] = values[i];
        }
        std::vector<int64_t> list;
        for (int64_t d = 0; d < c->ndof; ++d)
            if (flag[d]) list.push_back(d);
        // block rows touched: rows holding a constrained dof, or a constrained column
        std::vector<uint8_t> nodebc(c->nnodes, 0), touched(c->nnodes, 0);
        for (int64_t d : list) nodebc[d / 3] = 1;
        const FineMatrix &K = c->K;
        for (int64_t i = 0; i < c->nnodes; ++i) {
            if (nodebc[i]) { touched[i] = 1; continue; }
            const int64_t base = K.h_slice_ptr[i / SELL_C] + i % SELL_C;
            for (int k = 0; k < K.h_rowlen[i]; ++k)
                if (nodebc[K.h_colind[base + (int64_t)k * SELL_C]]) { touched[i] = 1; break; }
        }
        std::vector<int32_t> rows;
        for (int64_t i = 0; i < c->nnodes; ++i)
            if (touched[i]) rows.push_back((int32_t)i);
        hipStream_t s = c->stream;
        c->nbc = (int64_t)list.size();
        c->nbcrows = (int64_t)rows.size();
        c->bcflag.upload(flag.data(), c->ndof, s);
        c->bcval.upload(val.data(), c->ndof, s);
        c->bcdofs.alloc(std::max<int64_t>(1, c->nbc));
        c->bcdofs.upload(list.data(), c->nbc, s);
        c->bcrows.alloc(std::max<int64_t>(1, c->nbcrows));
        c->bcrows.upload(rows.data(), c->nbcrows, s);
        c->pc_ready = false;
        c->sync();
        return PFEM_OK;
    });
}

int pfem_set_neumann(pfem_ctx *c, int64_t n, const int64_t *dofs, const double *values) {
    if (!c || n < 0 || (n > 0 && (!dofs || !values))) return PFEM_ERR_BAD_ARG;
    for (int64_t i = 0; i < n; ++i)
        if (dofs[i] < 0 || dofs[i] >= c->ndof) return PFEM_ERR_BAD_ARG;
    return guarded(nullptr, [&]() -> int {
        std::vector<double> f(c->ndof, 0.0);  // assemble_neumann! (BoundaryConditions.jl:44-49)
        for (int64_t i = 0; i < n; ++i) f[dofs[i]] += values[i];
        double s2 = 0.0;
        for (double v : f) s2 += v * v;
        c->fext0_norm = std::sqrt(s2);
        c->Fext0.upload(f.data(), c->ndof, c->stream);
        c->sync();
        return PFEM_OK;
    });
}

int pfem_set_displacement(pfem_ctx *c, const double *U) {
    if (!c || !U) return PFEM_ERR_BAD_ARG;
    return guarded(nullptr, [&]() -> int { c->U.upload(U, c->ndof, c->stream); c->sync(); return PFEM_OK; });
}
int pfem_get_displacement(pfem_ctx *c, double *U) {
    if (!c || !U) return PFEM_ERR_BAD_ARG;
    return guarded(nullptr, [&]() -> int { c->U.download(U, c->ndof, c->stream); c->sync(); return PFEM_OK; });
}
int pfem_get_R(pfem_ctx *c, double *R) {
    if (!c || !R) return PFEM_ERR_BAD_ARG;
    return guarded(nullptr, [&]() -> int { c->R.download(R, c->ndof, c->stream); c->sync(); return PFEM_OK; });
}
int pfem_get_dU(pfem_ctx *c, double *d) {
    if (!c || !d) return PFEM_ERR_BAD_ARG;
    return guarded(nullptr, [&]() -> int { c->dU.download(d, c->ndof, c->stream); c->sync(); return PFEM_OK; });
}
int pfem_set_state(pfem_ctx *c, const double *epsp, const double *beta, const double *abar,
                   const double *cpinv) {
    if (!c) return PFEM_ERR_BAD_ARG;
    return guarded(nullptr, [&]() -> int {
        const int64_t ngp = 8 * c->nelem;
        if (epsp) c->epsp.upload(epsp, 6 * ngp, c->stream);
        if (beta && c->beta.p) c->beta.upload(beta, 6 * ngp, c->stream);
        if (abar) c->abar.upload(abar, ngp, c->stream);
        if (cpinv) c->cpinv.upload(cpinv, 6 * ngp, c->stream);
        c->sync();
        return PFEM_OK;
    });
}
int pfem_get_state(pfem_ctx *c, double *epsp, double *beta, double *abar, double *cpinv, double *sig) {
    if (!c) return PFEM_ERR_BAD_ARG;
    return guarded(nullptr, [&]() -> int {
        const int64_t ngp = 8 * c->nelem;
        if (epsp) c->epsp.download(epsp, 6 * ngp, c->stream);
        if (beta) {
            if (c->beta.p) c->beta.download(beta, 6 * ngp, c->stream);
            else std::memset(beta, 0, sizeof(double) * 6 * ngp);
        }
        if (abar) c->abar.download(abar, ngp, c->stream);
        if (cpinv) c->cpinv.download(cpinv, 6 * ngp, c->stream);
        if (sig) c->sig_c.download(sig, 6 * ngp, c->stream);
        c->sync();
        return PFEM_OK;
    });
}
int pfem_get_trial_stress(pfem_ctx *c, double *sig) {
    if (!c || !sig) return PFEM_ERR_BAD_ARG;
    return guarded(nullptr, [&]() -> int {
        c->sig.download(sig, 6 * 8 * c->nelem, c->stream);
        c->sync();
        return PFEM_OK;
    });
}
int pfem_gauss_stress(pfem_ctx *c, double *sig) {
    if (!c || !sig) return PFEM_ERR_BAD_ARG;
    return guarded(nullptr, [&]() -> int {
        ElementArgs a{};
        a.nelem = c->nelem; a.X = c->X.p; a.conn = c->conn.p; a.U = c->U.p; a.kind = c->kind;
        // the Gauss-point packet buffer (element scratch, rewritten by every assembly)
        // holds the result for the copy out
        launch_cauchy_kernel(a, c->sig_c.p, c->packet.p, c->stream);
        PF_HIP(hipMemcpyAsync(sig, c->packet.p, c->sig.bytes(), hipMemcpyDeviceToHost, c->stream));
        c->sync();
        return PFEM_OK;
    });
}
int pfem_rollback(pfem_ctx *c) {
    if (!c) return PFEM_ERR_BAD_ARG;
    return guarded(nullptr, [&]() -> int {
        PF_HIP(hipMemcpyAsync(c->U.p, c->U_c.p, c->ndof * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
        c->sync();
        return PFEM_OK;
    });
}
int pfem_get_colors(pfem_ctx *c, int32_t *colour, int32_t *ncolors) {
    if (!c) return PFEM_ERR_BAD_ARG;
    if (colour) std::copy(c->h_color.begin(), c->h_color.end(), colour);
    if (ncolors) *ncolors = c->ncolors;
    return PFEM_OK;
}

int pfem_assemble(pfem_ctx *c, const double *U, int commit, double *R_out, pfem_fail *fail) {
    if (!c) return PFEM_ERR_BAD_ARG;
    return guarded(fail, [&]() -> int {
        if (U) c->U.upload(U, c->ndof, c->stream);
        int rc = assemble_impl(c, commit, fail);
        if (rc) return rc;
        if (R_out) { c->R.download(R_out, c->ndof, c->stream); c->sync(); }
        return PFEM_OK;
    });
}

int pfem_subtract_external(pfem_ctx *c, double lambda) {
    if (!c) return PFEM_ERR_BAD_ARG;
    return guarded(nullptr, [&]() -> int {
        sub_external(c->ndof, lambda, c->Fext0.p, c->R.p, c->stream);
        c->sync();
        return PFEM_OK;
    });
}

static int need_K(pfem_ctx *c) {
    if (c->K_valid) return PFEM_OK;
    g_last_error = "K is invalid: the last assembly failed (inverted element)";
    return PFEM_ERR_BAD_ARG;
}

int pfem_impose_dirichlet(pfem_ctx *c, double lambda) {
    if (!c) return PFEM_ERR_BAD_ARG;
    if (int rc = need_K(c)) return rc;
    return guarded(nullptr, [&]() -> int {
        impose_set(c, 2, lambda);  // Solver.jl:354 dir_ramp
        impose_set(c, 1, lambda);  // :355 dir_fix
        c->sync();
        return PFEM_OK;
    });
}

int pfem_impose_dirichlet_set(pfem_ctx *c, int set, double lambda, const double *U, double *R) {
    if (!c || (set != PFEM_BC_RAMPED && set != PFEM_BC_FIXED)) return PFEM_ERR_BAD_ARG;
    if (int rc = need_K(c)) return rc;
    return guarded(nullptr, [&]() -> int {
        if (U) c->U.upload(U, c->ndof, c->stream);
        if (R) c->R.upload(R, c->ndof, c->stream);
        impose_set(c, set == PFEM_BC_RAMPED ? 2 : 1, lambda);
        if (R) c->R.download(R, c->ndof, c->stream);
        c->sync();
        return PFEM_OK;
    });
}

int pfem_residual_norm(pfem_ctx *c, double *out) {
    if (!c || !out) return PFEM_ERR_BAD_ARG;
    return guarded(nullptr, [&]() -> int {
        dot_dev<double, double>(c->ndof, c->R.p, c->R.p, c->partial.p, c->scal.p, c->stream);
        *out = std::sqrt(c->sync_read(c->scal.p));
        return PFEM_OK;
    });
}

int pfem_update_displacement(pfem_ctx *c) {
    if (!c) return PFEM_ERR_BAD_ARG;
    return guarded(nullptr, [&]() -> int {
        axpby<double>(c->ndof, 1.0, c->dU.p, 1.0, c->U.p, c->stream);
        c->sync();
        return PFEM_OK;
    });
}

int pfem_linear_solve(pfem_ctx *c, const double *b, double rtol, int rebuild_pc, int *niter, int *used,
                      double *relres, double *dU) {
    if (!c || !(rtol > 0)) return PFEM_ERR_BAD_ARG;
    if (int rc = need_K(c)) return rc;
    return guarded(nullptr, [&]() -> int {
        if (b)
            c->rhs.upload(b, c->ndof, c->stream);
        else  // b = -R (Solver.jl:378)
            axpby<double>(c->ndof, -1.0, c->R.p, 0.0, c->rhs.p, c->stream);
        const int rc = linear_solve_impl(c, rtol, rebuild_pc ? 1 : 0, niter, used, relres);
        if (dU && (rc == PFEM_OK || rc == PFEM_ERR_NOT_CONVERGED || rc == PFEM_ERR_NON_SPD)) {
            c->dU.download(dU, c->ndof, c->stream);
            c->sync();
        }
        return rc;
    });
}

int pfem_apply_K(pfem_ctx *c, int prec, const double *x, double *y) {
    if (!c || !x || !y || prec < 0 || prec > 2) return PFEM_ERR_BAD_ARG;
    if (int rc = need_K(c)) return rc;
    return guarded(nullptr, [&]() -> int {
        hipStream_t s = c->stream;
        const int64_t n = c->ndof;
        c->w_p.upload(x, n, s);
        if (prec == 0) {
            spmv_fine64(c->K, c->w_p.p, c->w_q.p, s);
        } else if (prec == 1) {
            if (!c->K.v32.p) c->K.v32.alloc(9 * c->K.nslots);
            convert<double, float>(9 * c->K.nslots, c->K.v64.p, c->K.v32.p, s);
            if (!c->f_p.p) { c->f_p.alloc(n); c->f_q.alloc(n); }
            convert<double, float>(n, c->w_p.p, c->f_p.p, s);
            spmv_fine32(c->K, c->f_p.p, c->f_q.p, s);
            convert<float, double>(n, c->f_q.p, c->w_q.p, s);
        } else {
            split_copy(c->K, s);
            spmv_fine_split(c->K, c->w_p.p, c->w_q.p, s);
        }
        c->w_q.download(y, n, s);
        c->sync();
        return PFEM_OK;
    });
}

int pfem_get_K_bsr(pfem_ctx *c, int64_t *nnzb, int32_t *rowptr, int32_t *colind, double *vals) {
    if (!c) return PFEM_ERR_BAD_ARG;
    if (vals)
        if (int rc = need_K(c)) return rc;
    return guarded(nullptr, [&]() -> int {
        const FineMatrix &K = c->K;
        if (nnzb) *nnzb = K.nnzb;
        if (!rowptr && !colind && !vals) return PFEM_OK;
        std::vector<double> hv;
        if (vals) {
            hv.resize(9 * K.nslots);
            K.v64.download(hv.data(), hv.size(), c->stream);
            c->sync();
        }
        int64_t p = 0;
        if (rowptr) rowptr[0] = 0;
        for (int64_t i = 0; i < K.nb; ++i) {
            const int64_t base = K.h_slice_ptr[i / SELL_C] + i % SELL_C;
            for (int k = 0; k < K.h_rowlen[i]; ++k, ++p) {
                const int64_t pos = base + (int64_t)k * SELL_C;
                if (colind) colind[p] = K.h_colind[pos];
                if (vals) std::memcpy(vals + 9 * p, hv.data() + 9 * pos, 9 * sizeof(double));
            }
            if (rowptr) rowptr[i + 1] = (int32_t)p;
        }
        return PFEM_OK;
    });
}

int pfem_get_K_csc(pfem_ctx *c, int64_t *nnz, int64_t *colptr, int64_t *rowval, double *nzval) {
    // Reference CSC (Assembly.jl:99-162): column 3j+cc holds rows 3nb+{0,1,2} of every
    // neighbour nb of node j ascending.  The pattern is symmetric, so the blocks of
    // block column j are the blocks (nb, j) of block rows nb in adj(j).
    if (!c) return PFEM_ERR_BAD_ARG;
    if (nzval)
        if (int rc = need_K(c)) return rc;
    return guarded(nullptr, [&]() -> int {
        const FineMatrix &K = c->K;
        if (nnz) *nnz = 9 * K.nnzb;
        if (!colptr && !rowval && !nzval) return PFEM_OK;
        std::vector<double> hv;
        if (nzval) {
            hv.resize(9 * K.nslots);
            K.v64.download(hv.data(), hv.size(), c->stream);
            c->sync();
        }
        auto slot_of = [&](int64_t i, int64_t j) -> int64_t {  // block (i, j)
            const int64_t base = K.h_slice_ptr[i / SELL_C] + i % SELL_C;
            int lo = 0, hi = K.h_rowlen[i] - 1;
            while (lo <= hi) {
                int mid = (lo + hi) >> 1;
                int32_t cj = K.h_colind[base + (int64_t)mid * SELL_C];
                if (cj == j) return base + (int64_t)mid * SELL_C;
                if (cj < j) lo = mid + 1; else hi = mid - 1;
            }
            return -1;
        };
        int64_t p = 0;
        if (colptr) colptr[0] = 0;
        for (int64_t j = 0; j < K.nb; ++j) {
            const int64_t basej = K.h_slice_ptr[j / SELL_C] + j % SELL_C;
            for (int cc = 0; cc < 3; ++cc) {
                for (int k = 0; k < K.h_rowlen[j]; ++k) {
                    const int64_t nbn = K.h_colind[basej + (int64_t)k * SELL_C];
                    const int64_t pos = nzval ? slot_of(nbn, j) : 0;
                    for (int r = 0; r < 3; ++r, ++p) {
                        if (rowval) rowval[p] = 3 * nbn + r;
                        if (nzval) nzval[p] = hv[9 * pos + 3 * r + cc];
                    }
                }
                if (colptr) colptr[3 * j + cc + 1] = p;
            }
        }
        return PFEM_OK;
    });
}

int pfem_get_stats(pfem_ctx *c, pfem_stats *st) {
    if (!c || !st) return PFEM_ERR_BAD_ARG;
    *st = c->st;
    size_t bytes = c->K.v64.bytes() + c->K.v32.bytes() + c->K.colind.bytes() + c->slots.bytes() +
                   c->packet.bytes() + c->g0.bytes() + 4 * c->epsp.bytes() + c->beta.bytes();
    st->device_bytes = (int64_t)bytes;
    st->t_create_pattern = c->t_create_pattern;
    st->built_on_device = c->built_on_device;
    return PFEM_OK;
}

int pfem_box_mesh(double lx, double ly, double lz, int64_t nx, int64_t ny, int64_t nz, double jitter, uint64_t seed,
                  double *nodes, int64_t *conn) {
    if (!(lx > 0) || !(ly > 0) || !(lz > 0) || nx < 1 || ny < 1 || nz < 1 || !(jitter >= 0) || !nodes || !conn)
        return PFEM_ERR_BAD_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return PFEM_ERR_NO_DEVICE;
    try {
        hipStream_t s;
        PF_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        struct SG { hipStream_t s; ~SG() { (void)hipStreamDestroy(s); } } sg{s};
        const int64_t nn = (nx + 1) * (ny + 1) * (nz + 1), ne = nx * ny * nz;
        dbuf<double> dn(3 * nn);
        dbuf<int64_t> dc(8 * ne);
        box_mesh_device(lx, ly, lz, nx, ny, nz, jitter, seed, dn.p, dc.p, s);
        dn.download(nodes, 3 * nn, s);
        dc.download(conn, 8 * ne, s);
        PF_HIP(hipStreamSynchronize(s));
        return PFEM_OK;
    } catch (const hip_failure &) {
        return PFEM_ERR_HIP;
    }
}

}  // extern "C"
