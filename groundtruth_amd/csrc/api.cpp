// The C ABI of libgtpeec (include/gtpeec.h): argument checks, uploads, and error codes for everything below it.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <string>

#include "stack.hpp"

namespace gt {
void gmres_batch(gt_stack* st, int nr, gt_op* const* ops, const zd* const* b, zd* const* x, const gt_gmres_params& p,
                 gt_gmres_result* res);
void gmres_batch_mixed(gt_stack* st, int nr, gt_op* const* ops, const zd* const* b, zd* const* x,
                       const gt_gmres_params& p, double inner_tol, int max_refine, gt_gmres_result* res);
void gmres_batch_fp32_basis(gt_stack* st, int nr, gt_op* const* ops, const zd* const* b, zd* const* x,
                            const gt_gmres_params& p, gt_gmres_result* res);
int64_t session_plan_bytes(const gt_stack* st, int cap, int restart, int basis, int max_extra);
gt_session* session_create(gt_stack* st, int cap, int restart, int basis, int max_extra);
int64_t session_bytes(const gt_session* ss);
int session_submit(gt_session* ss, gt_op* op, const zd* b, zd* x, const gt_gmres_params& p);
int session_step(gt_session* ss, int max_steps, int* ready);
bool session_poll(gt_session* ss, int slot, gt_gmres_result* res);
void session_destroy(gt_session* ss);
gt_stack* session_stack(gt_session* ss);
}

using namespace gt;

namespace {

thread_local std::string g_err;

template <class F>
int guard(F&& f) {
    try {
        f();
        g_err.clear();
        return GT_OK;
    } catch (const Error& e) {
        g_err = e.what();
        return e.code;
    } catch (const std::bad_alloc&) {
        g_err = "host allocation failed";
        return GT_ERR_OOM;
    } catch (const std::exception& e) {
        g_err = e.what();
        return GT_ERR_ARG;
    }
}

bool pow2(int v) { return v > 0 && (v & (v - 1)) == 0; }
int ilog2(int v) { int l = 0; while ((1 << l) < v) ++l; return l; }

void need(bool ok, const char* msg) {
    if (!ok) fail(GT_ERR_ARG, msg);
}

const DevBuf<zd>& twiddles(gt_stack* st, int P) {
    auto it = st->tw.find(P);
    if (it != st->tw.end()) return it->second;
    std::vector<zd> h(P);
    std::vector<zf> hf(P);
    for (int m = 0; m < P; ++m) {
        const long double a = 2.0L * 3.14159265358979323846264338327950288L * m / P;
        h[m] = zd((double)cosl(a), (double)-sinl(a));
        hf[m] = zf((float)h[m].re, (float)h[m].im);
    }
    DevBuf<zd> d;
    d.upload(h, st->ctx->stream);
    DevBuf<zf> df;
    df.upload(hf, st->ctx->stream);
    st->twf.emplace(P, std::move(df));
    return st->tw.emplace(P, std::move(d)).first->second;
}

void setup_orient(gt_stack* st, int oi, const int32_t* sij, int cnt, int lo, int n0, int n1, const int32_t* P,
                  const double* K, const KernelGen* kg, const gt_stack* donor = nullptr) {
    OrientHost& o = st->o[oi];
    o.present = cnt > 0;
    o.cnt = cnt;
    o.lo = lo;
    o.n0 = n0;
    o.n1 = n1;
    if (!o.present) return;
    need(K != nullptr || kg != nullptr || donor != nullptr, "kernel spectrum missing");
    need(pow2(P[0]) && pow2(P[1]), "FFT sizes must be powers of two");
    need(P[0] >= 2 * n0 - 1 && P[1] >= 2 * n1 - 1, "FFT sizes below 2n - 1 would wrap the convolution");
    need(n0 < 65536 && n1 < 65536, "grid above 65535 cells");
    o.P0 = P[0];
    o.P1 = P[1];
    o.log0 = ilog2(P[0]);
    o.log1 = ilog2(P[1]);
    const int S = st->S;
    if (o.P0 > LpPts<double>::big / 2)
        fail(GT_ERR_UNSUPPORTED, "column FFT size above this library's LDS budget (P0 <= 4096)");
    if ((size_t)S * o.P0 > (size_t)LpPts<double>::big) st->shape.split = true;   // the column pass through HBM
    if (o.P1 > LpPts<double>::big / 2)
        fail(GT_ERR_UNSUPPORTED, "row FFT size above this library's LDS budget (P1 <= 4096)");
    o.TR = std::max(1, std::min(8, LpPts<double>::rows / (o.P1 + 1)));
    // rows of the dense (sheet, i) grid as CSR over the sorted strips
    std::vector<int> rowptr((size_t)S * n0 + 1, 0), rowcol(cnt);
    int prev = -1;
    for (int k = 0; k < cnt; ++k) {
        const int s = sij[3 * k], i = sij[3 * k + 1], j = sij[3 * k + 2];
        need(s >= 0 && s < S && i >= 0 && i < n0 && j >= 0 && j < n1, "strip outside its grid");
        const int key = (s * n0 + i) * n1 + j;
        need(key > prev, "strips must be sorted by (sheet, i, j) without repeats");
        prev = key;
        rowptr[s * n0 + i + 1]++;
        rowcol[k] = (i << 16) | j;
        st->h_strip_sheet.push_back((unsigned char)s);
    }
    for (size_t r = 0; r + 1 < rowptr.size(); ++r) rowptr[r + 1] += rowptr[r];
    o.rowptr.upload(rowptr, st->ctx->stream);
    o.rowcol.upload(rowcol, st->ctx->stream);
    twiddles(st, o.P0);
    twiddles(st, o.P1);
    if (st->shape.split) twiddles(st, o.P0 / 4);         // the residue-class column pass
    if (donor) {
        const OrientHost& od = donor->o[oi];
        need(od.present && od.P0 == o.P0 && od.P1 == o.P1 && od.n0 == n0 && od.n1 == n1,
             "the donor stack's kernel spectra are of another shape");
        o.K = od.K;
        o.Kf = od.Kf;
        o.k_shared = true;
    } else if (kg) {
        kernel_quadrant_device(st, oi, *kg);
    } else {
        const size_t kq = (size_t)(o.P1 / 2 + 1) * (o.P0 / 2 + 1) * S * S;      // G = S * S: the identity map
        std::vector<double> kd(K, K + kq);
        std::vector<float> kf(kq);
        for (size_t e = 0; e < kq; ++e) kf[e] = (float)kd[e];
        o.K->upload(kd, st->ctx->stream);
        o.Kf->upload(kf, st->ctx->stream);
    }
    o.t_off = st->t_elems;                      // each orientation its own region: they run in one launch
    st->t_elems += (size_t)S * o.P1 * o.n0;
    st->shape.lds_rows = std::max(st->shape.lds_rows, (size_t)o.TR * (o.P1 + 1) * sizeof(zd));
    st->shape.lds_cols = std::max(st->shape.lds_cols, (size_t)S * o.P0 * sizeof(zd));
}

// B = Ar^T Ar from the branches' kept ends: its off-diagonal edges and its diagonal (degrees, edges to reference
// nodes included)
void laplacian(int nk, int nb, const int* a_keep, const int* b_keep, std::vector<int>& ea, std::vector<int>& eb,
               std::vector<double>& diag) {
    diag.assign(nk, 0.0);
    for (int b = 0; b < nb; ++b) {
        const int ka = a_keep[b], kb = b_keep[b];
        need(ka >= -1 && ka < nk && kb >= -1 && kb < nk, "kept node index out of range");
        if (ka >= 0) diag[ka] += 1.0;
        if (kb >= 0) diag[kb] += 1.0;
        if (ka >= 0 && kb >= 0) {
            need(ka != kb, "a branch joins a node to itself");
            ea.push_back(ka);
            eb.push_back(kb);
        }
    }
}

}  // namespace

struct gt_schur_host_s {
    SchurHost h;
};

LpArgs<double> gt_stack_s::lp_args(int mode) const {
    LpArgs<double> a{};
    for (int oi = 0; oi < 2; ++oi) {
        const OrientHost& h = o[oi];
        OrientArgs<double>& d = a.o[oi];
        d.n0 = h.n0;
        d.n1 = h.n1;
        d.P0 = h.P0;
        d.P1 = h.P1;
        d.log0 = h.log0;
        d.log1 = h.log1;
        d.TR = h.TR;
        d.lo = h.lo;
        d.cnt = h.cnt;
        d.rowptr = h.rowptr.p;
        d.rowcol = h.rowcol.p;
        d.K = h.K->p;
        d.T1 = work.T1.p + h.t_off;
        d.T2 = work.T2.p + h.t_off;
        d.tw0 = h.present ? tw.at(h.P0).p : nullptr;
        d.tw1 = h.present ? tw.at(h.P1).p : nullptr;
        d.twq = h.present && shape.split ? tw.at(h.P0 / 4).p : nullptr;
        a.nblk_rows[oi] = h.present ? S * ((h.n0 + h.TR - 1) / h.TR) : 0;
        a.nblk_cols[oi] = h.present ? h.P1 : 0;
    }
    a.S = S;
    a.mode = mode;
    a.t_stride = t_elems;
    a.G = kG;
    a.gmap = kgmap_dev.p;
    a.toeplitz = kG == S && !kgmap.empty();
    for (int t = 0; t < S; ++t)
        for (int u = 0; u < S; ++u) {
            const int g = kgmap.empty() ? 0 : kgmap[(size_t)t * S + u];
            a.gpad[t * LP_MAX_SHEETS + u] = g;
            if (g != (t > u ? t - u : u - t)) a.toeplitz = 0;
        }
    a.a_keep = a_keep.p;
    a.b_keep = b_keep.p;
    return a;
}

// The stack's own device memory plus its core's (a derived stack counts the shared core too: it is what the
// stack needs resident)
int64_t gt_stack_s::device_bytes() const {
    int64_t b = a_keep.bytes() + b_keep.bytes() + adj_ptr.bytes() + adj.bytes();
    b += work.bytes() + W.bytes() + G.bytes() + Yr.bytes();
    if (derived) return b;                 // the core is its base's
    for (const auto& h : o) b += h.rowptr.bytes() + h.rowcol.bytes() + (h.k_shared ? 0 : h.K->bytes() + h.Kf->bytes());
    for (const auto& t : tw) b += t.second.bytes();
    for (const auto& t : twf) b += t.second.bytes();
    b += node_sheet.bytes() + strip_sheet.bytes();
    b += core->nw_fac.bytes() + core->nw_dl.bytes() + core->nw_val.bytes() + core->nw_ptr.bytes() +
         core->nw_col.bytes();
    if (schur) b += schur->device_bytes();
    return b;
}

namespace {

// The incidence of a stack's branches on the device: a_keep, b_keep and the node adjacency (CSR of
// (branch << 1 | end)).
void upload_incidence(gt_stack* st) {
    const int nb = st->nb, nk = st->nk;
    hipStream_t s = st->ctx->stream;
    std::vector<int> aptr(nk + 1, 0);
    for (int b = 0; b < nb; ++b) {
        const int ka = st->h_a_keep[b], kb = st->h_b_keep[b];
        need(ka >= -1 && ka < nk && kb >= -1 && kb < nk, "kept node index out of range");
        if (ka >= 0) aptr[ka + 1]++;
        if (kb >= 0) aptr[kb + 1]++;
    }
    for (int v = 0; v < nk; ++v) aptr[v + 1] += aptr[v];
    std::vector<int> adj(std::max(1, aptr[nk]));
    std::vector<int> fill(aptr.begin(), aptr.end() - 1);
    for (int b = 0; b < nb; ++b) {
        if (st->h_a_keep[b] >= 0) adj[fill[st->h_a_keep[b]]++] = b << 1;
        if (st->h_b_keep[b] >= 0) adj[fill[st->h_b_keep[b]]++] = (b << 1) | 1;
    }
    st->a_keep.upload(st->h_a_keep.empty() ? std::vector<int>{0} : st->h_a_keep, s);
    st->b_keep.upload(st->h_b_keep.empty() ? std::vector<int>{0} : st->h_b_keep, s);
    st->adj_ptr.upload(aptr, s);
    st->adj.upload(adj, s);
}


}  // namespace

extern "C" {

const char* gt_last_error(void) { return g_err.c_str(); }

int gt_abi_version(void) { return GT_ABI_VERSION; }

int gt_ctx_create(int device, gt_ctx** out) {
    return guard([&] {
        need(out != nullptr, "out is NULL");
        *out = nullptr;
        int count = 0;
        GT_HIP(hipGetDeviceCount(&count));
        if (device < 0 || device >= count) fail(GT_ERR_DEVICE, "no HIP device " + std::to_string(device));
        GT_HIP(hipSetDevice(device));
        noalloc_scratch_setup(device);
        hipDeviceProp_t prop;
        GT_HIP(hipGetDeviceProperties(&prop, device));
        auto* c = new gt_ctx_s;
        c->device = device;
        c->lds_max = prop.sharedMemPerBlock;
        if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
            delete c;
            fail(GT_ERR_DEVICE, "hipStreamCreate failed");
        }
        *out = c;
    });
}

void gt_ctx_destroy(gt_ctx* ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

int gt_ctx_trim(gt_ctx* ctx) {
    return guard([&] {
        need(ctx, "NULL context");
        std::lock_guard<std::mutex> lk(ctx->mu);
        GT_HIP(hipSetDevice(ctx->device));
        GT_HIP(hipStreamSynchronize(ctx->stream));
        ctx_trim_gmres(ctx);
    });
}

int gt_ctx_sync(gt_ctx* ctx) {
    return guard([&] {
        need(ctx != nullptr, "ctx is NULL");
        std::lock_guard<std::mutex> lk(ctx->mu);
        GT_HIP(hipStreamSynchronize(ctx->stream));
    });
}

void* gt_ctx_stream(gt_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

int gt_ctx_schur_device_fronts(gt_ctx* ctx, int32_t m_min) {
    return guard([&] {
        need(ctx != nullptr, "ctx is NULL");
        need(m_min >= 1, "the smallest device front must be at least one row");
        std::lock_guard<std::mutex> lk(ctx->mu);
        ctx->schur_dev_min = m_min;
    });
}

int gt_ctx_schur_leaf(gt_ctx* ctx, int32_t nodes) {
    return guard([&] {
        need(ctx != nullptr, "ctx is NULL");
        need(nodes >= 4 && nodes <= 512, "a leaf of 4 to 512 nodes");
        std::lock_guard<std::mutex> lk(ctx->mu);
        ctx->schur_leaf = nodes;
    });
}

int gt_ctx_schur_ordering(gt_ctx* ctx, int32_t ordering) {
    return guard([&] {
        need(ctx != nullptr, "ctx is NULL");
        need(ordering == GT_ND_SHEETS_APART || ordering == GT_ND_JOINT, "unknown ordering");
        std::lock_guard<std::mutex> lk(ctx->mu);
        ctx->schur_sheets_apart = ordering == GT_ND_SHEETS_APART;
    });
}

int gt_ctx_woodbury(gt_ctx* ctx, int32_t form, int64_t stored_max_bytes) {
    return guard([&] {
        need(ctx != nullptr, "ctx is NULL");
        need(form == GT_WOODBURY_AUTO || form == GT_WOODBURY_STORED || form == GT_WOODBURY_IMPLICIT,
             "unknown Woodbury form");
        need(stored_max_bytes >= 0, "negative byte limit");
        std::lock_guard<std::mutex> lk(ctx->mu);
        ctx->woodbury = form;
        ctx->wb_stored_max = stored_max_bytes;
    });
}

int gt_stack_woodbury(const gt_stack* st, int32_t* form, int32_t* k) {
    return guard([&] {
        need(st && form && k, "NULL argument");
        need(st->schur != nullptr, "the stack has no Schur factor (gt_schur_factor first)");
        *form = st->wb_root ? GT_WOODBURY_ROOT : st->wb_implicit ? GT_WOODBURY_IMPLICIT : GT_WOODBURY_STORED;
        *k = st->k;
    });
}

int gt_ctx_graphs(gt_ctx* ctx, int enable) {
    return guard([&] {
        need(ctx != nullptr, "ctx is NULL");
        std::lock_guard<std::mutex> lk(ctx->mu);
        ctx->graphs = enable != 0;
    });
}

int gt_ctx_orthogonalization(gt_ctx* ctx, int mode) {
    return guard([&] {
        need(ctx != nullptr, "ctx is NULL");
        need(mode == GT_GS_CGS_DGKS || mode == GT_GS_DELAYED_CGS2, "mode is GT_GS_CGS_DGKS or GT_GS_DELAYED_CGS2");
        std::lock_guard<std::mutex> lk(ctx->mu);
        ctx->gs_delayed = mode == GT_GS_DELAYED_CGS2;
    });
}

int gt_ctx_lp_timing(gt_ctx* ctx, int enable) {
    return guard([&] {
        need(ctx != nullptr, "ctx is NULL");
        std::lock_guard<std::mutex> lk(ctx->mu);
        ctx->lp_timing = enable != 0;
        ctx->lp_used = 0;
        ctx->lp_rhs = 0;
    });
}

int gt_ctx_lp_time(gt_ctx* ctx, int64_t* launches, int64_t* rhs, double* ms) {
    return guard([&] {
        need(ctx && launches && rhs && ms, "NULL argument");
        std::lock_guard<std::mutex> lk(ctx->mu);
        GT_HIP(hipStreamSynchronize(ctx->stream));
        double total = 0;
        for (size_t i = 0; i < ctx->lp_used; ++i) {
            float t = 0;
            GT_HIP(hipEventElapsedTime(&t, ctx->lp_events[2 * i], ctx->lp_events[2 * i + 1]));
            total += t;
        }
        *launches = (int64_t)ctx->lp_used;
        *rhs = ctx->lp_rhs;
        *ms = total;
    });
}

}  // extern "C"

namespace {

void create_stack(gt_ctx* ctx, const gt_stack_desc* d, const gt_kernel_desc* kd, gt_stack** out,
                  const gt_stack* donor = nullptr) {
        need(ctx && d && out, "NULL argument");
        std::lock_guard<std::mutex> lk(ctx->mu);
        *out = nullptr;
        need(d->n_sheets >= 1 && d->n_sheets <= LP_MAX_SHEETS, "1 to 8 current sheets");
        need(d->nx >= 2 && d->ny >= 2, "grid below 2 x 2");
        need(d->nxb >= 0 && d->nyb >= 0 && d->nbr >= 0 && d->n_keep >= 0, "negative count");
        GT_HIP(hipSetDevice(ctx->device));
        std::unique_ptr<gt_stack_s> st(new gt_stack_s);
        st->ctx = ctx;
        st->S = d->n_sheets;
        st->nx = d->nx;
        st->ny = d->ny;
        st->nxb = d->nxb;
        st->nyb = d->nyb;
        st->nbr = d->nbr;
        st->ns = d->nxb + d->nyb;
        st->nb = st->ns + d->nbr;
        st->nk = d->n_keep;
        hipStream_t s = ctx->stream;
        KernelGen kg;
        const int S = st->S;
        st->kgmap.assign((size_t)S * S, 0);
        if (donor) {
            need(donor->ctx == ctx && donor->S == S && donor->nx == d->nx && donor->ny == d->ny,
                 "the donor stack is of another context, sheet count or grid");
            st->kgmap = donor->kgmap;
            st->kG = donor->kG;
        } else if (kd) {
            need(kd->d > 0 && kd->dz && kd->near_field, "kernel description incomplete");
            kg.d = kd->d;
            kg.mu = kd->mu;
            kg.dz = kd->dz;
            kg.nearv = kd->near_field;
            // one kernel per distinct face gap: sheet pairs with the same gap and near-field table share it
            constexpr int NF = 25;
            for (int a = 0; a < S; ++a)
                for (int b = a; b < S; ++b) {
                    const int e = a * S + b;
                    int g = 0;
                    for (; g < kg.G; ++g) {
                        const int r = kg.rep_s[g] * S + kg.rep_t[g];
                        if (kd->dz[r] == kd->dz[e] &&
                            std::memcmp(kd->near_field + (size_t)r * NF, kd->near_field + (size_t)e * NF,
                                        NF * sizeof(double)) == 0)
                            break;
                    }
                    if (g == kg.G) {
                        kg.rep_s[g] = a;
                        kg.rep_t[g] = b;
                        ++kg.G;
                    }
                    st->kgmap[(size_t)a * S + b] = st->kgmap[(size_t)b * S + a] = g;
                }
            st->kG = kg.G;
        } else {
            for (int e = 0; e < S * S; ++e) st->kgmap[e] = e;
            st->kG = S * S;
        }
        st->kgmap_dev.upload(st->kgmap, s);
        setup_orient(st.get(), 0, d->x_sij, d->nxb, 0, d->nx - 1, d->ny, d->px, d->kx, kd ? &kg : nullptr, donor);
        setup_orient(st.get(), 1, d->y_sij, d->nyb, d->nxb, d->nx, d->ny - 1, d->py, d->ky, kd ? &kg : nullptr, donor);
        st->strip_sheet.upload(st->h_strip_sheet.empty() ? std::vector<unsigned char>{0} : st->h_strip_sheet, s);
        const int nb = st->nb, nk = st->nk;
        need(nb == 0 || (d->a_keep && d->b_keep), "incidence missing");
        st->h_a_keep.assign(d->a_keep, d->a_keep + nb);
        st->h_b_keep.assign(d->b_keep, d->b_keep + nb);
        upload_incidence(st.get());
        std::vector<int> ns_(std::max(1, nk), 0);
        if (nk) {
            need(d->node_sheet != nullptr, "node_sheet missing");
            for (int v = 0; v < nk; ++v) {
                need(d->node_sheet[v] >= 0 && d->node_sheet[v] < st->S, "node sheet out of range");
                ns_[v] = d->node_sheet[v];
            }
        }
        st->node_sheet.upload(ns_, s);
        st->work.ensure(st.get(), 1);
        GT_HIP(hipStreamSynchronize(s));
        *out = st.release();
}

}  // namespace

extern "C" {

int gt_stack_create(gt_ctx* ctx, const gt_stack_desc* d, gt_stack** out) {
    return guard([&] { create_stack(ctx, d, nullptr, out); });
}

int gt_stack_create_gen(gt_ctx* ctx, const gt_stack_desc* d, const gt_kernel_desc* kd, gt_stack** out) {
    return guard([&] {
        need(kd != nullptr, "NULL argument");
        create_stack(ctx, d, kd, out);
    });
}

int gt_stack_create_like(gt_ctx* ctx, const gt_stack_desc* d, const gt_stack* donor, gt_stack** out) {
    return guard([&] {
        need(donor != nullptr, "NULL donor");
        need(d && d->px[0] == donor->o[0].P0 && d->px[1] == donor->o[0].P1 && d->py[0] == donor->o[1].P0 &&
                 d->py[1] == donor->o[1].P1,
             "the donor's FFT sizes differ");
        create_stack(ctx, d, nullptr, out, donor);
    });
}

int gt_node_field(gt_stack* st, const void* x_dev, double d, void* kx_dev, void* ky_dev) {
    return guard([&] {
        need(st && x_dev && kx_dev && ky_dev, "NULL argument");
        std::lock_guard<std::mutex> lk(st->ctx->mu);
        GT_HIP(hipSetDevice(st->ctx->device));
        node_field_device(st, (const zd*)x_dev, d, (zd*)kx_dev, (zd*)ky_dev);
        GT_HIP(hipStreamSynchronize(st->ctx->stream));
    });
}

int gt_loop_terms(gt_stack* st, const void* x_dev, const void* psi_dev, const void* bm_dev, int32_t n_inj,
                  const int32_t* inj_idx, const double* inj_amps, double d, const double* sheet_z, double* out) {
    return guard([&] {
        need(st && x_dev && psi_dev && sheet_z && out && n_inj >= 0 && (n_inj == 0 || (inj_idx && inj_amps)),
             "NULL argument");
        std::lock_guard<std::mutex> lk(st->ctx->mu);
        GT_HIP(hipSetDevice(st->ctx->device));
        loop_terms_device(st, (const zd*)x_dev, (const double*)psi_dev, (const double*)bm_dev, n_inj, inj_idx, inj_amps,
                          d, sheet_z, out);
    });
}

int gt_via_coupling(gt_ctx* ctx, int32_t nv, const double* vias, int32_t nseg, const gt_segment* segs, double mu,
                    double* out) {
    return guard([&] {
        need(ctx && nv >= 0 && nseg >= 0 && (nv == 0 || (vias && out)) && (nseg == 0 || segs), "NULL argument");
        std::lock_guard<std::mutex> lk(ctx->mu);
        GT_HIP(hipSetDevice(ctx->device));
        via_coupling_device(ctx, nv, vias, nseg, segs, mu, out);
    });
}

int gt_raster_label(gt_ctx* ctx, int32_t n_pieces, const int64_t* edge_ptr, const double* edges,
                    const int32_t* windows, int32_t nu, const double* xs, int32_t nv, const double* ys,
                    int32_t* labels) {
    return guard([&] {
        need(ctx, "NULL context");
        std::lock_guard<std::mutex> lk(ctx->mu);
        GT_HIP(hipSetDevice(ctx->device));
        raster_label(ctx, n_pieces, edge_ptr, edges, windows, nu, xs, nv, ys, labels);
    });
}

int gt_raster_fine(gt_ctx* ctx, int32_t n_pieces, const int64_t* edge_ptr, const double* edges,
                   const int32_t* windows, int32_t nx, int32_t ny, int32_t sub, const double* xf, const double* yf,
                   int32_t* part, double* fill, double* cross_x, double* off_x, double* cross_y, double* off_y,
                   int32_t* fine_labels) {
    return guard([&] {
        need(ctx, "NULL context");
        std::lock_guard<std::mutex> lk(ctx->mu);
        GT_HIP(hipSetDevice(ctx->device));
        raster_fine(ctx, n_pieces, edge_ptr, edges, windows, nx, ny, sub, xf, yf, part, fill, cross_x, off_x, cross_y,
                    off_y, fine_labels);
    });
}

int gt_raster_state(gt_ctx* ctx, int64_t n_edges, const double* edges, const int32_t* window, int32_t nu,
                    const double* xs, int32_t nv, const double* ys, int8_t* state) {
    return guard([&] {
        need(ctx, "NULL context");
        std::lock_guard<std::mutex> lk(ctx->mu);
        GT_HIP(hipSetDevice(ctx->device));
        raster_state(ctx, n_edges, edges, window, nu, xs, nv, ys, state);
    });
}

int gt_raster_cover(gt_ctx* ctx, int64_t n_edges, const double* edges, double u0, double du, int32_t n_lines,
                    double v0, double d, int32_t n_cells, double* out) {
    return guard([&] {
        need(ctx, "NULL context");
        std::lock_guard<std::mutex> lk(ctx->mu);
        GT_HIP(hipSetDevice(ctx->device));
        raster_cover(ctx, n_edges, edges, u0, du, n_lines, v0, d, n_cells, out);
    });
}

int gt_raster_points(gt_ctx* ctx, int64_t n_edges, const double* edges, int64_t n_points, const double* points,
                     int8_t* inside, int8_t* on_edge, double* nearest) {
    return guard([&] {
        need(ctx, "NULL context");
        std::lock_guard<std::mutex> lk(ctx->mu);
        GT_HIP(hipSetDevice(ctx->device));
        raster_points(ctx, n_edges, edges, n_points, points, inside, on_edge, nearest);
    });
}

int gt_strip_coupling(gt_stack* st, double d, double mu, int32_t nseg, const gt_segment* segs, const double* sheet_z,
                      const double* sheet_t, void* psi_dev) {
    return guard([&] {
        need(st && sheet_z && sheet_t && psi_dev && nseg >= 0 && (nseg == 0 || segs), "NULL argument");
        std::lock_guard<std::mutex> lk(st->ctx->mu);
        GT_HIP(hipSetDevice(st->ctx->device));
        strip_coupling_device(st, d, mu, nseg, segs, sheet_z, sheet_t, (double*)psi_dev);
    });
}

int gt_stack_kernel(const gt_stack* st, int32_t orient, double* out) {
    return guard([&] {
        need(st && out && (orient == 0 || orient == 1), "bad argument");
        std::lock_guard<std::mutex> lk(st->ctx->mu);
        const OrientHost& o = st->o[orient];
        if (!o.present) return;
        // the resident G kernels per pixel, expanded to the S x S sheet pairs
        std::vector<double> k(o.K->n);
        GT_HIP(hipMemcpyAsync(k.data(), o.K->p, o.K->bytes(), hipMemcpyDeviceToHost, st->ctx->stream));
        GT_HIP(hipStreamSynchronize(st->ctx->stream));
        const int S = st->S, G = st->kG;
        const size_t npix = o.K->n / G;
        for (size_t q = 0; q < npix; ++q)
            for (int e = 0; e < S * S; ++e) out[q * S * S + e] = k[q * G + st->kgmap[e]];
    });
}

void gt_stack_destroy(gt_stack* st) {
    if (!st) return;
    std::lock_guard<std::mutex> lk(st->ctx->mu);
    (void)hipSetDevice(st->ctx->device);
    (void)hipStreamSynchronize(st->ctx->stream);
    delete st;
}

int64_t gt_stack_device_bytes(const gt_stack* st) { return st ? st->device_bytes() : 0; }

int gt_lp_matvec(gt_stack* st, const void* I_dev, void* out_dev) {
    return guard([&] {
        need(st && I_dev && out_dev, "NULL argument");
        std::lock_guard<std::mutex> lk(st->ctx->mu);
        if (st->ns == 0) return;
        st->work.ensure(st, 1);                  // a what-if stack makes its scratch on its first own product
        LpArgs<double> a = st->lp_args(0);
        a.rhs.nr = 1;
        a.rhs.x[0] = (const zd*)I_dev;
        a.rhs.y[0] = (zd*)out_dev;
        launch_lp(a, st->shape, st->ctx->stream);
        narrow_apply<double>(st, 1, nullptr, a.rhs.x, a.rhs.y);
    });
}

int gt_lp_matvec_batch(gt_stack* st, int32_t nr, const void* const* I_dev, void* const* out_dev) {
    return guard([&] {
        need(st && I_dev && out_dev, "NULL argument");
        need(nr >= 1 && nr <= KMAX, "1 to 16 right-hand sides per batch");
        for (int i = 0; i < nr; ++i) need(I_dev[i] && out_dev[i] && I_dev[i] != out_dev[i], "NULL or aliased vector");
        std::lock_guard<std::mutex> lk(st->ctx->mu);
        if (st->ns == 0) return;
        st->work.ensure(st, nr);
        LpArgs<double> a = st->lp_args(0);
        a.rhs.nr = nr;
        for (int i = 0; i < nr; ++i) {
            a.rhs.x[i] = (const zd*)I_dev[i];
            a.rhs.y[i] = (zd*)out_dev[i];
        }
        launch_lp(a, st->shape, st->ctx->stream);
        narrow_apply<double>(st, nr, nullptr, a.rhs.x, a.rhs.y);
    });
}

int gt_lp_matvec_batch_noalloc(gt_stack* st, int32_t nr, const void* const* I_dev, void* const* out_dev) {
    return guard([&] {
        need(st && I_dev && out_dev, "NULL argument");
        need(nr >= 1 && nr <= KMAX, "1 to 16 right-hand sides per batch");
        for (int i = 0; i < nr; ++i) need(I_dev[i] && out_dev[i] && I_dev[i] != out_dev[i], "NULL or aliased vector");
        std::lock_guard<std::mutex> lk(st->ctx->mu);
        if (st->ns == 0) return;
        const size_t need_bytes = 2 * sizeof(zd) * (size_t)nr * std::max<size_t>(1, st->t_elems);
        if (need_bytes > GT_LP_NOALLOC_SCRATCH_BYTES)
            fail(GT_ERR_ARG, "gt_lp_matvec_batch_noalloc: " + std::to_string(nr) + " right-hand sides need " +
                                 std::to_string(need_bytes) + " bytes of scratch, above its " +
                                 std::to_string((size_t)GT_LP_NOALLOC_SCRATCH_BYTES) + "; use fewer per call");
        zd* T1 = static_cast<zd*>(noalloc_scratch(st->ctx->device));
        zd* T2 = T1 + (size_t)nr * std::max<size_t>(1, st->t_elems);
        GT_HIP(hipSetDevice(st->ctx->device));
        LpArgs<double> a = st->lp_args(0);
        for (int oi = 0; oi < 2; ++oi) {
            a.o[oi].T1 = T1 + st->o[oi].t_off;
            a.o[oi].T2 = T2 + st->o[oi].t_off;
        }
        a.rhs.nr = nr;
        for (int i = 0; i < nr; ++i) {
            a.rhs.x[i] = (const zd*)I_dev[i];
            a.rhs.y[i] = (zd*)out_dev[i];
        }
        launch_lp(a, st->shape, st->ctx->stream);
        narrow_apply<double>(st, nr, nullptr, a.rhs.x, a.rhs.y);
    });
}

int gt_stack_set_narrow(gt_stack* st, const double* r_factor, const double* dL, int64_t nnz, const int32_t* rowptr,
                        const int32_t* col, const double* val) {
    return guard([&] {
        need(st, "NULL stack");
        need(nnz >= 0 && (nnz == 0 || (rowptr && col && val)), "bad corrections");
        std::lock_guard<std::mutex> lk(st->ctx->mu);
        GT_HIP(hipSetDevice(st->ctx->device));
        const int ns = st->ns;
        if (ns == 0) return;
        std::vector<double> fac(ns, 0.0), dl(ns, 0.0);
        for (int s = 0; s < ns; ++s) {
            if (r_factor) fac[s] = r_factor[s] - 1.0;
            if (dL) dl[s] = dL[s];
        }
        std::vector<int> ptr(ns + 1, 0), cc(std::max<int64_t>(1, nnz), 0);
        std::vector<double> vv(std::max<int64_t>(1, nnz), 0.0);
        if (nnz) {
            need(rowptr[0] == 0 && rowptr[ns] == nnz, "CSR row pointers do not span the entries");
            for (int s = 0; s < ns; ++s) need(rowptr[s + 1] >= rowptr[s], "CSR row pointers decrease");
            for (int64_t k = 0; k < nnz; ++k) need(col[k] >= 0 && col[k] < ns, "CSR column off the strips");
            std::copy(rowptr, rowptr + ns + 1, ptr.begin());
            std::copy(col, col + nnz, cc.begin());
            std::copy(val, val + nnz, vv.begin());
        }
        gt::StackCore& c = *st->core;
        hipStream_t s = st->ctx->stream;
        c.nw_fac.upload(fac, s);
        c.nw_dl.upload(dl, s);
        c.nw_ptr.upload(ptr, s);
        c.nw_col.upload(cc, s);
        c.nw_val.upload(vv, s);
        c.narrow = true;
    });
}

int gt_stack_lp_autotune(gt_stack* st, int32_t reps, int32_t* rw_out, int32_t* cw_out) {
    return guard([&] {
        need(st && rw_out && cw_out && reps >= 1, "bad argument");
        std::lock_guard<std::mutex> lk(st->ctx->mu);
        GT_HIP(hipSetDevice(st->ctx->device));
        hipStream_t s = st->ctx->stream;
        const int nrs[3] = {1, 4, 8};
        // the shapes choose among the one-wave-per-transform passes, which exist for transforms of 256 to 1024
        // points: with every transform longer, every shape launches the same kernels and there is nothing to time
        bool shaped = false;
        {
            const LpArgs<double> a0 = st->lp_args(0);
            for (int oi = 0; oi < 2; ++oi)
                if (a0.nblk_rows[oi] > 0 && (a0.o[oi].P0 <= 1024 || a0.o[oi].P1 <= 1024)) shaped = true;
        }
        if (st->ns > 0 && shaped) {
            const int cap0 = st->work.cap;          // the products' workspace goes back to this afterwards
            st->work.ensure(st, 8);
            DevBuf<zd> X((size_t)st->ns * 8), Y((size_t)st->ns * 8);
            GT_HIP(hipMemsetAsync(X.p, 0, X.bytes(), s));
            LpArgs<double> a = st->lp_args(0);
            for (int i = 0; i < 8; ++i) {
                a.rhs.x[i] = X.p + (size_t)i * st->ns;
                a.rhs.y[i] = Y.p + (size_t)i * st->ns;
            }
            hipEvent_t e0, e1;
            GT_HIP(hipEventCreate(&e0));
            GT_HIP(hipEventCreate(&e1));
            const int shapes[3] = {1, 2, 4};
            for (int b = 0; b < 3; ++b) {
                a.rhs.nr = nrs[b];
                float best = 0;
                int brw = st->shape.rw[b], bcw = st->shape.cw[b];
                for (int rw : shapes)
                    for (int cw : shapes) {
                        LpShape sh = st->shape;
                        sh.rw[b] = rw;
                        sh.cw[b] = cw;
                        launch_lp(a, sh, s);
                        launch_lp(a, sh, s);
                        GT_HIP(hipEventRecord(e0, s));
                        for (int r = 0; r < reps; ++r) launch_lp(a, sh, s);
                        GT_HIP(hipEventRecord(e1, s));
                        GT_HIP(hipEventSynchronize(e1));
                        float ms = 0;
                        GT_HIP(hipEventElapsedTime(&ms, e0, e1));
                        if (best == 0 || ms < best) {
                            best = ms;
                            brw = rw;
                            bcw = cw;
                        }
                    }
                st->shape.rw[b] = brw;
                st->shape.cw[b] = bcw;
            }
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            GT_HIP(hipStreamSynchronize(s));
            st->work.shrink(st, cap0);
        }
        for (int b = 0; b < 3; ++b) {
            rw_out[b] = st->shape.rw[b];
            cw_out[b] = st->shape.cw[b];
        }
    });
}

int gt_stack_set_lp_plan(gt_stack* st, const int32_t* rw, const int32_t* cw) {
    return guard([&] {
        need(st && rw && cw, "NULL argument");
        for (int b = 0; b < 3; ++b) {
            need(rw[b] == 1 || rw[b] == 2 || rw[b] == 4, "rows per workgroup: 1, 2 or 4");
            need(cw[b] == 1 || cw[b] == 2 || cw[b] == 4, "waves per workgroup: 1, 2 or 4");
        }
        std::lock_guard<std::mutex> lk(st->ctx->mu);
        for (int b = 0; b < 3; ++b) {
            st->shape.rw[b] = rw[b];
            st->shape.cw[b] = cw[b];
        }
    });
}

namespace {

// The real Laplacian B = Ar^T Ar of the stack (strips and branches at weight 1), factorised on the device
void factor_laplacian(gt_stack* st, const int32_t* node_xyz) {
    const int nk = st->nk, nb = st->nb;
    std::vector<int> ea, eb;
    std::vector<double> diag;
    laplacian(nk, nb, st->h_a_keep.data(), st->h_b_keep.data(), ea, eb, diag);
    auto S = std::make_unique<Schur>();
    S->h.dev_min = st->ctx->schur_dev_min;
    S->h.leaf = st->ctx->schur_leaf;
    S->h.sheets_apart = st->ctx->schur_sheets_apart;
    S->h.build(nk, ea, eb, diag, node_xyz);
    S->upload(st->ctx->stream);
    S->factor_device(st->ctx->stream);
    S->top_device(st->ctx->stream);
    st->schur = std::move(S);
}

}  // namespace

int gt_schur_factor_nodes(gt_stack* st, const int32_t* node_xyz, int32_t* n_cols, int32_t* cols, double* G_out) {
    return guard([&] {
        need(st && node_xyz && n_cols, "NULL argument");
        need(!st->derived, "a derived stack shares its base's factor");
        std::lock_guard<std::mutex> lk(st->ctx->mu);
        factor_laplacian(st, node_xyz);
        // the distinct kept end nodes of the branches, ascending: one unit column each (wb_units with kb = -1)
        std::vector<int> nodes;
        for (int m = 0; m < st->nbr; ++m)
            for (int e : {st->h_a_keep[st->ns + m], st->h_b_keep[st->ns + m]})
                if (e >= 0) nodes.push_back(e);
        std::sort(nodes.begin(), nodes.end());
        nodes.erase(std::unique(nodes.begin(), nodes.end()), nodes.end());
        const int k = (int)nodes.size();
        if (k > WB_NODES_MAX_K) fail(GT_ERR_UNSUPPORTED, "more than 2048 branch end nodes on one stack");
        need(k == 0 || cols != nullptr, "NULL cols");
        st->wb_implicit = true;
        st->wb_nodes = true;
        std::vector<int> none(k, -1);
        woodbury_extend(st, nullptr, k, nodes.data(), none.data());
        if (k > 0 && !st->wb_root)
            fail(GT_ERR_UNSUPPORTED, "sheets of different admittance need every branch end in the factor's last "
                                     "supernode (branches that join nodes of one sheet are not supported)");
        *n_cols = k;
        for (int c = 0; c < k; ++c) cols[c] = nodes[c];
        if (G_out) woodbury_gram_device(st, G_out);
        GT_HIP(hipStreamSynchronize(st->ctx->stream));
    });
}

int gt_schur_factor(gt_stack* st, const int32_t* node_xyz, double* G_out) {
    return guard([&] {
        need(st && node_xyz, "NULL argument");
        need(!st->derived, "a derived stack shares its base's factor");
        std::lock_guard<std::mutex> lk(st->ctx->mu);
        const int nk = st->nk;
        factor_laplacian(st, node_xyz);
        // Woodbury columns W = B^-1 U (U: the branches' incidence) by device solves, and G = U^T W; W stays resident
        // unless it is too large (then only G is kept and each product solves for B^-1 U t)
        const int k = st->nbr;
        const int form = st->ctx->woodbury;
        st->wb_implicit = form == GT_WOODBURY_IMPLICIT ||
                          (form == GT_WOODBURY_AUTO && (k > WB_STORED_MAX_K ||
                                                        (int64_t)k * nk * (int64_t)sizeof(double) > st->ctx->wb_stored_max));
        need(k <= (st->wb_implicit ? WB_IMPLICIT_MAX_K : WB_STORED_MAX_K),
             st->wb_implicit ? "more than 8192 branches on one stack" : "more than 256 branches with W resident");
        std::vector<int> ka(k), kb(k);
        for (int m = 0; m < k; ++m) {
            ka[m] = st->h_a_keep[st->ns + m];
            kb[m] = st->h_b_keep[st->ns + m];
        }
        woodbury_extend(st, nullptr, k, ka.data(), kb.data());
        if (G_out) woodbury_gram_device(st, G_out);
        GT_HIP(hipStreamSynchronize(st->ctx->stream));
    });
}

int gt_stack_derive(gt_stack* base, int32_t n_extra, const int32_t* a_keep_extra, const int32_t* b_keep_extra,
                    double* G_out, gt_stack** out) {
    return guard([&] {
        need(base && out && n_extra >= 0 && (n_extra == 0 || (a_keep_extra && b_keep_extra)), "bad argument");
        std::lock_guard<std::mutex> lk(base->ctx->mu);
        need(base->schur != nullptr, "the base stack has no Schur factor (gt_schur_factor first)");
        if (base->wb_nodes)
            fail(GT_ERR_UNSUPPORTED, "a stack of sheets of different admittance derives no what-ifs: make a stack");
        *out = nullptr;
        GT_HIP(hipSetDevice(base->ctx->device));
        std::unique_ptr<gt_stack_s> st(new gt_stack_s(base->core));
        st->derived = true;
        st->lineage.push_back(base->id);
        st->lineage.insert(st->lineage.end(), base->lineage.begin(), base->lineage.end());
        st->ctx = base->ctx;
        st->S = base->S;
        st->nx = base->nx;
        st->ny = base->ny;
        st->nxb = base->nxb;
        st->nyb = base->nyb;
        st->ns = base->ns;
        st->nk = base->nk;
        st->nbr = base->nbr + n_extra;
        st->nb = st->ns + st->nbr;
        const int nk = st->nk;
        st->h_a_keep = base->h_a_keep;
        st->h_b_keep = base->h_b_keep;
        for (int e = 0; e < n_extra; ++e) {
            need(a_keep_extra[e] >= -1 && a_keep_extra[e] < nk && b_keep_extra[e] >= -1 && b_keep_extra[e] < nk,
                 "kept node index out of range");
            need(a_keep_extra[e] != b_keep_extra[e], "a branch joins a node to itself");
            st->h_a_keep.push_back(a_keep_extra[e]);
            st->h_b_keep.push_back(b_keep_extra[e]);
        }
        upload_incidence(st.get());
        // W' = [W  B^-1 U']: the base's columns copied on the device, the added ones by device solves (or, in the
        // implicit form, G' from the base's G and the added columns' solves)
        st->wb_implicit = base->wb_implicit;
        hipStream_t s = st->ctx->stream;
        woodbury_extend(st.get(), base, n_extra, a_keep_extra, b_keep_extra);
        if (G_out) woodbury_gram_device(st.get(), G_out);
        // no scratch of its own: its solves run in its base's batches (check_batch), and its own products make it
        // when first called
        GT_HIP(hipStreamSynchronize(s));
        *out = st.release();
    });
}

int gt_stack_derive_bytes(const gt_stack* base, int32_t n_extra, const int32_t* a_keep_extra,
                          const int32_t* b_keep_extra, int64_t* held, int64_t* peak) {
    return guard([&] {
        need(base && held && peak && n_extra >= 0 && (n_extra == 0 || (a_keep_extra && b_keep_extra)),
             "bad argument");
        std::lock_guard<std::mutex> lk(base->ctx->mu);
        need(base->schur != nullptr, "the base stack has no Schur factor (gt_schur_factor first)");
        const int64_t nb = base->nb + n_extra, nk = base->nk, k = base->k + n_extra;
        int64_t ends = 0;
        for (int b = 0; b < base->nb; ++b) ends += (base->h_a_keep[b] >= 0) + (base->h_b_keep[b] >= 0);
        for (int e = 0; e < n_extra; ++e) ends += (a_keep_extra[e] >= 0) + (b_keep_extra[e] >= 0);
        const int64_t i4 = sizeof(int), d8 = sizeof(double);
        int64_t h = 2 * i4 * std::max<int64_t>(1, nb) + i4 * (nk + 1) + i4 * std::max<int64_t>(1, ends);
        h += base->wb_implicit ? d8 * std::max<int64_t>(1, k * k) : d8 * std::max<int64_t>(1, k * nk);
        int64_t t = 0;                           // set-up only: the added columns' solves, 16 at a time
        if (n_extra > 0 && nk > 0) {
            const int chunk = std::min<int>(n_extra, KMAX);
            t = 2 * (int64_t)sizeof(zd) * nk * chunk;
            t += std::max<int64_t>(0, base->schur->bytes_after(chunk) - base->schur->bytes_after(0));
        }
        *held = h;
        *peak = h + t;
    });
}

int gt_matvec_on(gt_stack* host, gt_op* op, const void* x_dev, void* y_dev) {
    return guard([&] {
        need(host && op && x_dev && y_dev && x_dev != y_dev, "NULL or aliased argument");
        std::lock_guard<std::mutex> lk(host->ctx->mu);
        const zd* x = (const zd*)x_dev;
        zd* y = (zd*)y_dev;
        ops_matvec(host, 1, &op, &x, &y);
    });
}

int gt_schur_stats(const gt_stack* st, int64_t* nnz_L, int32_t* n_super, int32_t* n_levels, int32_t* max_cols) {
    return guard([&] {
        need(st && st->schur, "no Schur factor");
        const SchurHost& h = st->schur->h;
        if (nnz_L) *nnz_L = h.nnz_L;
        if (n_super) *n_super = (int32_t)h.c0.size();
        if (n_levels) *n_levels = h.levels;
        if (max_cols) *max_cols = h.max_cols;
    });
}

int gt_schur_top_stats(const gt_stack* st, int32_t* levels, int32_t* cols, int64_t* entries) {
    return guard([&] {
        need(st && st->schur, "no Schur factor");
        const SchurHost& h = st->schur->h;
        if (levels) *levels = h.top_T;
        if (cols) *cols = (int32_t)h.top_col.size();
        if (entries) *entries = h.top_fnnz + h.top_bnnz;
    });
}

int gt_schur_solve_host(const gt_stack* st, const double* b, double* x) {
    return guard([&] {
        need(st && st->schur && b && x, "NULL argument or no factor");
        std::lock_guard<std::mutex> lk(st->ctx->mu);
        st->schur->fetch_panels(st->ctx->stream);     // the device fronts' panels, once
        st->schur->h.solve_host(b, x);
    });
}

int gt_schur_solve(gt_stack* st, const void* b_dev, void* x_dev) {
    return guard([&] {
        need(st && b_dev && x_dev && b_dev != x_dev, "NULL or aliased argument");
        std::lock_guard<std::mutex> lk(st->ctx->mu);
        schur_apply(st, (const zd*)b_dev, (zd*)x_dev, 1);
    });
}

int gt_schur_host_factor(int32_t n_keep, int32_t nb, const int32_t* a_keep, const int32_t* b_keep,
                         const int32_t* node_xyz, gt_schur_host** out) {
    return guard([&] {
        need(out && (nb == 0 || (a_keep && b_keep)) && (n_keep == 0 || node_xyz), "NULL argument");
        *out = nullptr;
        std::vector<int> ea, eb;
        std::vector<double> diag;
        laplacian(n_keep, nb, a_keep, b_keep, ea, eb, diag);
        auto h = std::make_unique<gt_schur_host_s>();
        h->h.build(n_keep, ea, eb, diag, node_xyz);
        *out = h.release();
    });
}

int gt_schur_host_solve(const gt_schur_host* h, const double* b, double* x, int32_t k) {
    return guard([&] {
        need(h && b && x && k >= 0, "bad argument");
        const size_t n = (size_t)h->h.n;
        for (int32_t c = 0; c < k; ++c) h->h.solve_host(b + 2 * n * c, x + 2 * n * c);
    });
}

int gt_schur_host_stats(const gt_schur_host* hh, int64_t* nnz_L, int32_t* n_super, int32_t* n_levels,
                        int32_t* max_cols) {
    return guard([&] {
        need(hh != nullptr, "NULL handle");
        const SchurHost& h = hh->h;
        if (nnz_L) *nnz_L = h.nnz_L;
        if (n_super) *n_super = (int32_t)h.c0.size();
        if (n_levels) *n_levels = h.levels;
        if (max_cols) *max_cols = h.max_cols;
    });
}

void gt_schur_host_free(gt_schur_host* h) { delete h; }

int gt_op_create(gt_stack* st, const gt_op_desc* d, gt_op** out) {
    return guard([&] {
        need(st && d && out, "NULL argument");
        std::lock_guard<std::mutex> lk(st->ctx->mu);
        *out = nullptr;
        need(d->r_sheet && d->dinv_sheet, "per-sheet impedances missing");
        need(st->nbr == 0 || (d->z_tail && d->dinv_tail), "branch impedances missing");
        need(d->k == 0 || (d->k == st->k && d->q != nullptr), "Woodbury rank must be the stack's branch count");
        need(d->k <= (st->wb_implicit ? WB_IMPLICIT_MAX_K : WB_STORED_MAX_K), "Woodbury rank above the stack's limit");
        std::unique_ptr<gt_op_s> op(new gt_op_s);
        op->st = st;
        op->w = d->w;
        hipStream_t s = st->ctx->stream;
        auto up = [&](DevBuf<zd>& dst, const double* src, size_t cnt) {
            std::vector<zd> v(std::max<size_t>(1, cnt));
            for (size_t i = 0; i < cnt; ++i) v[i] = zd(src[2 * i], src[2 * i + 1]);
            dst.upload(v, s);
        };
        up(op->r_sheet, d->r_sheet, st->S);
        up(op->dinv_sheet, d->dinv_sheet, st->S);
        up(op->z_tail, d->z_tail, st->nbr);
        up(op->dinv_tail, d->dinv_tail, st->nbr);
        up(op->q, d->q, (size_t)d->k * d->k);
        auto upf = [&](DevBuf<zf>& dst, const double* src, size_t cnt) {
            std::vector<zf> v(std::max<size_t>(1, cnt));
            for (size_t i = 0; i < cnt; ++i) v[i] = zf((float)src[2 * i], (float)src[2 * i + 1]);
            dst.upload(v, s);
        };
        upf(op->r_sheet_f, d->r_sheet, st->S);
        upf(op->z_tail_f, d->z_tail, st->nbr);
        op->s_scale = zd(d->s_scale[0], d->s_scale[1]);
        need((d->s_in == nullptr) == (d->s_out == nullptr), "s_in and s_out go together");
        need(!st->wb_nodes || d->s_in != nullptr, "a node-column stack's operator needs per-sheet scales");
        if (d->s_in) {
            up(op->s_in, d->s_in, st->S);
            up(op->s_out, d->s_out, st->S);
            op->per_sheet = true;
        }
        op->k = d->k;
        GT_HIP(hipStreamSynchronize(s));
        *out = op.release();
    });
}

void gt_op_destroy(gt_op* op) {
    if (!op) return;
    std::lock_guard<std::mutex> lk(op->st->ctx->mu);
    delete op;
}

int gt_matvec(gt_op* op, const void* x_dev, void* y_dev) {
    return guard([&] {
        need(op && x_dev && y_dev && x_dev != y_dev, "NULL or aliased argument");
        std::lock_guard<std::mutex> lk(op->st->ctx->mu);
        const zd* x = (const zd*)x_dev;
        zd* y = (zd*)y_dev;
        ops_matvec(op->st, 1, &op, &x, &y);
    });
}

int gt_precond(gt_op* op, const void* r_dev, void* z_dev) {
    return guard([&] {
        need(op && r_dev && z_dev && r_dev != z_dev, "NULL or aliased argument");
        std::lock_guard<std::mutex> lk(op->st->ctx->mu);
        need(op->st->schur != nullptr, "the stack has no Schur factor (gt_schur_factor first)");
        const zd* r = (const zd*)r_dev;
        zd* z = (zd*)z_dev;
        ops_precond(op->st, 1, &op, &r, &z);
    });
}

int gt_gmres(gt_op* op, const double* b_host, double* x_host, const gt_gmres_params* p, gt_gmres_result* res) {
    return guard([&] {
        need(op && b_host && x_host && p && res, "NULL argument");
        std::lock_guard<std::mutex> lk(op->st->ctx->mu);
        need(op->st->schur != nullptr, "the stack has no Schur factor (gt_schur_factor first)");
        need(p->tol > 0 && p->max_cycles >= 0 && p->extra_cycles >= 0, "bad GMRES parameters");
        GT_HIP(hipSetDevice(op->st->ctx->device));
        const int n = op->st->nb + op->st->nk;
        hipStream_t s = op->st->ctx->stream;
        DevBuf<zd> b(n), x(n);
        GT_HIP(hipMemcpyAsync(b.p, b_host, sizeof(zd) * n, hipMemcpyHostToDevice, s));
        const zd* bp = b.p;
        zd* xp = x.p;
        gmres_batch(op->st, 1, &op, &bp, &xp, *p, res);
        GT_HIP(hipMemcpyAsync(x_host, x.p, sizeof(zd) * n, hipMemcpyDeviceToHost, s));
        GT_HIP(hipStreamSynchronize(s));
    });
}

int gt_gmres_dev(gt_op* op, const void* b_dev, void* x_dev, const gt_gmres_params* p, gt_gmres_result* res) {
    return guard([&] {
        need(op && b_dev && x_dev && p && res && b_dev != x_dev, "NULL or aliased argument");
        std::lock_guard<std::mutex> lk(op->st->ctx->mu);
        need(op->st->schur != nullptr, "the stack has no Schur factor (gt_schur_factor first)");
        need(p->tol > 0 && p->max_cycles >= 0 && p->extra_cycles >= 0, "bad GMRES parameters");
        GT_HIP(hipSetDevice(op->st->ctx->device));
        const zd* bp = (const zd*)b_dev;
        zd* xp = (zd*)x_dev;
        gmres_batch(op->st, 1, &op, &bp, &xp, *p, res);
    });
}

int gt_gmres_batch(gt_stack* st, int32_t nr, gt_op* const* ops, const void* const* b_dev, void* const* x_dev,
                   const gt_gmres_params* p, gt_gmres_result* res) {
    return guard([&] {
        need(st && ops && b_dev && x_dev && p && res, "NULL argument");
        std::lock_guard<std::mutex> lk(st->ctx->mu);
        need(nr >= 1 && nr <= KMAX, "1 to 16 right-hand sides per batch");
        need(st->schur != nullptr, "the stack has no Schur factor (gt_schur_factor first)");
        need(p->tol > 0 && p->max_cycles >= 0 && p->extra_cycles >= 0, "bad GMRES parameters");
        std::vector<const zd*> b(nr);
        std::vector<zd*> x(nr);
        for (int i = 0; i < nr; ++i) {
            need(ops[i] && ops[i]->st == st, "every operator must belong to the stack");
            need(b_dev[i] && x_dev[i] && b_dev[i] != x_dev[i], "NULL or aliased vector");
            b[i] = (const zd*)b_dev[i];
            x[i] = (zd*)x_dev[i];
        }
        GT_HIP(hipSetDevice(st->ctx->device));
        gmres_batch(st, nr, ops, b.data(), x.data(), *p, res);
    });
}

int gt_gmres_batch_fp32_basis(gt_stack* st, int32_t nr, gt_op* const* ops, const void* const* b_dev,
                              void* const* x_dev, const gt_gmres_params* p, gt_gmres_result* res) {
    return guard([&] {
        need(st && ops && b_dev && x_dev && p && res, "NULL argument");
        std::lock_guard<std::mutex> lk(st->ctx->mu);
        need(nr >= 1 && nr <= KMAX, "1 to 16 right-hand sides per batch");
        need(st->schur != nullptr, "the stack has no Schur factor (gt_schur_factor first)");
        need(p->tol > 0 && p->max_cycles >= 0 && p->extra_cycles >= 0, "bad GMRES parameters");
        std::vector<const zd*> b(nr);
        std::vector<zd*> x(nr);
        for (int i = 0; i < nr; ++i) {
            need(ops[i] && ops[i]->st == st, "every operator must belong to the stack");
            need(b_dev[i] && x_dev[i] && b_dev[i] != x_dev[i], "NULL or aliased vector");
            b[i] = (const zd*)b_dev[i];
            x[i] = (zd*)x_dev[i];
        }
        GT_HIP(hipSetDevice(st->ctx->device));
        gmres_batch_fp32_basis(st, nr, ops, b.data(), x.data(), *p, res);
    });
}

int gt_gmres_batch_mixed(gt_stack* st, int32_t nr, gt_op* const* ops, const void* const* b_dev, void* const* x_dev,
                         const gt_gmres_params* p, double inner_tol, int32_t max_refine, gt_gmres_result* res) {
    return guard([&] {
        need(st && ops && b_dev && x_dev && p && res, "NULL argument");
        std::lock_guard<std::mutex> lk(st->ctx->mu);
        need(nr >= 1 && nr <= KMAX, "1 to 16 right-hand sides per batch");
        need(st->schur != nullptr, "the stack has no Schur factor (gt_schur_factor first)");
        need(p->tol > 0 && p->max_cycles >= 0 && p->extra_cycles >= 0, "bad GMRES parameters");
        need(max_refine >= 1 && max_refine <= 64, "1 to 64 refinement rounds");
        std::vector<const zd*> b(nr);
        std::vector<zd*> x(nr);
        for (int i = 0; i < nr; ++i) {
            need(ops[i] && ops[i]->st == st, "every operator must belong to the stack");
            need(b_dev[i] && x_dev[i] && b_dev[i] != x_dev[i], "NULL or aliased vector");
            b[i] = (const zd*)b_dev[i];
            x[i] = (zd*)x_dev[i];
        }
        GT_HIP(hipSetDevice(st->ctx->device));
        gmres_batch_mixed(st, nr, ops, b.data(), x.data(), *p, inner_tol, max_refine, res);
    });
}

int gt_session_open(gt_stack* st, int32_t capacity, int32_t restart, int32_t basis, int32_t max_extra,
                    gt_session** out) {
    return guard([&] {
        need(st && out, "NULL argument");
        std::lock_guard<std::mutex> lk(st->ctx->mu);
        *out = nullptr;
        need(st->schur != nullptr, "the stack has no Schur factor (gt_schur_factor first)");
        GT_HIP(hipSetDevice(st->ctx->device));
        *out = session_create(st, capacity, restart, basis, max_extra);
    });
}

int gt_session_create(gt_stack* st, int32_t capacity, int32_t restart, gt_session** out) {
    return gt_session_open(st, capacity, restart, GT_BASIS_FP64, 0, out);
}

void gt_session_destroy(gt_session* ss) {
    if (!ss) return;
    gt_ctx* ctx = session_stack(ss)->ctx;
    std::lock_guard<std::mutex> lk(ctx->mu);
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    session_destroy(ss);
}

int gt_session_submit(gt_session* ss, gt_op* op, const void* b_dev, void* x_dev, const gt_gmres_params* p,
                      int32_t* slot) {
    return guard([&] {
        need(ss && op && b_dev && x_dev && p && slot && b_dev != x_dev, "NULL or aliased argument");
        std::lock_guard<std::mutex> lk(session_stack(ss)->ctx->mu);
        need(p->tol > 0 && p->max_cycles >= 0 && p->extra_cycles >= 0, "bad GMRES parameters");
        *slot = session_submit(ss, op, (const zd*)b_dev, (zd*)x_dev, *p);
    });
}

int gt_session_step(gt_session* ss, int32_t max_steps, int32_t* running, int32_t* ready) {
    return guard([&] {
        need(ss && running && ready, "NULL argument");
        need(max_steps >= 0, "negative step count");
        gt_ctx* ctx = session_stack(ss)->ctx;
        std::lock_guard<std::mutex> lk(ctx->mu);
        GT_HIP(hipSetDevice(ctx->device));
        int rd = 0;
        *running = session_step(ss, max_steps, &rd);
        *ready = rd;
    });
}

int gt_session_poll(gt_session* ss, int32_t slot, gt_gmres_result* res, int32_t* done) {
    return guard([&] {
        need(ss && res && done, "NULL argument");
        std::lock_guard<std::mutex> lk(session_stack(ss)->ctx->mu);
        *done = session_poll(ss, slot, res) ? 1 : 0;
    });
}

int64_t gt_session_device_bytes(const gt_session* ss) { return ss ? session_bytes(ss) : 0; }

int64_t gt_session_open_bytes(const gt_stack* st, int32_t capacity, int32_t restart, int32_t basis,
                              int32_t max_extra) {
    if (!st || capacity < 1 || capacity > KMAX || restart < 0 || max_extra < 0) return -1;
    if (basis != GT_BASIS_FP64 && basis != GT_BASIS_FP32) return -1;
    try {
        return session_plan_bytes(st, capacity, restart, basis, max_extra);
    } catch (...) {
        return -1;
    }
}

int64_t gt_session_plan_bytes(const gt_stack* st, int32_t capacity, int32_t restart) {
    return gt_session_open_bytes(st, capacity, restart, GT_BASIS_FP64, 0);
}

}  // extern "C"
