// The operator's device work for a batch of right-hand sides, each with its own frequency (gt_op) on one stack:
// the system product K x (peec.py:3043-3062) and the preconditioner M^-1 r (peec.py:3064-3083):
// y = S^-1 (r2 - Ar^T D^-1 r1), x1 = D^-1 (r1 + Ar y), with S^-1 = (1/d_s) (B + U C U^T)^-1 applied as B^-1 (the
// device triangular solves, all right-hand sides in one pass over the panels) plus the Woodbury correction
// z - W Q (U^T z) (W = B^-1 U resident, Q each frequency's k x k matrix). Loops over the batch sit inside the
// kernels, so the incidence and W are read once per batch.
#include "stack.hpp"

namespace gt {

namespace {

struct PreArgs {
    int ns, nb, nk, k, nr;
    const zd* r[KMAX];        // n: [r1 ; r2]
    zd* out[KMAX];            // n: [x1 ; y]
    const zd* dinv_sheet[KMAX];
    const zd* dinv_tail[KMAX];
    const zd* q[KMAX];        // k x k row-major
    zd s_scale[KMAX];
    zd* s;                    // nk x nr interleaved: Schur right-hand sides
    const zd* z;              // nk x nr interleaved: B^-1 s
    zd* t;                    // nr x k: Q (U^T z)
    const double* W;          // k columns of nk
    const int* a_keep;
    const int* b_keep;
    const int* adj_ptr;
    const int* adj;
    const int* node_sheet;
    const unsigned char* strip_sheet;
};

// s = r2 - Ar^T (D^-1 r1): a kept node's incident strips lie on its own sheet (one admittance), its branches carry
// their own
__global__ __launch_bounds__(256) void pre_schur_rhs(PreArgs a) {
    const int n = blockIdx.x * 256 + threadIdx.x;
    if (n >= a.nk) return;
    const int e0 = a.adj_ptr[n], e1 = a.adj_ptr[n + 1];
    const int sh = a.node_sheet[n];
    for (int p = 0; p < a.nr; ++p) {
        const zd* r = a.r[p];
        zd acc = r[a.nb + n];
        const zd ds = a.dinv_sheet[p][sh];
        for (int e = e0; e < e1; ++e) {
            const int code = a.adj[e], b = code >> 1;
            const zd v = (b < a.ns ? ds : a.dinv_tail[p][b - a.ns]) * r[b];
            if (code & 1) acc = acc + v;
            else acc = acc - v;
        }
        a.s[(size_t)n * a.nr + p] = acc;
    }
}

// t_p = Q_p (U^T z_p): one workgroup per right-hand side (k, the branch count, is small)
__global__ __launch_bounds__(256) void pre_woodbury_small(PreArgs a) {
    __shared__ double ut_raw[512];
    zd* ut = reinterpret_cast<zd*>(ut_raw);
    const int p = blockIdx.x;
    for (int m = threadIdx.x; m < a.k; m += 256) {
        const int b = a.ns + m;
        const int ka = a.a_keep[b], kb = a.b_keep[b];
        zd v;
        if (ka >= 0) v = v + a.z[(size_t)ka * a.nr + p];
        if (kb >= 0) v = v - a.z[(size_t)kb * a.nr + p];
        ut[m] = v;
    }
    __syncthreads();
    const zd* q = a.q[p];
    for (int i = threadIdx.x; i < a.k; i += 256) {
        zd acc;
        for (int m = 0; m < a.k; ++m) acc = acc + q[i * a.k + m] * ut[m];
        a.t[p * a.k + i] = acc;
    }
}

// y_p = s_scale_p (z_p - W t_p), written to the node part of each output. A thread takes one node for the whole
// batch: each W value is read once and applied to every right-hand side (t of the batch staged in LDS, in chunks
// of WB branches). The sums over m run in the same order for every batch size.
constexpr int WB = 64;

__global__ __launch_bounds__(256) void pre_woodbury_apply(PreArgs a) {
    __shared__ double ts_raw[2 * WB * KMAX];
    zd* ts = reinterpret_cast<zd*>(ts_raw);
    const int n = blockIdx.x * 256 + threadIdx.x;
    double vre[KMAX], vim[KMAX];
#pragma unroll
    for (int p = 0; p < KMAX; ++p) {
        const zd z = n < a.nk && p < a.nr ? a.z[(size_t)n * a.nr + p] : zd();
        vre[p] = z.re;
        vim[p] = z.im;
    }
    for (int m0 = 0; m0 < a.k; m0 += WB) {
        const int mc = min(WB, a.k - m0);
        __syncthreads();
        for (int e = threadIdx.x; e < mc * KMAX; e += 256) {      // right-hand sides past the batch: zeros
            const int p = e / mc, m = e - p * mc;
            ts[m * KMAX + p] = p < a.nr ? a.t[p * a.k + m0 + m] : zd();
        }
        __syncthreads();
        if (n < a.nk) {
            for (int m = 0; m < mc; ++m) {
                const double w = a.W[(size_t)(m0 + m) * a.nk + n];
                const zd* tm = ts + m * KMAX;
#pragma unroll
                for (int p = 0; p < KMAX; ++p) {
                    vre[p] -= w * tm[p].re;
                    vim[p] -= w * tm[p].im;
                }
            }
        }
    }
    if (n >= a.nk) return;
#pragma unroll
    for (int p = 0; p < KMAX; ++p)
        if (p < a.nr) a.out[p][a.nb + n] = a.s_scale[p] * zd(vre[p], vim[p]);
}

// x1 = D^-1 (r1 + Ar y)
__global__ __launch_bounds__(256) void pre_branches(PreArgs a) {
    const int b = blockIdx.x * 256 + threadIdx.x;
    if (b >= a.nb) return;
    const int ka = a.a_keep[b], kb = a.b_keep[b];
    const int sh = b < a.ns ? a.strip_sheet[b] : -1;
    for (int p = 0; p < a.nr; ++p) {
        zd v = a.r[p][b];
        const zd* y = a.out[p] + a.nb;
        if (ka >= 0) v = v + y[ka];
        if (kb >= 0) v = v - y[kb];
        const zd d = sh >= 0 ? a.dinv_sheet[p][sh] : a.dinv_tail[p][b - a.ns];
        a.out[p][b] = d * v;
    }
}

}  // namespace

void Work::ensure(const gt_stack* st, int nr) {
    if (nr <= cap) return;
    T1.alloc(std::max<size_t>(1, st->t_elems) * nr);
    T2.alloc(std::max<size_t>(1, st->t_elems) * nr);
    s.alloc((size_t)std::max(1, st->nk) * nr);
    z.alloc((size_t)std::max(1, st->nk) * nr);
    t.alloc((size_t)std::max(1, st->nbr) * nr);
    cap = nr;
}

void ops_matvec(gt_stack* st, int nr, gt_op* const* ops, const zd* const* x, zd* const* y) {
    if (nr < 1 || nr > KMAX) fail(GT_ERR_ARG, "1 to 16 right-hand sides per batch");
    hipStream_t s = st->ctx->stream;
    st->work.ensure(st, nr);
    RhsSet<double> rs{};
    rs.nr = nr;
    for (int p = 0; p < nr; ++p) {
        rs.x[p] = x[p];
        rs.y[p] = y[p];
        rs.w[p] = ops[p]->w;
        rs.r_sheet[p] = ops[p]->r_sheet.p;
        rs.z_tail[p] = ops[p]->z_tail.p;
    }
    LpArgs<double> la = st->lp_args(1);
    la.rhs = rs;
    if (st->ns > 0) {
        gt_ctx* c = st->ctx;
        if (c->lp_timing) {
            if (c->lp_used * 2 + 2 > c->lp_events.size()) {
                for (int e = 0; e < 2; ++e) {
                    hipEvent_t ev;
                    GT_HIP(hipEventCreate(&ev));
                    c->lp_events.push_back(ev);
                }
            }
            GT_HIP(hipEventRecord(c->lp_events[2 * c->lp_used], s));
            launch_lp(la, st->shape, s);
            GT_HIP(hipEventRecord(c->lp_events[2 * c->lp_used + 1], s));
            c->lp_used++;
            c->lp_rhs += nr;
        } else {
            launch_lp(la, st->shape, s);
        }
    }
    TailArgs<double> ta{st->ns, st->nb, st->nbr, st->nk, st->a_keep.p, st->b_keep.p, st->adj_ptr.p, st->adj.p, rs};
    launch_tail_bottom(ta, s);
}

void schur_apply(gt_stack* st, const zd* b, zd* x, int nr) {
    if (!st->schur) fail(GT_ERR_ARG, "the stack has no Schur factor (gt_schur_factor first)");
    schur_solve_device(*st->schur, b, x, nr, st->ctx->stream);
}

void ops_precond(gt_stack* st, int nr, gt_op* const* ops, const zd* const* r, zd* const* z) {
    if (nr < 1 || nr > KMAX) fail(GT_ERR_ARG, "1 to 16 right-hand sides per batch");
    hipStream_t s = st->ctx->stream;
    st->work.ensure(st, nr);
    PreArgs a{};
    a.ns = st->ns;
    a.nb = st->nb;
    a.nk = st->nk;
    a.nr = nr;
    a.k = ops[0]->k;
    for (int p = 0; p < nr; ++p) {
        if (ops[p]->st != st || ops[p]->k != a.k) fail(GT_ERR_ARG, "a batch's operators must share one stack");
        a.r[p] = r[p];
        a.out[p] = z[p];
        a.dinv_sheet[p] = ops[p]->dinv_sheet.p;
        a.dinv_tail[p] = ops[p]->dinv_tail.p;
        a.q[p] = ops[p]->q.p;
        a.s_scale[p] = ops[p]->s_scale;
    }
    a.s = st->work.s.p;
    a.z = st->work.z.p;
    a.t = st->work.t.p;
    a.W = st->W.p;
    a.a_keep = st->a_keep.p;
    a.b_keep = st->b_keep.p;
    a.adj_ptr = st->adj_ptr.p;
    a.adj = st->adj.p;
    a.node_sheet = st->node_sheet.p;
    a.strip_sheet = st->strip_sheet.p;
    if (st->nk > 0) {
        hipLaunchKernelGGL(pre_schur_rhs, dim3(blocks_for(st->nk, 256)), dim3(256), 0, s, a);
        GT_LAUNCH("pre_schur_rhs");
        schur_apply(st, st->work.s.p, st->work.z.p, nr);
        if (a.k > 0) {
            hipLaunchKernelGGL(pre_woodbury_small, dim3(nr), dim3(256), 0, s, a);
            GT_LAUNCH("pre_woodbury_small");
        }
        hipLaunchKernelGGL(pre_woodbury_apply, dim3(blocks_for(st->nk, 256)), dim3(256), 0, s, a);
        GT_LAUNCH("pre_woodbury_apply");
    }
    if (st->nb > 0) {
        hipLaunchKernelGGL(pre_branches, dim3(blocks_for(st->nb, 256)), dim3(256), 0, s, a);
        GT_LAUNCH("pre_branches");
    }
}

}  // namespace gt
