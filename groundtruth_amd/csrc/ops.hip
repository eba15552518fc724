// The operator's device work: the system product K x (peec.py:3043-3062) and the preconditioner M^-1 r
// (peec.py:3064-3083): y = S^-1 (r2 - Ar^T D^-1 r1), x1 = D^-1 (r1 + Ar y), with
// S^-1 = (1/d_s) (B + U C U^T)^-1 applied as B^-1 (device triangular solves) plus the Woodbury correction
// z - W Q (U^T z) (W = B^-1 U resident, Q the frequency's k x k matrix).
#include "stack.hpp"

namespace gt {

namespace {

struct PreArgs {
    int ns, nb, nk, k;
    const zd* r;              // n: [r1 ; r2]
    zd* out;                  // n: [x1 ; y]
    zd* s;                    // nk: Schur right-hand side
    const zd* z;              // nk: B^-1 s
    zd* t;                    // k: U^T z, then Q U^T z
    const zd* dinv_sheet;
    const zd* dinv_tail;
    const zd* q;              // k x k row-major
    const double* W;          // k columns of nk
    zd s_scale;
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
    zd acc = a.r[a.nb + n];
    const zd ds = a.dinv_sheet[a.node_sheet[n]];
    for (int e = a.adj_ptr[n]; e < a.adj_ptr[n + 1]; ++e) {
        const int code = a.adj[e], b = code >> 1;
        const zd v = (b < a.ns ? ds : a.dinv_tail[b - a.ns]) * a.r[b];
        if (code & 1) acc = acc + v;
        else acc = acc - v;
    }
    a.s[n] = acc;
}

// t = Q (U^T z): one workgroup (k is the number of branches, small)
__global__ __launch_bounds__(256) void pre_woodbury_small(PreArgs a) {
    __shared__ double ut_raw[512];
    zd* ut = reinterpret_cast<zd*>(ut_raw);
    for (int m = threadIdx.x; m < a.k; m += 256) {
        const int b = a.ns + m;
        const int ka = a.a_keep[b], kb = a.b_keep[b];
        zd v;
        if (ka >= 0) v = v + a.z[ka];
        if (kb >= 0) v = v - a.z[kb];
        ut[m] = v;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < a.k; i += 256) {
        zd acc;
        for (int m = 0; m < a.k; ++m) acc = acc + a.q[i * a.k + m] * ut[m];
        a.t[i] = acc;
    }
}

// y = s_scale (z - W t), written to the node part of the output
__global__ __launch_bounds__(256) void pre_woodbury_apply(PreArgs a) {
    __shared__ double ts_raw[512];
    zd* ts = reinterpret_cast<zd*>(ts_raw);
    for (int m = threadIdx.x; m < a.k; m += 256) ts[m] = a.t[m];
    __syncthreads();
    const int n = blockIdx.x * 256 + threadIdx.x;
    if (n >= a.nk) return;
    zd v = a.z[n];
    for (int m = 0; m < a.k; ++m) {
        const double w = a.W[(size_t)m * a.nk + n];
        v.re -= w * ts[m].re;
        v.im -= w * ts[m].im;
    }
    a.out[a.nb + n] = a.s_scale * v;
}

// x1 = D^-1 (r1 + Ar y)
__global__ __launch_bounds__(256) void pre_branches(PreArgs a) {
    const int b = blockIdx.x * 256 + threadIdx.x;
    if (b >= a.nb) return;
    zd v = a.r[b];
    const int ka = a.a_keep[b], kb = a.b_keep[b];
    if (ka >= 0) v = v + a.out[a.nb + ka];
    if (kb >= 0) v = v - a.out[a.nb + kb];
    const zd d = b < a.ns ? a.dinv_sheet[a.strip_sheet[b]] : a.dinv_tail[b - a.ns];
    a.out[b] = d * v;
}

}  // namespace

void op_matvec(gt_op* op, const zd* x, zd* y) {
    gt_stack* st = op->st;
    hipStream_t s = st->ctx->stream;
    LpArgs<double> la = st->lp_args(x, y, 1);
    la.r_sheet = op->r_sheet.p;
    la.w = op->w;
    if (st->ns > 0) launch_lp(la, st->shape, s);
    TailArgs<double> ta{st->ns, st->nb, st->nbr, st->nk, x, y, op->z_tail.p, st->a_keep.p, st->b_keep.p,
                        st->adj_ptr.p, st->adj.p};
    launch_tail_bottom(ta, s);
}

void schur_apply(gt_stack* st, const zd* b, zd* x) {
    if (!st->schur) fail(GT_ERR_ARG, "the stack has no Schur factor (gt_schur_factor first)");
    schur_solve_device(*st->schur, b, x, st->ctx->stream);
}

void op_precond(gt_op* op, const zd* r, zd* z) {
    gt_stack* st = op->st;
    hipStream_t s = st->ctx->stream;
    PreArgs a{st->ns, st->nb, st->nk, op->k, r, z, st->work.s.p, st->work.z.p, st->work.t.p, op->dinv_sheet.p,
              op->dinv_tail.p, op->q.p, st->W.p, op->s_scale, st->a_keep.p, st->b_keep.p, st->adj_ptr.p,
              st->adj.p, st->node_sheet.p, st->strip_sheet.p};
    if (st->nk > 0) {
        hipLaunchKernelGGL(pre_schur_rhs, dim3(blocks_for(st->nk, 256)), dim3(256), 0, s, a);
        GT_LAUNCH("pre_schur_rhs");
        schur_apply(st, st->work.s.p, st->work.z.p);
        if (op->k > 0) {
            hipLaunchKernelGGL(pre_woodbury_small, dim3(1), dim3(256), 0, s, a);
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
