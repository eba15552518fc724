// The operator's device work for a batch of right-hand sides, each with its own frequency (gt_op) on one stack:
// the system product K x (peec.py:3043-3062) and the preconditioner M^-1 r (peec.py:3064-3083):
// y = S^-1 (r2 - Ar^T D^-1 r1), x1 = D^-1 (r1 + Ar y), with S^-1 = (1/d_s) (B + U C U^T)^-1 applied as B^-1 (the
// device triangular solves, all right-hand sides in one pass over the panels) plus the Woodbury correction
// z - W Q (U^T z) (W = B^-1 U resident, Q each frequency's k x k matrix). Loops over the batch sit inside the
// kernels, so the incidence and W are read once per batch. Where W would be too large (many branches on many nodes)
// the implicit form applies B^-1 (U t) by a second batched Schur solve instead.
#include <type_traits>

#include "stack.hpp"

namespace gt {

namespace {

__device__ __forceinline__ zd ldz(const zd& v) { return v; }
__device__ __forceinline__ zd ldz(const cpx<float>& v) { return zd(v.re, v.im); }
template <class T>
__device__ __forceinline__ cpx<T> stz(const zd& v) { return cpx<T>((T)v.re, (T)v.im); }

// T: the storage of the input and output vectors (double, or float for the mixed-precision inner solve); the
// arithmetic and the Schur solve are in double. The right-hand sides may belong to different stacks of one core (a
// board and what-ifs derived from it): the strips, the kept nodes and the Schur factor are common; each has its own
// branches (so its own node offset nb), incidence, node adjacency and Woodbury columns, the first k_sh of which are
// the batch's host stack's (copied into every stack derived from it).
template <class T>
struct PreArgs {
    int ns, nk, nr, k_sh, kt;
    int nb[KMAX], k[KMAX];
    const cpx<T>* r[KMAX];    // n: [r1 ; r2]
    cpx<T>* out[KMAX];        // n: [x1 ; y]
    const zd* dinv_sheet[KMAX];
    const zd* dinv_tail[KMAX];
    const zd* q[KMAX];        // k x k row-major
    zd s_scale[KMAX];
    const zd* s_in[KMAX];     // per-sheet scales (node-column stacks), or null: s_scale alone
    const zd* s_out[KMAX];
    zd* s;                    // nk x nr interleaved: Schur right-hand sides
    const zd* z;              // nk x nr interleaved: B^-1 s
    zd* t;                    // nr x kt: Q (U^T z)
    const zd* z2;             // implicit form: B^-1 (U t), interleaved like z
    // root form (right-hand sides with root[p]): between the Schur solve's passes, u = Y_r^T w_r, t = Q u,
    // w_r -= Y_r t on the forward result zp's last-supernode rows
    int root[KMAX];
    int qmode;                // which right-hand sides pre_woodbury_q takes: 0 all, 1 the root form's, 2 the others
    const double* Yr;
    int yr_c0, yr_n, yr_k;
    zd* zp;
    zd* u;
    const double* W;          // the shared columns: k_sh of nk
    const double* Wx[KMAX];   // each right-hand side's own columns k_sh .. k - 1
    const int* a_keep[KMAX];
    const int* b_keep[KMAX];
    const int* adj_ptr[KMAX];
    const int* adj[KMAX];
    const int* node_sheet;
    const unsigned char* strip_sheet;
};

// s = r2 - Ar^T (D^-1 r1): a kept node's incident strips lie on its own sheet (one admittance), its branches carry
// their own
template <class T>
__global__ __launch_bounds__(64) void pre_schur_rhs(PreArgs<T> a) {
    // a wave's 64 nodes: their right-hand sides staged in LDS, then written as one contiguous run of s (the
    // interleaved rows of consecutive nodes) instead of one strided store per right-hand side
    __shared__ double stg_raw[2 * 64 * KMAX];
    zd* stg = reinterpret_cast<zd*>(stg_raw);
    const int n0 = blockIdx.x * 64, n = n0 + threadIdx.x;
    if (n < a.nk) {
        const int sh = a.node_sheet[n];
        for (int p = 0; p < a.nr; ++p) {
            const cpx<T>* r = a.r[p];
            const int e0 = a.adj_ptr[p][n], e1 = a.adj_ptr[p][n + 1];
            const int* adj = a.adj[p];
            zd acc = ldz(r[a.nb[p] + n]);
            const zd ds = a.dinv_sheet[p][sh];
            for (int e = e0; e < e1; ++e) {
                const int code = adj[e], b = code >> 1;
                const zd v = (b < a.ns ? ds : a.dinv_tail[p][b - a.ns]) * ldz(r[b]);
                if (code & 1) acc = acc + v;
                else acc = acc - v;
            }
            stg[threadIdx.x * a.nr + p] = a.s_in[p] ? a.s_in[p][sh] * acc : acc;
        }
    }
    __syncthreads();
    const int cnt = min(64, a.nk - n0) * a.nr;
    zd* dst = a.s + (size_t)n0 * a.nr;
    for (int e = threadIdx.x; e < cnt; e += 64) dst[e] = stg[e];
}

// t_p = Q_p (U_p^T z_p): one workgroup per right-hand side (k, the branch count, is small)
template <class T>
__global__ __launch_bounds__(256) void pre_woodbury_small(PreArgs<T> a) {
    __shared__ double ut_raw[2 * WB_STORED_MAX_K];
    zd* ut = reinterpret_cast<zd*>(ut_raw);
    const int p = blockIdx.x;
    const int k = a.k[p];
    for (int m = threadIdx.x; m < k; m += 256) {
        const int b = a.ns + m;
        const int ka = a.a_keep[p][b], kb = a.b_keep[p][b];
        zd v;
        if (ka >= 0) v = v + a.z[(size_t)ka * a.nr + p];
        if (kb >= 0) v = v - a.z[(size_t)kb * a.nr + p];
        ut[m] = v;
    }
    __syncthreads();
    const zd* q = a.q[p];
    for (int i = threadIdx.x; i < k; i += 256) {
        zd acc;
        for (int m = 0; m < k; ++m) acc = acc + q[i * k + m] * ut[m];
        a.t[p * a.kt + i] = acc;
    }
}

// The implicit form's t_p = Q_p (U_p^T z_p) for up to KM branches (U^T z staged in LDS: 32 KB for 2048, 128 KB for
// the 8192 of the largest stacks): a workgroup takes 32 rows of one right-hand side, eight lanes a row (m strided by
// eight, then the eight partial sums added by a fixed butterfly)
template <class T, int KM>
__global__ __launch_bounds__(256) void pre_woodbury_q(PreArgs<T> a) {
    __shared__ double ut_raw[2 * KM];
    zd* ut = reinterpret_cast<zd*>(ut_raw);
    const int p = blockIdx.x;
    const int k = a.k[p], i0 = blockIdx.y * 32;
    if (i0 >= k || (a.qmode == 1 && !a.root[p]) || (a.qmode == 2 && a.root[p])) return;
    for (int m = threadIdx.x; m < k; m += 256) {
        zd v;
        if (a.root[p]) {
            v = a.u[p * a.kt + m];
        } else {
            const int b = a.ns + m;
            const int ka = a.a_keep[p][b], kb = a.b_keep[p][b];
            if (ka >= 0) v = v + a.z[(size_t)ka * a.nr + p];
            if (kb >= 0) v = v - a.z[(size_t)kb * a.nr + p];
        }
        ut[m] = v;
    }
    __syncthreads();
    const int i = i0 + (threadIdx.x >> 3), l = threadIdx.x & 7;
    double re = 0, im = 0;
    if (i < k) {
        const zd* q = a.q[p] + (size_t)i * k;
        for (int m = l; m < k; m += 8) {
            const zd qm = q[m], um = ut[m];
            re += qm.re * um.re - qm.im * um.im;
            im += qm.re * um.im + qm.im * um.re;
        }
    }
#pragma unroll
    for (int off = 4; off > 0; off >>= 1) {
        re += __shfl_xor(re, off, 8);
        im += __shfl_xor(im, off, 8);
    }
    if (i < k && l == 0) a.t[p * a.kt + i] = zd(re, im);
}

// The implicit form's s_p = U_p t_p at each kept node: its Woodbury branches (adjacency entries past the strips) in
// adjacency order, + at a branch's a end, - at its b end
template <class T>
__global__ __launch_bounds__(256) void pre_woodbury_scatter(PreArgs<T> a) {
    const int n = blockIdx.x * 256 + threadIdx.x;
    if (n >= a.nk) return;
    for (int p = 0; p < a.nr; ++p) {
        const int e0 = a.adj_ptr[p][n], e1 = a.root[p] ? e0 : a.adj_ptr[p][n + 1];
        const int* adj = a.adj[p];
        const zd* tp = a.t + p * a.kt;
        zd acc;
        for (int e = e0; e < e1; ++e) {
            const int code = adj[e], m = (code >> 1) - a.ns;
            if (m < 0 || m >= a.k[p]) continue;
            if (code & 1) acc = acc - tp[m];
            else acc = acc + tp[m];
        }
        a.s[(size_t)n * a.nr + p] = acc;
    }
}

// The implicit form's y_p = s_scale_p (z_p - B^-1 U_p t_p); the root form's (and every one's without a second solve,
// z2 null) y_p = s_scale_p z_p
template <class T>
__global__ __launch_bounds__(256) void pre_woodbury_diff(PreArgs<T> a) {
    const int n = blockIdx.x * 256 + threadIdx.x;
    if (n >= a.nk) return;
    for (int p = 0; p < a.nr; ++p) {
        const size_t i = (size_t)n * a.nr + p;
        const zd v = a.z2 && !a.root[p] ? a.z[i] - a.z2[i] : a.z[i];
        const zd sc = a.s_out[p] ? a.s_out[p][a.node_sheet[n]] : a.s_scale[p];
        a.out[p][a.nb[p] + n] = stz<T>(sc * v);
    }
}

// The root form's u_p = Y_r^T w_p,r: one workgroup per column m, the yr_n-long sums strided over the threads and
// added by a fixed tree
template <class T>
__global__ __launch_bounds__(256) void pre_root_u(PreArgs<T> a) {
    __shared__ double red[2 * 256];
    const int m = blockIdx.x;
    const double* y = a.Yr + (size_t)m * a.yr_n;
    for (int p = 0; p < a.nr; ++p) {
        if (!a.root[p]) continue;
        double re = 0, im = 0;
        for (int r = threadIdx.x; r < a.yr_n; r += 256) {
            const zd w = a.zp[(size_t)(a.yr_c0 + r) * a.nr + p];
            re += y[r] * w.re;
            im += y[r] * w.im;
        }
        red[threadIdx.x] = re;
        red[256 + threadIdx.x] = im;
        __syncthreads();
        for (int h = 128; h > 0; h >>= 1) {
            if (threadIdx.x < h) {
                red[threadIdx.x] += red[threadIdx.x + h];
                red[256 + threadIdx.x] += red[256 + threadIdx.x + h];
            }
            __syncthreads();
        }
        if (threadIdx.x == 0) a.u[p * a.kt + m] = zd(red[0], red[256]);
        __syncthreads();
    }
}

// The root form's w_p,r -= Y_r t_p: a thread per row r of the last supernode, the sum over m in ascending order
template <class T>
__global__ __launch_bounds__(256) void pre_root_sub(PreArgs<T> a) {
    const int r = blockIdx.x * 256 + threadIdx.x;
    if (r >= a.yr_n) return;
    for (int p = 0; p < a.nr; ++p) {
        if (!a.root[p]) continue;
        const zd* tp = a.t + p * a.kt;
        double re = 0, im = 0;
        for (int m = 0; m < a.yr_k; ++m) {
            const double y = a.Yr[(size_t)m * a.yr_n + r];
            re += y * tp[m].re;
            im += y * tp[m].im;
        }
        zd& w = a.zp[(size_t)(a.yr_c0 + r) * a.nr + p];
        w = w - zd(re, im);
    }
}

// y_p = s_scale_p (z_p - W_p t_p), written to the node part of each output. A thread takes one node for the whole
// batch: each shared W value is read once and applied to every right-hand side (t of the batch staged in LDS, in
// chunks of WB branches); a right-hand side's own columns follow. The sums over m run in ascending m for every
// batch.
constexpr int WB = 64;

template <class T>
__global__ __launch_bounds__(256) void pre_woodbury_apply(PreArgs<T> a) {
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
    for (int m0 = 0; m0 < a.k_sh; m0 += WB) {
        const int mc = min(WB, a.k_sh - m0);
        __syncthreads();
        for (int e = threadIdx.x; e < mc * KMAX; e += 256) {      // right-hand sides past the batch: zeros
            const int p = e / mc, m = e - p * mc;
            ts[m * KMAX + p] = p < a.nr ? a.t[p * a.kt + m0 + m] : zd();
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
    for (int p = 0; p < KMAX; ++p) {
        if (p < a.nr) {
            const double* Wx = a.Wx[p];
            const zd* tp = a.t + p * a.kt;
            for (int m = a.k_sh; m < a.k[p]; ++m) {
                const double w = Wx[(size_t)(m - a.k_sh) * a.nk + n];
                vre[p] -= w * tp[m].re;
                vim[p] -= w * tp[m].im;
            }
            const zd sc = a.s_out[p] ? a.s_out[p][a.node_sheet[n]] : a.s_scale[p];
            a.out[p][a.nb[p] + n] = stz<T>(sc * zd(vre[p], vim[p]));
        }
    }
}

// x1 = D^-1 (r1 + Ar y)
template <class T>
__global__ __launch_bounds__(256) void pre_branches(PreArgs<T> a) {
    const int b = blockIdx.x * 256 + threadIdx.x;
    const int sh = b < a.ns ? a.strip_sheet[b] : -1;
    for (int p = 0; p < a.nr; ++p) {
        if (b >= a.nb[p]) continue;
        const int ka = a.a_keep[p][b], kb = a.b_keep[p][b];
        zd v = ldz(a.r[p][b]);
        const cpx<T>* y = a.out[p] + a.nb[p];
        if (ka >= 0) v = v + ldz(y[ka]);
        if (kb >= 0) v = v - ldz(y[kb]);
        const zd d = sh >= 0 ? a.dinv_sheet[p][sh] : a.dinv_tail[p][b - a.ns];
        a.out[p][b] = stz<T>(d * v);
    }
}

}  // namespace

void Work::ensure(const gt_stack* st, int nr, int kt) {
    kt = std::max(kt, st->nbr);
    if (nr > cap) {
        T1.alloc(std::max<size_t>(1, st->t_elems) * nr);
        T2.alloc(std::max<size_t>(1, st->t_elems) * nr);
        s.alloc((size_t)std::max(1, st->nk) * nr);
        z.alloc((size_t)std::max(1, st->nk) * nr);
    }
    if (nr > cap || kt > t_cols) {
        t_cols = std::max(kt, t_cols);
        t.alloc((size_t)std::max(1, t_cols) * std::max(nr, cap));
        u.alloc((size_t)std::max(1, t_cols) * std::max(nr, cap));
    }
    cap = std::max(cap, nr);
    if (st->wb_implicit && z2.n < (size_t)std::max(1, st->nk) * cap) z2.alloc((size_t)std::max(1, st->nk) * cap);
}

void Work::shrink(const gt_stack* st, int nr) {
    if (nr >= cap) return;
    const int kt = t_cols;
    T1.release();
    T2.release();
    s.release();
    z.release();
    z2.release();
    t.release();
    u.release();
    cap = 0;
    t_cols = 0;
    if (nr > 0) ensure(st, nr, kt);
}

void Work::reset(const gt_stack* st) {
    T1.release();
    T2.release();
    s.release();
    z.release();
    z2.release();
    t.release();
    u.release();
    cap = 0;
    t_cols = 0;
    ensure(st, 1);
}

int64_t Work::bytes_after(const gt_stack* st, int nr, int kt) const {
    kt = std::max(kt, st->nbr);
    const int64_t zb = sizeof(zd);
    int64_t b = nr > cap ? (int64_t)(2 * std::max<size_t>(1, st->t_elems) + 2 * (size_t)std::max(1, st->nk)) *
                               nr * zb
                         : (int64_t)(T1.bytes() + T2.bytes() + s.bytes() + z.bytes());
    if (nr > cap || kt > t_cols) b += 2 * (int64_t)std::max(1, std::max(kt, t_cols)) * std::max(nr, cap) * zb;
    else b += (int64_t)(t.bytes() + u.bytes());
    if (st->wb_implicit) b += std::max<int64_t>((int64_t)z2.bytes(), (int64_t)std::max(1, st->nk) * std::max(nr, cap) * zb);
    return b;
}

// The right-hand sides of one batch on the host stack st: each operator's stack is st or one derived from it (a
// what-if sharing st's core, with st's Woodbury columns first)
void check_batch(const gt_stack* st, int nr, gt_op* const* ops) {
    if (nr < 1 || nr > KMAX) fail(GT_ERR_ARG, "1 to 16 right-hand sides per batch");
    for (int p = 0; p < nr; ++p) {
        const gt_stack* o = ops[p]->st;
        if (o == st) continue;
        bool ok = o->core == st->core && o->k >= st->k && o->k_root == st->k_root;
        if (ok) {
            ok = false;
            for (uint64_t id : o->lineage) ok = ok || id == st->id;
        }
        if (!ok) fail(GT_ERR_ARG, "a batch's operators must belong to one stack or to what-ifs derived from it");
    }
}

namespace {

// The L_p arguments of a product in storage type T (the double spectra and twiddles, or their float copies; the
// scratch T1 / T2 viewed as T)
template <class T>
LpArgs<T> lp_args_t(const gt_stack* st) {
    if constexpr (std::is_same<T, double>::value) {
        return st->lp_args(1);
    } else {
        const LpArgs<double> d = st->lp_args(1);
        LpArgs<float> a{};
        for (int oi = 0; oi < 2; ++oi) {
            const OrientHost& h = st->o[oi];
            const OrientArgs<double>& od = d.o[oi];
            OrientArgs<float>& o = a.o[oi];
            o.n0 = od.n0;
            o.n1 = od.n1;
            o.P0 = od.P0;
            o.P1 = od.P1;
            o.log0 = od.log0;
            o.log1 = od.log1;
            o.TR = od.TR;
            o.lo = od.lo;
            o.cnt = od.cnt;
            o.rowptr = od.rowptr;
            o.rowcol = od.rowcol;
            o.K = h.Kf->p;
            o.T1 = reinterpret_cast<zf*>(st->work.T1.p) + h.t_off;
            o.T2 = reinterpret_cast<zf*>(st->work.T2.p) + h.t_off;
            o.tw0 = h.present ? st->twf.at(h.P0).p : nullptr;
            o.tw1 = h.present ? st->twf.at(h.P1).p : nullptr;
            o.twq = h.present && st->shape.split ? st->twf.at(h.P0 / 4).p : nullptr;
            a.nblk_rows[oi] = d.nblk_rows[oi];
            a.nblk_cols[oi] = d.nblk_cols[oi];
        }
        a.S = d.S;
        a.mode = 1;
        a.t_stride = d.t_stride;
        a.G = d.G;
        a.gmap = d.gmap;
        std::copy(std::begin(d.gpad), std::end(d.gpad), std::begin(a.gpad));
        a.toeplitz = d.toeplitz;
        a.a_keep = d.a_keep;
        a.b_keep = d.b_keep;
        return a;
    }
}

}  // namespace

namespace {

template <class T>
struct NarrowArgs {
    const double* fac;              // ns: 1 / r_width - 1
    const double* dl;               // ns: the narrow self term (H)
    const int* ptr;                 // side-by-side corrections (H), CSR over the strips
    const int* col;
    const double* val;
    const unsigned char* sheet;     // ns
    int ns;
    const cpx<T>* x[KMAX];
    cpx<T>* y[KMAX];
    const cpx<T>* r_sheet[KMAX];    // NULL: the L_p product's corrections alone
    T w[KMAX];
};

// y_s += z_sheet (1/r_width - 1) x_s + j w (dL_s x_s + sum_k M_sk x_k) (peec.py:2129-2132, :2159); one thread per strip,
// blockIdx.y the right-hand side; the neighbours summed in CSR order
template <class T>
__global__ __launch_bounds__(256) void op_narrow(NarrowArgs<T> a) {
    const int s = blockIdx.x * 256 + threadIdx.x, p = blockIdx.y;
    if (s >= a.ns) return;
    const cpx<T>* x = a.x[p];
    const cpx<T> xs = x[s];
    const T dl = (T)a.dl[s];
    T lr = dl * xs.re, li = dl * xs.im;
    for (int k = a.ptr[s]; k < a.ptr[s + 1]; ++k) {
        const T m = (T)a.val[k];
        const cpx<T> xk = x[a.col[k]];
        lr += m * xk.re;
        li += m * xk.im;
    }
    cpx<T> add;
    if (a.r_sheet[p]) {
        const cpx<T> z = a.r_sheet[p][a.sheet[s]];
        const T f = (T)a.fac[s], w = a.w[p];
        add = cpx<T>(f * (z.re * xs.re - z.im * xs.im) - w * li, f * (z.re * xs.im + z.im * xs.re) + w * lr);
    } else {
        add = cpx<T>(lr, li);
    }
    cpx<T>* y = a.y[p];
    y[s] = y[s] + add;
}

}  // namespace

template <class T>
void narrow_apply(const gt_stack* st, int nr, gt_op* const* ops, const cpx<T>* const* x, cpx<T>* const* y) {
    if (!st->core->narrow || st->ns == 0) return;
    const StackCore& c = *st->core;
    NarrowArgs<T> a{};
    a.fac = c.nw_fac.p;
    a.dl = c.nw_dl.p;
    a.ptr = c.nw_ptr.p;
    a.col = c.nw_col.p;
    a.val = c.nw_val.p;
    a.sheet = c.strip_sheet.p;
    a.ns = st->ns;
    for (int p = 0; p < nr; ++p) {
        a.x[p] = x[p];
        a.y[p] = y[p];
        if (ops) {
            if constexpr (std::is_same<T, double>::value) a.r_sheet[p] = ops[p]->r_sheet.p;
            else a.r_sheet[p] = ops[p]->r_sheet_f.p;
            a.w[p] = (T)ops[p]->w;
        }
    }
    hipLaunchKernelGGL((op_narrow<T>), dim3((unsigned)((st->ns + 255) / 256), nr), dim3(256), 0, st->ctx->stream,
                       a);
    GT_LAUNCH("op_narrow");
}

template <class T>
void ops_matvec(gt_stack* st, int nr, gt_op* const* ops, const cpx<T>* const* x, cpx<T>* const* y) {
    check_batch(st, nr, ops);
    hipStream_t s = st->ctx->stream;
    st->work.ensure(st, nr);
    RhsSet<T> rs{};
    rs.nr = nr;
    for (int p = 0; p < nr; ++p) {
        rs.x[p] = x[p];
        rs.y[p] = y[p];
        rs.nb[p] = ops[p]->st->nb;
        rs.w[p] = (T)ops[p]->w;
        if constexpr (std::is_same<T, double>::value) {
            rs.r_sheet[p] = ops[p]->r_sheet.p;
            rs.z_tail[p] = ops[p]->z_tail.p;
        } else {
            rs.r_sheet[p] = ops[p]->r_sheet_f.p;
            rs.z_tail[p] = ops[p]->z_tail_f.p;
        }
    }
    LpArgs<T> la = lp_args_t<T>(st);
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
    TailArgs<T> ta{};
    ta.ns = st->ns;
    ta.nk = st->nk;
    for (int p = 0; p < nr; ++p) {
        const gt_stack* o = ops[p]->st;
        ta.nbr[p] = o->nbr;
        ta.a_keep[p] = o->a_keep.p;
        ta.b_keep[p] = o->b_keep.p;
        ta.adj_ptr[p] = o->adj_ptr.p;
        ta.adj[p] = o->adj.p;
    }
    ta.rhs = rs;
    launch_tail_bottom(ta, s);
    narrow_apply<T>(st, nr, ops, x, y);
}

void schur_apply(gt_stack* st, const zd* b, zd* x, int nr) {
    if (!st->schur) fail(GT_ERR_ARG, "the stack has no Schur factor (gt_schur_factor first)");
    schur_solve_device(*st->schur, b, x, nr, st->ctx->stream, st->ctx->graphs);
}

namespace {

struct UnitCols {
    int ka[KMAX], kb[KMAX];   // kept end nodes of the added branches (-1: a reference node)
};

// b (n_keep x nr interleaved) = the incidence columns U' e_c of nr added branches: +1 at node a, -1 at node b
__global__ __launch_bounds__(256) void wb_units(zd* __restrict__ b, int nk, int nr, UnitCols u) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= (size_t)nk * nr) return;
    const int v = (int)(i / nr), c = (int)(i % nr);
    const double val = (v == u.ka[c] ? 1.0 : 0.0) - (v == u.kb[c] ? 1.0 : 0.0);
    b[i] = zd(val, 0.0);
}

// W[c] (a column of n_keep, c < nr) = Re x[:, c] of the interleaved solutions
__global__ __launch_bounds__(256) void wb_take(const zd* __restrict__ x, int nk, int nr, double* __restrict__ W) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= (size_t)nk * nr) return;
    const int v = (int)(i / nr), c = (int)(i % nr);
    W[(size_t)c * nk + v] = x[i].re;
}

// G = U^T W (k x k, row-major): G[r][c] = W[c][a_r] - W[c][b_r] over the branches r (kept ends; -1 a reference node)
__global__ __launch_bounds__(256) void wb_gram(const double* __restrict__ W, int nk, int k, const int* __restrict__ ka,
                                               const int* __restrict__ kb, double* __restrict__ G) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= k * k) return;
    const int r = i / k, c = i % k;
    double v = 0;
    if (ka[r] >= 0) v += W[(size_t)c * nk + ka[r]];
    if (kb[r] >= 0) v -= W[(size_t)c * nk + kb[r]];
    G[i] = v;
}

// The implicit form's G[r][c0 + c] = (U^T Z)[r][c] over the k branches r, from nr solved columns Z (interleaved)
__global__ __launch_bounds__(256) void wb_gram_cols(const zd* __restrict__ z, int nr, int k, int c0,
                                                    const int* __restrict__ ka, const int* __restrict__ kb,
                                                    double* __restrict__ G) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= k * nr) return;
    const int r = i / nr, c = i % nr;
    double v = 0;
    if (ka[r] >= 0) v += z[(size_t)ka[r] * nr + c].re;
    if (kb[r] >= 0) v -= z[(size_t)kb[r] * nr + c].re;
    G[(size_t)r * k + c0 + c] = v;
}

// A derived stack's G (k x k): the base's k0 x k0 block copied
__global__ __launch_bounds__(256) void wb_gram_base(const double* __restrict__ G0, int k0, double* __restrict__ G,
                                                    int k) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= k0 * k0) return;
    G[(size_t)(i / k0) * k + i % k0] = G0[i];
}

// ... and its added rows c >= k0 over the base's columns r < k0 by symmetry: G[c][r] = G[r][c]
__global__ __launch_bounds__(256) void wb_gram_mirror(double* __restrict__ G, int k, int k0) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= (k - k0) * k0) return;
    const int c = k0 + i / k0, r = i % k0;
    G[(size_t)c * k + r] = G[(size_t)r * k + c];
}

// The root form's Y_r columns c0 .. c0 + nr - 1: Re of the forward results' last-supernode rows
__global__ __launch_bounds__(256) void wb_take_root(const zd* __restrict__ zp, int nr, int yc0, int yn, int c0,
                                                    double* __restrict__ Yr) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= yn * nr) return;
    const int r = i / nr, c = i % nr;
    Yr[(size_t)(c0 + c) * yn + r] = zp[(size_t)(yc0 + r) * nr + c].re;
}

// The root form's G = Y_r^T Y_r (k x k row-major; each entry's sum over rows in ascending order, so G is symmetric)
__global__ __launch_bounds__(256) void wb_gram_root(const double* __restrict__ Yr, int yn, int k, double* __restrict__ G) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= k * k) return;
    const int r = i / k, c = i % k;
    const double* a = Yr + (size_t)min(r, c) * yn;
    const double* b = Yr + (size_t)max(r, c) * yn;
    double v = 0;
    for (int q = 0; q < yn; ++q) v += a[q] * b[q];
    G[i] = v;
}

// Every end of the branches in the factor's last supernode (whose rows are then Y's only nonzero ones)
bool ends_in_last_supernode(const gt_stack* st, int n, const int* ka, const int* kb) {
    const SchurHost& H = st->schur->h;
    if (H.c0.empty()) return false;
    const int c0 = H.c0.back();
    if (c0 + H.nc.back() != H.n) return false;
    std::vector<char> last(H.n, 0);
    for (int c = c0; c < H.n; ++c) last[H.perm[c]] = 1;
    for (int m = 0; m < n; ++m)
        if ((ka[m] >= 0 && !last[ka[m]]) || (kb[m] >= 0 && !last[kb[m]])) return false;
    return true;
}

}  // namespace

void woodbury_extend(gt_stack* st, const gt_stack* base, int n_extra, const int* a_extra, const int* b_extra) {
    hipStream_t s = st->ctx->stream;
    const int nk = st->nk, k0 = base ? base->k : 0, k = k0 + n_extra;
    const bool imp = st->wb_implicit;
    if (imp) {
        if (k > WB_IMPLICIT_MAX_K) fail(GT_ERR_ARG, "more than 8192 branches on one stack");
        if (base && !base->wb_implicit) fail(GT_ERR_ARG, "a derived stack takes its base's Woodbury form");
        st->G.alloc(std::max<size_t>(1, (size_t)k * k));
        GT_HIP(hipMemsetAsync(st->G.p, 0, st->G.bytes(), s));
        if (k0 > 0) {
            hipLaunchKernelGGL(wb_gram_base, dim3(blocks_for((size_t)k0 * k0, 256)), dim3(256), 0, s, base->G.p, k0,
                               st->G.p, k);
            GT_LAUNCH("wb_gram_base");
        }
    } else {
        st->W.alloc(std::max<size_t>(1, (size_t)k * nk));
        if (k0 > 0 && nk > 0)
            GT_HIP(hipMemcpyAsync(st->W.p, base->W.p, sizeof(double) * (size_t)k0 * nk, hipMemcpyDeviceToDevice, s));
    }
    st->wb_root = imp && !base && n_extra > 0 && nk > 0 && ends_in_last_supernode(st, n_extra, a_extra, b_extra);
    if (st->wb_root) {
        const SchurHost& H = st->schur->h;
        st->yr_c0 = H.c0.back();
        st->yr_n = H.n - st->yr_c0;
        st->Yr.alloc((size_t)st->yr_n * k);
    }
    if (n_extra > 0 && nk > 0) {
        // the columns' solves in batches of up to 16, on scratch of their own (the product workspace and the
        // Schur solve's buffers go back to what they held); the root form takes only the forward passes
        const int chunk = std::min(n_extra, KMAX);
        const int schur_cap = st->schur->cap;
        DevBuf<zd> sb((size_t)nk * chunk), zb((size_t)nk * chunk);
        for (int m0 = 0; m0 < n_extra; m0 += KMAX) {
            const int nr = std::min(KMAX, n_extra - m0);
            UnitCols u{};
            for (int c = 0; c < nr; ++c) {
                u.ka[c] = a_extra[m0 + c];
                u.kb[c] = b_extra[m0 + c];
            }
            const size_t cnt = (size_t)nk * nr;
            hipLaunchKernelGGL(wb_units, dim3(blocks_for(cnt, 256)), dim3(256), 0, s, sb.p, nk, nr, u);
            GT_LAUNCH("wb_units");
            if (st->wb_root) {
                schur_forward_last_device(*st->schur, sb.p, nr, s);
                hipLaunchKernelGGL(wb_take_root, dim3(blocks_for((size_t)st->yr_n * nr, 256)), dim3(256), 0, s,
                                   st->schur->zp.p, nr, st->yr_c0, st->yr_n, m0, st->Yr.p);
                GT_LAUNCH("wb_take_root");
                continue;
            }
            schur_apply(st, sb.p, zb.p, nr);
            if (imp) {
                hipLaunchKernelGGL(wb_gram_cols, dim3(blocks_for((size_t)k * nr, 256)), dim3(256), 0, s, zb.p, nr,
                                   k, k0 + m0, st->a_keep.p + st->ns, st->b_keep.p + st->ns, st->G.p);
                GT_LAUNCH("wb_gram_cols");
            } else {
                hipLaunchKernelGGL(wb_take, dim3(blocks_for(cnt, 256)), dim3(256), 0, s, zb.p, nk, nr,
                                   st->W.p + (size_t)(k0 + m0) * nk);
                GT_LAUNCH("wb_take");
            }
        }
        if (st->wb_root) {
            hipLaunchKernelGGL(wb_gram_root, dim3(blocks_for((size_t)k * k, 256)), dim3(256), 0, s, st->Yr.p,
                               st->yr_n, k, st->G.p);
            GT_LAUNCH("wb_gram_root");
        }
        GT_HIP(hipStreamSynchronize(s));
        st->schur->shrink(schur_cap);
        if (imp && k0 > 0) {
            hipLaunchKernelGGL(wb_gram_mirror, dim3(blocks_for((size_t)n_extra * k0, 256)), dim3(256), 0, s, st->G.p,
                               k, k0);
            GT_LAUNCH("wb_gram_mirror");
        }
    }
    st->k = k;
    st->k_root = base ? base->k_root : k;
}

void woodbury_gram_device(const gt_stack* st, double* G_host) {
    const int k = st->k;
    if (k == 0) return;
    hipStream_t s = st->ctx->stream;
    if (st->wb_implicit) {
        GT_HIP(hipMemcpyAsync(G_host, st->G.p, sizeof(double) * (size_t)k * k, hipMemcpyDeviceToHost, s));
        GT_HIP(hipStreamSynchronize(s));
        return;
    }
    DevBuf<double> G((size_t)k * k);
    hipLaunchKernelGGL(wb_gram, dim3(blocks_for((size_t)k * k, 256)), dim3(256), 0, s, st->W.p, st->nk, k,
                       st->a_keep.p + st->ns, st->b_keep.p + st->ns, G.p);
    GT_LAUNCH("wb_gram");
    GT_HIP(hipMemcpyAsync(G_host, G.p, sizeof(double) * (size_t)k * k, hipMemcpyDeviceToHost, s));
    GT_HIP(hipStreamSynchronize(s));
}

template <class T>
void ops_precond(gt_stack* st, int nr, gt_op* const* ops, const cpx<T>* const* r, cpx<T>* const* z) {
    check_batch(st, nr, ops);
    hipStream_t s = st->ctx->stream;
    PreArgs<T> a{};
    a.ns = st->ns;
    a.nk = st->nk;
    a.nr = nr;
    const bool imp = st->wb_implicit;    // every stack of a batch has its lineage's form
    a.k_sh = imp ? 0 : st->k_root;       // the lineage's first columns: the same split alone and in a board's batch
    a.kt = 0;
    int nbmax = 0;
    for (int p = 0; p < nr; ++p) {
        const gt_stack* o = ops[p]->st;
        if (ops[p]->k != o->k) fail(GT_ERR_ARG, "the operator's Woodbury size differs from its stack's");
        if (!imp && o->k > WB_STORED_MAX_K) fail(GT_ERR_ARG, "more than 256 branches on one stack");
        a.kt = std::max(a.kt, o->k);
        nbmax = std::max(nbmax, o->nb);
        a.nb[p] = o->nb;
        a.k[p] = o->k;
        a.r[p] = r[p];
        a.out[p] = z[p];
        a.dinv_sheet[p] = ops[p]->dinv_sheet.p;
        a.dinv_tail[p] = ops[p]->dinv_tail.p;
        a.q[p] = ops[p]->q.p;
        a.s_scale[p] = ops[p]->s_scale;
        a.s_in[p] = ops[p]->per_sheet ? ops[p]->s_in.p : nullptr;
        a.s_out[p] = ops[p]->per_sheet ? ops[p]->s_out.p : nullptr;
        a.Wx[p] = !imp && o->W.p ? o->W.p + (size_t)a.k_sh * st->nk : nullptr;
        a.a_keep[p] = o->a_keep.p;
        a.b_keep[p] = o->b_keep.p;
        a.adj_ptr[p] = o->adj_ptr.p;
        a.adj[p] = o->adj.p;
    }
    st->work.ensure(st, nr, a.kt);
    a.s = st->work.s.p;
    a.z = st->work.z.p;
    a.t = st->work.t.p;
    a.W = st->W.p;
    a.node_sheet = st->node_sheet.p;
    a.strip_sheet = st->strip_sheet.p;
    if (st->nk > 0) {
        hipLaunchKernelGGL(pre_schur_rhs<T>, dim3(blocks_for(st->nk, 64)), dim3(64), 0, s, a);
        GT_LAUNCH("pre_schur_rhs");
        if (imp && a.kt > 0) {
            // the host stack's own right-hand sides in the root form, the others (what-ifs) with a second solve;
            // each right-hand side's arithmetic is the same in any batch
            bool any_root = false, any_second = false;
            for (int p = 0; p < nr; ++p) {
                a.root[p] = ops[p]->st == st && st->wb_root && a.k[p] > 0;
                any_root = any_root || a.root[p];
                any_second = any_second || (!a.root[p] && a.k[p] > 0);
            }
            if (any_root) {
                Schur& S = *st->schur;
                a.Yr = st->Yr.p;
                a.yr_c0 = st->yr_c0;
                a.yr_n = st->yr_n;
                a.yr_k = st->k;
                a.u = st->work.u.p;
                schur_forward_device(S, st->work.s.p, nr, s);
                a.zp = S.zp.p;
                hipLaunchKernelGGL(pre_root_u<T>, dim3(st->k), dim3(256), 0, s, a);
                GT_LAUNCH("pre_root_u");
                a.qmode = 1;
                if (a.kt <= WB_IMPLICIT_SMALL_K)
                    hipLaunchKernelGGL((pre_woodbury_q<T, WB_IMPLICIT_SMALL_K>), dim3(nr, (a.kt + 31) / 32), dim3(256),
                                       0, s, a);
                else
                    hipLaunchKernelGGL((pre_woodbury_q<T, WB_IMPLICIT_MAX_K>), dim3(nr, (a.kt + 31) / 32), dim3(256),
                                       0, s, a);
                GT_LAUNCH("pre_woodbury_q");
                hipLaunchKernelGGL(pre_root_sub<T>, dim3(blocks_for(st->yr_n, 256)), dim3(256), 0, s, a);
                GT_LAUNCH("pre_root_sub");
                schur_backward_device(S, st->work.z.p, nr, s);
            } else {
                schur_apply(st, st->work.s.p, st->work.z.p, nr);
            }
            if (any_second) {
                a.z2 = st->work.z2.p;
                a.qmode = 2;
                if (a.kt <= WB_IMPLICIT_SMALL_K)
                    hipLaunchKernelGGL((pre_woodbury_q<T, WB_IMPLICIT_SMALL_K>), dim3(nr, (a.kt + 31) / 32), dim3(256),
                                       0, s, a);
                else
                    hipLaunchKernelGGL((pre_woodbury_q<T, WB_IMPLICIT_MAX_K>), dim3(nr, (a.kt + 31) / 32), dim3(256),
                                       0, s, a);
                GT_LAUNCH("pre_woodbury_q");
                hipLaunchKernelGGL(pre_woodbury_scatter<T>, dim3(blocks_for(st->nk, 256)), dim3(256), 0, s, a);
                GT_LAUNCH("pre_woodbury_scatter");
                schur_apply(st, st->work.s.p, st->work.z2.p, nr);
            }
            hipLaunchKernelGGL(pre_woodbury_diff<T>, dim3(blocks_for(st->nk, 256)), dim3(256), 0, s, a);
            GT_LAUNCH("pre_woodbury_diff");
        } else {
            schur_apply(st, st->work.s.p, st->work.z.p, nr);
            if (a.kt > 0) {
                hipLaunchKernelGGL(pre_woodbury_small<T>, dim3(nr), dim3(256), 0, s, a);
                GT_LAUNCH("pre_woodbury_small");
            }
            hipLaunchKernelGGL(pre_woodbury_apply<T>, dim3(blocks_for(st->nk, 256)), dim3(256), 0, s, a);
            GT_LAUNCH("pre_woodbury_apply");
        }
    }
    if (nbmax > 0) {
        hipLaunchKernelGGL(pre_branches<T>, dim3(blocks_for(nbmax, 256)), dim3(256), 0, s, a);
        GT_LAUNCH("pre_branches");
    }
}

template void ops_matvec<double>(gt_stack*, int, gt_op* const*, const zd* const*, zd* const*);
template void narrow_apply<double>(const gt_stack*, int, gt_op* const*, const zd* const*, zd* const*);
template void narrow_apply<float>(const gt_stack*, int, gt_op* const*, const zf* const*, zf* const*);
template void ops_matvec<float>(gt_stack*, int, gt_op* const*, const zf* const*, zf* const*);
template void ops_precond<double>(gt_stack*, int, gt_op* const*, const zd* const*, zd* const*);
template void ops_precond<float>(gt_stack*, int, gt_op* const*, const zf* const*, zf* const*);

}  // namespace gt
