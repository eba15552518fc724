// Restarted, right-preconditioned GMRES on the device for a batch of right-hand sides on one stack, each with its
// own operator (frequency) and right-hand side (peec.py:3093-3224 `_gmres`, called as peec.py:2581 does:
// (K M^-1) y = b, x = M^-1 y).
//
// Every right-hand side follows the reference's control flow exactly (CGS with a DGKS re-pass below 0.7, Givens
// rotations, the true-residual check at the aim with its miss budget, restarts, extra cycles while each falls to
// 0.7), as its own state machine on the host. The machines advance in lockstep: each phase (cycle starts, Arnoldi
// steps, re-passes, true-residual checks, cycle ends) is one batched launch sequence over the right-hand sides
// that are in it, so the kernel spectra, the incidence and the Schur panels are read once per batch. A right-hand
// side's arithmetic does not depend on what else is in its batch: its result is bit-identical alone and in any mix.
// The bases and vectors stay in HBM; per phase one copy brings the Hessenberg columns and norms to the host.
#include <cmath>
#include <complex>
#include <limits>
#include <vector>

#include "krylov.hpp"
#include "stack.hpp"

namespace gt {

namespace {

using cd = std::complex<double>;

constexpr double PROGRESS = 0.7;

enum State { START, ARNOLDI, DONE };

// A device pointer in the role of the buffers it replaces (`.p`)
struct Vec {
    zd* p = nullptr;
};

struct Rhs {
    gt_op* op = nullptr;
    const zd* b = nullptr;                 // right-hand side (device)
    Vec V, w, x, xt, r, t, u;              // basis (restart + 1 rows) and vectors, in the context's workspace
    State state = START;
    bool conv = false;
    double bnorm = 0, last = 0, aim = 0, resid = 0, hn = 0;
    int cycle = 0, its = 0, j = 0, k = 0, misses = 0, products = 0;
    std::vector<cd> H, sn, g, h;
    std::vector<double> cs;
};

}  // namespace

// The device (and pinned host) memory of a GMRES batch. The context keeps it between calls, grown to the largest
// vector length, basis and batch it has served: a solve allocates nothing once its size has been seen.
struct GmresWork {
    size_t n = 0;                          // vector length the slots hold
    int rows = 0;                          // basis rows the slots hold
    int slots = 0;                         // right-hand-side slots allocated at that size
    DevBuf<zd> V[KMAX], vec[KMAX];         // per slot: the basis, and six vectors of n
    Krylov kr;
    int kr_rows = 0;
    size_t kr_n = 0;
    DevBuf<zd> ydev;                       // combination coefficients, rows per right-hand side
    HostBuf<cd> ybuf;
    HostBuf<double> red;

    void ensure(size_t n_, int rows_, int nr) {
        if (n_ > n || rows_ > rows) {
            for (int i = 0; i < slots; ++i) {
                V[i].release();
                vec[i].release();
            }
            slots = 0;
            n = std::max(n, n_);
            rows = std::max(rows, rows_);
            ydev.alloc((size_t)KMAX * rows);
            ybuf.alloc((size_t)KMAX * rows);
        }
        for (; slots < nr; ++slots) {
            V[slots].alloc((size_t)rows * n);
            vec[slots].alloc(6 * n);
        }
        if (kr_n != n_ || kr_rows != rows_) {
            kr.init((int)n_, rows_);
            red.alloc((size_t)KMAX * kr.P);
            kr_n = n_;
            kr_rows = rows_;
        }
    }
};

namespace {

struct Batch {
    gt_stack* st;
    hipStream_t s;
    int n, restart, nr;
    gt_gmres_params p;
    std::vector<Rhs> R;
    GmresWork& gw;
    Krylov& kr;
    DevBuf<zd>& ydev;
    HostBuf<cd>& ybuf;
    HostBuf<double>& red;

    Batch(gt_stack* st_, int nr_, int restart_, const gt_gmres_params& p_, GmresWork& gw_)
        : st(st_), s(st_->ctx->stream), n(st_->nb + st_->nk), restart(restart_), nr(nr_), p(p_), R(nr_), gw(gw_),
          kr(gw_.kr), ydev(gw_.ydev), ybuf(gw_.ybuf), red(gw_.red) {
        gw.ensure((size_t)n, restart + 1, nr);
        for (int i = 0; i < nr; ++i) {
            Rhs& q = R[i];
            q.V.p = gw.V[i].p;
            zd* v = gw.vec[i].p;
            q.w.p = v;
            q.x.p = v + (size_t)n;
            q.xt.p = v + 2 * (size_t)n;
            q.r.p = v + 3 * (size_t)n;
            q.t.p = v + 4 * (size_t)n;
            q.u.p = v + 5 * (size_t)n;
            q.H.assign((size_t)(restart + 1) * restart, 0);
            q.sn.assign(restart, 0);
            q.g.assign(restart + 1, 0);
            q.cs.assign(restart, 0);
        }
    }

    zd* row(Rhs& q, int j) { return q.V.p + (size_t)j * n; }

    void fetch(int cnt) {
        GT_HIP(hipMemcpyAsync(red.p, kr.red.p, sizeof(double) * kr.P * cnt, hipMemcpyDeviceToHost, s));
        spin_sync(s);
    }
    double val(int slot_in_batch, int c) const { return red[(size_t)slot_in_batch * kr.P + c]; }

    // out_q = K_q M_q^-1 v_q for the listed right-hand sides (t_q the intermediate)
    void apply(const std::vector<int>& who, const std::vector<const zd*>& in, const std::vector<zd*>& out) {
        const int c = (int)who.size();
        std::vector<gt_op*> ops(c);
        std::vector<zd*> tmp(c);
        for (int i = 0; i < c; ++i) {
            ops[i] = R[who[i]].op;
            tmp[i] = R[who[i]].t.p;
            R[who[i]].products++;
        }
        ops_precond(st, c, ops.data(), in.data(), tmp.data());
        std::vector<const zd*> tin(tmp.begin(), tmp.end());
        ops_matvec(st, c, ops.data(), tin.data(), out.data());
    }

    void norms(const std::vector<int>& who, const std::vector<const zd*>& v, std::vector<double>& out) {
        KrSet k{};
        k.nr = (int)who.size();
        for (int i = 0; i < k.nr; ++i) k.w[i] = const_cast<zd*>(v[i]);
        kr.norm2(k, s);
        fetch(k.nr);
        out.resize(who.size());
        for (int i = 0; i < k.nr; ++i) out[i] = std::sqrt(val(i, kr.slot_norm()));
    }

    // dst_q = from_q + V_q[0:k_q]^T y_q   (from_q copied first when distinct)
    void combine(const std::vector<int>& who, const std::vector<std::vector<cd>>& ys, const std::vector<zd*>& dst,
                 const std::vector<const zd*>& from) {
        VecOps c{};
        c.nr = (int)who.size();
        bool any = false;
        for (int i = 0; i < c.nr; ++i) {
            c.dst[i] = dst[i];
            c.a[i] = from[i];
            if (from[i] != dst[i]) any = true;
        }
        if (any) k_copy(c, n, s);
        spin_sync(s);                        // the previous copy out of ybuf has finished
        for (int i = 0; i < c.nr; ++i) {
            c.a[i] = R[who[i]].V.p;
            c.k[i] = (int)ys[i].size();
            for (size_t r = 0; r < ys[i].size(); ++r) ybuf[(size_t)i * (restart + 1) + r] = ys[i][r];
        }
        GT_HIP(hipMemcpyAsync(ydev.p, ybuf.p, sizeof(zd) * c.nr * (restart + 1), hipMemcpyHostToDevice, s));
        k_combine(c, ydev.p, restart + 1, n, s);
    }

    std::vector<cd> back_substitute(const Rhs& q, int k) const {
        std::vector<cd> y(k);
        const int ld = restart;
        for (int i = k - 1; i >= 0; --i) {
            cd acc = q.g[i];
            for (int c = i + 1; c < k; ++c) acc -= q.H[(size_t)i * ld + c] * y[c];
            y[i] = acc / q.H[(size_t)i * ld + i];
        }
        return y;
    }

    // r_q = b_q - K_q M_q^-1 x_q for the listed right-hand sides
    void residuals(const std::vector<int>& who, const std::vector<zd*>& xs) {
        std::vector<const zd*> in(xs.begin(), xs.end());
        std::vector<zd*> out;
        for (int q : who) out.push_back(R[q].u.p);
        apply(who, in, out);
        VecOps v{};
        v.nr = (int)who.size();
        for (int i = 0; i < v.nr; ++i) {
            v.dst[i] = R[who[i]].r.p;
            v.a[i] = R[who[i]].b;
            v.b[i] = R[who[i]].u.p;
        }
        k_sub(v, n, s);
    }

    void run() {
        const double tol = p.tol;
        const int max_cycles = p.max_cycles, extra = p.extra_cycles;
        const int max_its = p.max_its > 0 ? p.max_its : std::numeric_limits<int>::max();
        const int check_misses = p.check_misses > 0 ? p.check_misses : 4;
        std::vector<int> all(nr);
        for (int i = 0; i < nr; ++i) all[i] = i;
        // x = 0, r = b, |b|
        {
            VecOps v{};
            v.nr = nr;
            for (int i = 0; i < nr; ++i) v.dst[i] = R[i].x.p;
            k_zero(v, n, s);
            for (int i = 0; i < nr; ++i) {
                v.dst[i] = R[i].r.p;
                v.a[i] = R[i].b;
            }
            k_copy(v, n, s);
            std::vector<const zd*> bs;
            for (int i = 0; i < nr; ++i) bs.push_back(R[i].b);
            std::vector<double> b0;
            norms(all, bs, b0);
            for (int i = 0; i < nr; ++i) {
                Rhs& q = R[i];
                q.bnorm = p.bnorm > 0 ? p.bnorm : b0[i];
                q.last = std::numeric_limits<double>::infinity();
                if (q.bnorm == 0 || b0[i] <= tol * q.bnorm) {
                    q.state = DONE;
                    q.conv = true;
                    q.resid = q.bnorm > 0 ? b0[i] / q.bnorm : 0.0;
                }
            }
        }
        for (;;) {
            // ---- cycle starts: |r|, the stopping tests, V[0] = r / |r|
            std::vector<int> S;
            for (int i = 0; i < nr; ++i)
                if (R[i].state == START) S.push_back(i);
            if (!S.empty()) {
                std::vector<const zd*> rs;
                for (int q : S) rs.push_back(R[q].r.p);
                std::vector<double> beta;
                norms(S, rs, beta);
                VecOps v{};
                for (size_t i = 0; i < S.size(); ++i) {
                    Rhs& q = R[S[i]];
                    const double b = beta[i];
                    q.resid = b / q.bnorm;
                    if (b <= tol * q.bnorm) { q.state = DONE; q.conv = true; continue; }
                    if (q.its >= max_its || q.cycle >= max_cycles + extra ||
                        (q.cycle >= max_cycles && b > PROGRESS * q.last)) {
                        q.state = DONE;
                        q.conv = false;
                        continue;
                    }
                    q.last = b;
                    std::fill(q.H.begin(), q.H.end(), cd(0));
                    std::fill(q.cs.begin(), q.cs.end(), 0.0);
                    std::fill(q.sn.begin(), q.sn.end(), cd(0));
                    std::fill(q.g.begin(), q.g.end(), cd(0));
                    q.g[0] = b;
                    q.j = 0;
                    q.k = 0;
                    q.aim = tol * q.bnorm;
                    q.misses = 0;
                    q.cycle++;
                    q.state = ARNOLDI;
                    v.dst[v.nr] = row(q, 0);
                    v.a[v.nr] = q.r.p;
                    v.f[v.nr] = 1.0 / b;
                    v.nr++;
                }
                if (v.nr) k_scale(v, n, s);
            }
            std::vector<int> A;
            for (int i = 0; i < nr; ++i)
                if (R[i].state == ARNOLDI) A.push_back(i);
            if (A.empty()) break;
            // ---- one Arnoldi step each: w = K M^-1 V[j], one CGS pass, |w| before and after
            {
                std::vector<const zd*> in;
                std::vector<zd*> out;
                for (int q : A) {
                    in.push_back(row(R[q], R[q].j));
                    out.push_back(R[q].w.p);
                }
                apply(A, in, out);
            }
            // the fused pass also forms the second pass's coefficients V^H w', used when the re-pass is taken
            int jmax = 0;
            for (int q : A) jmax = std::max(jmax, R[q].j);
            const bool fuse = Krylov::fused(jmax + 1);
            auto project = [&](const std::vector<int>& who) {
                KrSet k{};
                k.nr = (int)who.size();
                for (int i = 0; i < k.nr; ++i) {
                    Rhs& q = R[who[i]];
                    k.V[i] = q.V.p;
                    k.m[i] = q.j + 1;
                    k.w[i] = q.w.p;
                }
                kr.dots(k, s);
                if (fuse) {
                    kr.update_dots(k, s);
                    kr.repass(k, s);
                } else {
                    kr.project_update(k, s);
                }
                fetch(k.nr);
            };
            project(A);
            // CGS, then the DGKS re-pass where |w| fell below 0.7 of its value before the projection (on the fused
            // path the device took the test and the re-pass already; its decision is read back)
            std::vector<int> D;
            for (size_t i = 0; i < A.size(); ++i) {
                Rhs& q = R[A[i]];
                q.h.assign(q.j + 1, 0);
                for (int r = 0; r <= q.j; ++r) q.h[r] = cd(val((int)i, 2 * r), val((int)i, 2 * r + 1));
                const double w_norm = std::sqrt(val((int)i, kr.slot_w()));
                q.hn = std::sqrt(val((int)i, kr.slot_w_after()));
                if (fuse) {
                    if (val((int)i, kr.slot_repass()) != 0.0) {
                        for (int r = 0; r <= q.j; ++r)
                            q.h[r] += cd(val((int)i, kr.slot_h2() + 2 * r), val((int)i, kr.slot_h2() + 2 * r + 1));
                        q.hn = std::sqrt(val((int)i, kr.slot_w_after2()));
                    }
                } else if (q.hn < 0.7 * w_norm) {
                    D.push_back(A[i]);
                }
            }
            if (!D.empty()) {                 // DGKS: a second pass only after cancellation
                project(D);
                for (size_t i = 0; i < D.size(); ++i) {
                    Rhs& q = R[D[i]];
                    for (int r = 0; r <= q.j; ++r) q.h[r] += cd(val((int)i, 2 * r), val((int)i, 2 * r + 1));
                    q.hn = std::sqrt(val((int)i, kr.slot_w_after()));
                }
            }
            // ---- Givens rotations and the decisions
            std::vector<int> C, E, Nx;        // true-residual checks, cycle ends, next basis vector
            for (int qi : A) {
                Rhs& q = R[qi];
                const int j = q.j, ld = restart;
                std::vector<cd> col(j + 2);
                for (int i = 0; i <= j; ++i) col[i] = q.h[i];
                col[j + 1] = q.hn;
                for (int i = 0; i < j; ++i) {
                    const cd a = col[i], c = col[i + 1];
                    col[i] = q.cs[i] * a + q.sn[i] * c;
                    col[i + 1] = -std::conj(q.sn[i]) * a + q.cs[i] * c;
                }
                const cd a = col[j];
                const double rho = std::hypot(std::abs(a), q.hn);
                if (rho == 0.0) {
                    q.cs[j] = 1.0;
                    q.sn[j] = 0.0;
                } else if (std::abs(a) == 0.0) {
                    q.cs[j] = 0.0;
                    q.sn[j] = 1.0;
                    col[j] = q.hn;
                } else {
                    const cd phase = a / std::abs(a);
                    q.cs[j] = std::abs(a) / rho;
                    q.sn[j] = phase * q.hn / rho;
                    col[j] = phase * rho;
                }
                col[j + 1] = 0.0;
                for (int i = 0; i <= j + 1; ++i) q.H[(size_t)i * ld + j] = col[i];
                q.g[j + 1] = -std::conj(q.sn[j]) * q.g[j];
                q.g[j] = q.cs[j] * q.g[j];
                q.its++;
                q.k = j + 1;
                double hnorm = 0;
                for (int i = 0; i <= j; ++i) hnorm += std::norm(q.h[i]);
                hnorm = std::sqrt(hnorm);
                if (q.hn <= 1e-14 * hnorm || q.its >= max_its) E.push_back(qi);
                else if (std::abs(q.g[j + 1]) <= q.aim) C.push_back(qi);
                else Nx.push_back(qi);
            }
            // ---- true-residual checks: xt = x + V y, r = b - K M^-1 xt
            if (!C.empty()) {
                std::vector<std::vector<cd>> ys;
                std::vector<zd*> dst;
                std::vector<const zd*> from;
                for (int qi : C) {
                    ys.push_back(back_substitute(R[qi], R[qi].k));
                    dst.push_back(R[qi].xt.p);
                    from.push_back(R[qi].x.p);
                }
                combine(C, ys, dst, from);
                residuals(C, dst);
                std::vector<const zd*> rs;
                for (int qi : C) rs.push_back(R[qi].r.p);
                std::vector<double> rt;
                norms(C, rs, rt);
                VecOps cp{};
                for (size_t i = 0; i < C.size(); ++i) {
                    Rhs& q = R[C[i]];
                    q.resid = rt[i] / q.bnorm;
                    if (rt[i] <= tol * q.bnorm) {
                        cp.dst[cp.nr] = q.x.p;
                        cp.a[cp.nr++] = q.xt.p;
                        q.state = DONE;
                        q.conv = true;
                        continue;
                    }
                    q.misses++;
                    if (q.misses >= check_misses) {     // restart from the check: x = xt, r = rt
                        cp.dst[cp.nr] = q.x.p;
                        cp.a[cp.nr++] = q.xt.p;
                        q.state = START;
                        continue;
                    }
                    q.aim = 0.5 * std::abs(q.g[q.j + 1]);
                    Nx.push_back(C[i]);
                }
                if (cp.nr) k_copy(cp, n, s);
            }
            // ---- the next basis vector, V[j + 1] = w / hn; a full cycle ends
            if (!Nx.empty()) {
                VecOps v{};
                for (int qi : Nx) {
                    Rhs& q = R[qi];
                    v.dst[v.nr] = row(q, q.j + 1);
                    v.a[v.nr] = q.w.p;
                    v.f[v.nr++] = 1.0 / q.hn;
                }
                k_scale(v, n, s);
                for (int qi : Nx) {
                    Rhs& q = R[qi];
                    if (q.j + 1 >= restart) E.push_back(qi);
                    else q.j++;
                }
            }
            // ---- cycle ends: x += V y (least squares by back substitution), r = b - K M^-1 x
            if (!E.empty()) {
                std::vector<std::vector<cd>> ys;
                std::vector<zd*> dst;
                std::vector<const zd*> from;
                for (int qi : E) {
                    Rhs& q = R[qi];
                    int k = q.k;
                    while (k > 0 && q.H[(size_t)(k - 1) * restart + (k - 1)] == cd(0)) --k;
                    ys.push_back(k ? back_substitute(q, k) : std::vector<cd>());
                    dst.push_back(q.x.p);
                    from.push_back(q.x.p);
                    q.state = START;
                }
                combine(E, ys, dst, from);
                residuals(E, dst);
            }
        }
    }
};

}  // namespace

int default_restart(int n) { return (int)std::min(200.0, std::max(100.0, std::floor(4e8 / (16.0 * n)))); }

void gmres_batch(gt_stack* st, int nr, gt_op* const* ops, const zd* const* b, zd* const* x, const gt_gmres_params& p,
                 gt_gmres_result* res) {
    if (nr < 1 || nr > KMAX) fail(GT_ERR_ARG, "1 to 16 right-hand sides per batch");
    const int n = st->nb + st->nk;
    if (!st->ctx->gmres) st->ctx->gmres.reset(new GmresWork);
    Batch bt(st, nr, p.restart > 0 ? p.restart : default_restart(n), p, *st->ctx->gmres);
    for (int i = 0; i < nr; ++i) {
        if (ops[i]->st != st) fail(GT_ERR_ARG, "a batch's operators must share one stack");
        bt.R[i].op = ops[i];
        bt.R[i].b = b[i];
    }
    bt.run();
    // x = M^-1 y (zero when y is)
    std::vector<const zd*> ys;
    for (int i = 0; i < nr; ++i) ys.push_back(bt.R[i].x.p);
    ops_precond(st, nr, ops, ys.data(), x);
    GT_HIP(hipStreamSynchronize(bt.s));
    for (int i = 0; i < nr; ++i) {
        res[i].iterations = bt.R[i].its;
        res[i].converged = bt.R[i].conv ? 1 : 0;
        res[i].products = bt.R[i].products;
        res[i].residual = bt.R[i].resid;
    }
}

}  // namespace gt

gt_ctx_s::gt_ctx_s() = default;
gt_ctx_s::~gt_ctx_s() {
    for (hipEvent_t e : lp_events) (void)hipEventDestroy(e);
}
