// Restarted, right-preconditioned GMRES on the device for a batch of right-hand sides on one stack, each with its
// own operator (frequency) and right-hand side (peec.py:3093-3224 `_gmres`, called as peec.py:2581 does:
// (K M^-1) y = b, x = M^-1 y).
//
// Every right-hand side follows the reference's control flow (Givens rotations, the true-residual check at the aim
// with its miss budget, restarts, extra cycles while each falls to 0.7), as its own state machine on the host. Its
// orthogonalisation is the reference's CGS with a DGKS re-pass below 0.7, or (the context's default) delayed CGS2:
// the same second pass, always taken and one step late, fused with the next step's projection. The machines advance in lockstep: each phase (cycle starts, Arnoldi
// steps, re-passes, true-residual checks, cycle ends) is one batched launch sequence over the right-hand sides
// that are in it, so the kernel spectra, the incidence and the Schur panels are read once per batch. A right-hand
// side's arithmetic does not depend on what else is in its batch: its result is bit-identical alone and in any mix.
// The bases and vectors stay in HBM; per phase one copy brings the Hessenberg columns and norms to the host.
#include <cmath>
#include <complex>
#include <limits>
#include <memory>
#include <type_traits>
#include <vector>

#include "krylov.hpp"
#include "stack.hpp"

namespace gt {

namespace {

using cd = std::complex<double>;

constexpr double PROGRESS = 0.7;

// FREE: no work in the slot; PENDING: submitted, its set-up (x = 0, r = b, |b|) not yet launched; START: at a cycle
// start; ARNOLDI: inside a cycle; DONE: stopped, x = M^-1 y not yet formed; READY: finished, results held.
enum State { FREE, PENDING, START, ARNOLDI, DONE, READY };

// A device pointer in the role of the buffers it replaces (`.p`)
template <class T>
struct Vec {
    cpx<T>* p = nullptr;
};

template <class TV, class TW>
struct Rhs {
    gt_op* op = nullptr;
    const cpx<TW>* b = nullptr;            // right-hand side (device)
    gt_gmres_params p{};                   // its tolerance and budgets
    Vec<TV> V;                             // basis (restart + 1 rows)
    Vec<TW> w, x, xt, r, t, u;             // vectors, in the workspace
    State state = FREE;
    bool conv = false;
    double bnorm = 0, last = 0, aim = 0, resid = 0, hn = 0;
    int cycle = 0, its = 0, j = 0, k = 0, misses = 0, products = 0;
    std::vector<cd> H, sn, g, h;
    std::vector<double> cs;
    // delayed CGS2: the unrotated Hessenberg columns, the last column's first-pass coefficients and the scale of its
    // subdiagonal, and this step's t, p, u^H v and alpha
    std::vector<cd> Hraw, sprev, dt, dp;
    cd duv = 0;
    double alpha_prev = 1, dalpha = 1;
};

}  // namespace

// The device (and pinned host) memory of a GMRES batch. The context keeps it between calls, grown to the largest
// vector length, basis and batch it has served: a solve allocates nothing once its size has been seen.
struct GmresWork {
    size_t n = 0;                          // vector length the slots hold
    int rows = 0;                          // basis rows the slots hold
    int slots = 0;                         // right-hand-side slots allocated at that size
    int vbytes = 0;                        // bytes per stored basis value the slots hold (16 or 8)
    DevBuf<zd> V[KMAX], vec[KMAX];         // per slot: the basis, and six vectors of n
    Krylov kr;
    int kr_rows = 0;
    size_t kr_n = 0;
    DevBuf<zd> ydev;                       // combination coefficients, rows per right-hand side
    HostBuf<cd> ybuf;
    HostBuf<double> red;
    DevBuf<zd> ref[KMAX];                  // mixed precision: the FP64 refinement vectors of each slot (4 n)
    size_t ref_n = 0;
    int ref_slots = 0;

    void ensure_ref(size_t n_, int nr) {
        if (n_ > ref_n) {
            for (int i = 0; i < ref_slots; ++i) ref[i].release();
            ref_slots = 0;
            ref_n = n_;
        }
        for (; ref_slots < nr; ++ref_slots) ref[ref_slots].alloc(4 * ref_n);
    }

    // complex doubles of one slot's basis: rows of the padded stride, vbytes per stored value (16 or 8)
    static size_t basis_count(size_t n_, int rows_, int vbytes_) {
        return (size_t)rows_ * basis_ld(n_) * (size_t)vbytes_ / sizeof(zd);
    }
    // device bytes of a workspace of `slots` slots of `rows` basis rows over vectors of n
    static int64_t plan_bytes(size_t n_, int rows_, int slots_, int vbytes_ = sizeof(zd)) {
        const int64_t z = sizeof(zd);
        const int64_t P = 4 * (int64_t)rows_ + 8;
        const int64_t nblk = Krylov::part_blocks((int)n_);
        return (int64_t)slots_ * ((int64_t)basis_count(n_, rows_, vbytes_) + 6 * (int64_t)n_) * z +
               (int64_t)KMAX * rows_ * z +
               (int64_t)KMAX * nblk * P * 8 + (int64_t)KMAX * P * 8;
    }
    int64_t device_bytes() const {
        int64_t b = ydev.bytes() + kr.part.bytes() + kr.red.bytes();
        for (int i = 0; i < KMAX; ++i) b += V[i].bytes() + vec[i].bytes() + ref[i].bytes();
        return b;
    }

    void ensure(size_t n_, int rows_, int nr, int vbytes_ = sizeof(zd)) {
        if (n_ > n || rows_ > rows || vbytes_ > vbytes) {
            for (int i = 0; i < slots; ++i) {
                V[i].release();
                vec[i].release();
            }
            slots = 0;
            n = std::max(n, n_);
            rows = std::max(rows, rows_);
            vbytes = std::max(vbytes, vbytes_);
            ydev.alloc((size_t)KMAX * rows);
            ybuf.alloc((size_t)KMAX * rows);
        }
        for (; slots < nr; ++slots) {
            V[slots].alloc(basis_count(n, rows, vbytes));
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

// TV: the storage of the basis, TW: of the vectors (double; float and float for the mixed-precision inner solve;
// float and double for the solve whose basis alone is single precision). The Hessenberg matrix, the rotations and
// every reduction are double.
template <class TV, class TW>
struct Batch {
    using cz = cpx<TW>;
    using cv = cpx<TV>;
    using RhsT = Rhs<TV, TW>;
    gt_stack* st;
    hipStream_t s;
    int n, restart, cap;
    size_t ld;                             // basis row stride
    std::vector<RhsT> R;
    GmresWork& gw;
    bool delayed;                          // delayed CGS2 (else CGS with the reference's DGKS re-pass)
    Krylov& kr;
    DevBuf<zd>& ydev;
    HostBuf<cd>& ybuf;
    HostBuf<double>& red;

    // cap slots over the workspace gw, every one FREE, for vectors of up to n_max (0: the host stack's): the slots may
    // hold solves of what-if stacks derived from st, each over its own length
    Batch(gt_stack* st_, int cap_, int restart_, GmresWork& gw_, int n_max = 0)
        : st(st_), s(st_->ctx->stream), n(std::max(n_max, st_->nb + st_->nk)), restart(restart_), cap(cap_),
          ld(basis_ld(n)), R(cap_), gw(gw_), delayed(st_->ctx->gs_delayed),
          kr(gw_.kr), ydev(gw_.ydev), ybuf(gw_.ybuf), red(gw_.red) {
        gw.ensure((size_t)n, restart + 1, cap, (int)sizeof(cv));
        for (int i = 0; i < cap; ++i) {
            RhsT& q = R[i];
            q.V.p = reinterpret_cast<cv*>(gw.V[i].p);
            cz* v = reinterpret_cast<cz*>(gw.vec[i].p);
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

    // A solve enters slot i (FREE or READY): its set-up runs at the next init().
    void submit(int i, gt_op* op, const cz* b, const gt_gmres_params& p) {
        RhsT& q = R[i];
        q.op = op;
        q.b = b;
        q.p = p;
        q.state = PENDING;
        q.conv = false;
        q.bnorm = q.last = q.aim = q.resid = q.hn = 0;
        q.cycle = q.its = q.j = q.k = q.misses = q.products = 0;
    }

    bool running() const {
        for (const RhsT& q : R)
            if (q.state == PENDING || q.state == START || q.state == ARNOLDI) return true;
        return false;
    }

    cv* row(RhsT& q, int j) { return q.V.p + (size_t)j * ld; }
    // a slot's vector length: its operator's stack's unknowns
    int len(int slot) const { return R[slot].op->st->nb + R[slot].op->st->nk; }

    void fetch(int cnt) {
        GT_HIP(hipMemcpyAsync(red.p, kr.red.p, sizeof(double) * kr.P * cnt, hipMemcpyDeviceToHost, s));
        spin_sync(s);
    }
    double val(int slot_in_batch, int c) const { return red[(size_t)slot_in_batch * kr.P + c]; }

    // out_q = K_q M_q^-1 v_q for the listed right-hand sides (t_q the intermediate)
    void apply(const std::vector<int>& who, const std::vector<const cz*>& in, const std::vector<cz*>& out) {
        const int c = (int)who.size();
        std::vector<gt_op*> ops(c);
        std::vector<cz*> tmp(c);
        for (int i = 0; i < c; ++i) {
            ops[i] = R[who[i]].op;
            tmp[i] = R[who[i]].t.p;
            R[who[i]].products++;
        }
        ops_precond<TW>(st, c, ops.data(), in.data(), tmp.data());
        std::vector<const cz*> tin(tmp.begin(), tmp.end());
        ops_matvec<TW>(st, c, ops.data(), tin.data(), out.data());
    }

    void norms(const std::vector<int>& who, const std::vector<const cz*>& v, std::vector<double>& out) {
        KrSetT<TV, TW> k{};
        k.nr = (int)who.size();
        for (int i = 0; i < k.nr; ++i) {
            k.w[i] = const_cast<cz*>(v[i]);
            k.n[i] = len(who[i]);
        }
        kr.norm2(k, s);
        fetch(k.nr);
        out.resize(who.size());
        for (int i = 0; i < k.nr; ++i) out[i] = std::sqrt(val(i, kr.slot_norm()));
    }

    // dst_q = from_q + V_q[0:k_q]^T y_q   (from_q copied first when distinct)
    void combine(const std::vector<int>& who, const std::vector<std::vector<cd>>& ys, const std::vector<cz*>& dst,
                 const std::vector<const cz*>& from) {
        VecOpsT<TW> cp{};
        cp.nr = (int)who.size();
        bool any = false;
        for (int i = 0; i < cp.nr; ++i) {
            cp.dst[i] = dst[i];
            cp.a[i] = from[i];
            cp.n[i] = len(who[i]);
            if (from[i] != dst[i]) any = true;
        }
        if (any) k_copy(cp, n, s);
        VecOpsT<TW, TV> c{};
        c.nr = cp.nr;
        for (int i = 0; i < c.nr; ++i) {
            c.dst[i] = dst[i];
            c.n[i] = len(who[i]);
        }
        spin_sync(s);                        // the previous copy out of ybuf has finished
        for (int i = 0; i < c.nr; ++i) {
            c.a[i] = R[who[i]].V.p;
            c.k[i] = (int)ys[i].size();
            for (size_t r = 0; r < ys[i].size(); ++r) ybuf[(size_t)i * (restart + 1) + r] = ys[i][r];
        }
        GT_HIP(hipMemcpyAsync(ydev.p, ybuf.p, sizeof(zd) * c.nr * (restart + 1), hipMemcpyHostToDevice, s));
        k_combine(c, ydev.p, restart + 1, n, (int)ld, s);
    }

    std::vector<cd> back_substitute(const RhsT& q, int k) const {
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
    void residuals(const std::vector<int>& who, const std::vector<cz*>& xs) {
        std::vector<const cz*> in(xs.begin(), xs.end());
        std::vector<cz*> out;
        for (int q : who) out.push_back(R[q].u.p);
        apply(who, in, out);
        VecOpsT<TW> v{};
        v.nr = (int)who.size();
        for (int i = 0; i < v.nr; ++i) {
            v.dst[i] = R[who[i]].r.p;
            v.a[i] = R[who[i]].b;
            v.b[i] = R[who[i]].u.p;
            v.n[i] = len(who[i]);
        }
        k_sub(v, n, s);
    }

    static int max_its_of(const RhsT& q) {
        return q.p.max_its > 0 ? q.p.max_its : std::numeric_limits<int>::max();
    }

    // The PENDING slots' set-up: x = 0, r = b, |b|; a right-hand side already within its tolerance is DONE.
    void init() {
        std::vector<int> P;
        for (int i = 0; i < cap; ++i)
            if (R[i].state == PENDING) P.push_back(i);
        if (P.empty()) return;
        VecOpsT<TW> v{};
        v.nr = (int)P.size();
        for (int i = 0; i < v.nr; ++i) {
            v.dst[i] = R[P[i]].x.p;
            v.n[i] = len(P[i]);
        }
        k_zero(v, n, s);
        for (int i = 0; i < v.nr; ++i) {
            v.dst[i] = R[P[i]].r.p;
            v.a[i] = R[P[i]].b;
        }
        k_copy(v, n, s);
        std::vector<const cz*> bs;
        for (int q : P) bs.push_back(R[q].b);
        std::vector<double> b0;
        norms(P, bs, b0);
        for (size_t i = 0; i < P.size(); ++i) {
            RhsT& q = R[P[i]];
            q.bnorm = q.p.bnorm > 0 ? q.p.bnorm : b0[i];
            q.last = std::numeric_limits<double>::infinity();
            q.state = START;
            if (q.bnorm == 0 || b0[i] <= q.p.tol * q.bnorm) {
                q.state = DONE;
                q.conv = true;
                q.resid = q.bnorm > 0 ? b0[i] / q.bnorm : 0.0;
            }
        }
    }

    // Gram-Schmidt for the slots `who` (all on the fused pass or all not), the values fetched to the host
    void project(const std::vector<int>& who, bool fuse) {
        KrSetT<TV, TW> k{};
        k.nr = (int)who.size();
        k.ld = (int)ld;
        for (int i = 0; i < k.nr; ++i) {
            RhsT& q = R[who[i]];
            k.V[i] = q.V.p;
            k.m[i] = q.j + 1;
            k.w[i] = q.w.p;
            k.n[i] = len(who[i]);
        }
        kr.dots(k, s);
        if (fuse) {
            kr.update_dots(k, s);
            kr.repass(k, s);
        } else {
            kr.project_update(k, s);
        }
        fetch(k.nr);
    }

    // ---- cycle starts: |r|, the stopping tests, V[0] = r / |r|
    void start_cycles() {
        std::vector<int> S;
        for (int i = 0; i < cap; ++i)
            if (R[i].state == START) S.push_back(i);
        if (S.empty()) return;
        std::vector<const cz*> rs;
        for (int q : S) rs.push_back(R[q].r.p);
        std::vector<double> beta;
        norms(S, rs, beta);
        VecOpsT<TV, TW> v{};
        for (size_t i = 0; i < S.size(); ++i) {
            RhsT& q = R[S[i]];
            const double b = beta[i];
            q.resid = b / q.bnorm;
            if (b <= q.p.tol * q.bnorm) { q.state = DONE; q.conv = true; continue; }
            if (q.its >= max_its_of(q) || q.cycle >= q.p.max_cycles + q.p.extra_cycles ||
                (q.cycle >= q.p.max_cycles && b > PROGRESS * q.last)) {
                q.state = DONE;
                q.conv = false;
                continue;
            }
            q.last = b;
            std::fill(q.H.begin(), q.H.end(), cd(0));
            std::fill(q.cs.begin(), q.cs.end(), 0.0);
            std::fill(q.sn.begin(), q.sn.end(), cd(0));
            std::fill(q.g.begin(), q.g.end(), cd(0));
            if (delayed) q.Hraw.assign((size_t)(restart + 1) * restart, cd(0));
            q.g[0] = b;
            q.j = 0;
            q.k = 0;
            q.aim = q.p.tol * q.bnorm;
            q.misses = 0;
            q.cycle++;
            q.state = ARNOLDI;
            v.dst[v.nr] = row(q, 0);
            v.a[v.nr] = q.r.p;
            v.f[v.nr] = 1.0 / b;
            v.n[v.nr] = len(S[i]);
            v.nr++;
        }
        if (v.nr) k_scale(v, n, s);
    }

    // w = K M^-1 V[j] for the slots listed (a single-precision basis row widened into u for the product)
    void product(const std::vector<int>& A) {
        std::vector<const cz*> in;
        std::vector<cz*> out;
        if constexpr (std::is_same<TV, TW>::value) {
            for (int q : A) in.push_back(row(R[q], R[q].j));
        } else {
            VecOpsT<TW, TV> cv{};
            for (int q : A) {
                cv.dst[cv.nr] = R[q].u.p;
                cv.a[cv.nr] = row(R[q], R[q].j);
                cv.n[cv.nr] = len(q);
                cv.f[cv.nr++] = 1.0;
                in.push_back(R[q].u.p);
            }
            k_scale(cv, n, s);
        }
        for (int q : A) out.push_back(R[q].w.p);
        apply(A, in, out);
    }

    // Column jc of the Hessenberg matrix (q.h, jc + 1 values, and q.hn below them): the previous rotations, a new
    // one, the residual estimate, and the decision: cycle end (E), true-residual check (C) or go on (Nx).
    void rotate(int qi, int jc, std::vector<int>& C, std::vector<int>& E, std::vector<int>& Nx) {
        RhsT& q = R[qi];
        const int ld = restart;
        std::vector<cd> col(jc + 2);
        for (int i = 0; i <= jc; ++i) col[i] = q.h[i];
        col[jc + 1] = q.hn;
        for (int i = 0; i < jc; ++i) {
            const cd a = col[i], c = col[i + 1];
            col[i] = q.cs[i] * a + q.sn[i] * c;
            col[i + 1] = -std::conj(q.sn[i]) * a + q.cs[i] * c;
        }
        const cd a = col[jc];
        const double rho = std::hypot(std::abs(a), q.hn);
        if (rho == 0.0) {
            q.cs[jc] = 1.0;
            q.sn[jc] = 0.0;
        } else if (std::abs(a) == 0.0) {
            q.cs[jc] = 0.0;
            q.sn[jc] = 1.0;
            col[jc] = q.hn;
        } else {
            const cd phase = a / std::abs(a);
            q.cs[jc] = std::abs(a) / rho;
            q.sn[jc] = phase * q.hn / rho;
            col[jc] = phase * rho;
        }
        col[jc + 1] = 0.0;
        for (int i = 0; i <= jc + 1; ++i) q.H[(size_t)i * ld + jc] = col[i];
        q.g[jc + 1] = -std::conj(q.sn[jc]) * q.g[jc];
        q.g[jc] = q.cs[jc] * q.g[jc];
        q.its++;
        q.k = jc + 1;
        double hnorm = 0;
        for (int i = 0; i <= jc; ++i) hnorm += std::norm(q.h[i]);
        hnorm = std::sqrt(hnorm);
        if (q.hn <= 1e-14 * hnorm || q.its >= max_its_of(q)) E.push_back(qi);
        else if (std::abs(q.g[jc + 1]) <= q.aim) C.push_back(qi);
        else Nx.push_back(qi);
    }

    // ---- true-residual checks: xt = x + V y, r = b - K M^-1 xt; those that go on join Nx
    void checks(const std::vector<int>& C, std::vector<int>& Nx) {
        if (C.empty()) return;
        std::vector<std::vector<cd>> ys;
        std::vector<cz*> dst;
        std::vector<const cz*> from;
        for (int qi : C) {
            ys.push_back(back_substitute(R[qi], R[qi].k));
            dst.push_back(R[qi].xt.p);
            from.push_back(R[qi].x.p);
        }
        combine(C, ys, dst, from);
        residuals(C, dst);
        std::vector<const cz*> rs;
        for (int qi : C) rs.push_back(R[qi].r.p);
        std::vector<double> rt;
        norms(C, rs, rt);
        VecOpsT<TW> cp{};
        for (size_t i = 0; i < C.size(); ++i) {
            RhsT& q = R[C[i]];
            q.resid = rt[i] / q.bnorm;
            if (rt[i] <= q.p.tol * q.bnorm) {
                cp.dst[cp.nr] = q.x.p;
                cp.n[cp.nr] = len(C[i]);
                cp.a[cp.nr++] = q.xt.p;
                q.state = DONE;
                q.conv = true;
                continue;
            }
            q.misses++;
            if (q.misses >= (q.p.check_misses > 0 ? q.p.check_misses : 4)) {     // restart from the check: x = xt, r = rt
                cp.dst[cp.nr] = q.x.p;
                cp.n[cp.nr] = len(C[i]);
                cp.a[cp.nr++] = q.xt.p;
                q.state = START;
                continue;
            }
            q.aim = 0.5 * std::abs(q.g[q.k]);
            Nx.push_back(C[i]);
        }
        if (cp.nr) k_copy(cp, n, s);
    }

    // ---- cycle ends: x += V y (least squares by back substitution), r = b - K M^-1 x
    void cycle_ends(const std::vector<int>& E) {
        if (E.empty()) return;
        std::vector<std::vector<cd>> ys;
        std::vector<cz*> dst;
        std::vector<const cz*> from;
        for (int qi : E) {
            RhsT& q = R[qi];
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

    // One Arnoldi step of every slot inside a cycle, after the cycle starts; false when no slot is left in one.
    bool step() {
        start_cycles();
        std::vector<int> A;
        for (int i = 0; i < cap; ++i)
            if (R[i].state == ARNOLDI) A.push_back(i);
        if (A.empty()) return false;
        if (delayed) {
            step_delayed(A);
            return true;
        }
        // ---- one Arnoldi step each: w = K M^-1 V[j], one CGS pass, |w| before and after
        product(A);
        // CGS with the DGKS re-pass where |w| fell below 0.7 of its value before the projection. Up to 128 basis rows
        // the update pass also forms the re-pass's coefficients V^H w' and the device takes the test and the re-pass
        // (its decision read back); longer bases test on the host. The path depends on the slot's own j only, so a
        // right-hand side's arithmetic does not depend on the rest of the batch.
        std::vector<int> D;
        for (const bool fuse : {true, false}) {
            std::vector<int> G;
            for (int q : A)
                if (Krylov::fused(R[q].j + 1) == fuse) G.push_back(q);
            if (G.empty()) continue;
            project(G, fuse);
            for (size_t i = 0; i < G.size(); ++i) {
                RhsT& q = R[G[i]];
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
                    D.push_back(G[i]);
                }
            }
        }
        if (!D.empty()) {                     // DGKS on the host's test: the second pass
            project(D, false);
            for (size_t i = 0; i < D.size(); ++i) {
                RhsT& q = R[D[i]];
                for (int r = 0; r <= q.j; ++r) q.h[r] += cd(val((int)i, 2 * r), val((int)i, 2 * r + 1));
                q.hn = std::sqrt(val((int)i, kr.slot_w_after()));
            }
        }
        std::vector<int> C, E, Nx;        // true-residual checks, cycle ends, next basis vector
        for (int qi : A) rotate(qi, R[qi].j, C, E, Nx);
        checks(C, Nx);
        // ---- the next basis vector, V[j + 1] = w / hn; a full cycle ends
        if (!Nx.empty()) {
            VecOpsT<TV, TW> v{};
            for (int qi : Nx) {
                RhsT& q = R[qi];
                v.dst[v.nr] = row(q, q.j + 1);
                v.a[v.nr] = q.w.p;
                v.n[v.nr] = len(qi);
                v.f[v.nr++] = 1.0 / q.hn;
            }
            k_scale(v, n, s);
            for (int qi : Nx) {
                RhsT& q = R[qi];
                if (q.j + 1 >= restart) E.push_back(qi);
                else q.j++;
            }
        }
        cycle_ends(E);
        return true;
    }

    // The delayed CGS2 step (classical Gram-Schmidt with its second pass one step late, as low-synchronisation
    // GMRES does it): row j holds u_j, the once-projected and normalised vector, rows 0..j-1 are final. v = A u_j;
    // pass A: t = Q^H u_j, p = Q^H v, u^H u, u^H v; alpha = |u_j - Q t| = sqrt(u^H u - |t|^2); column j - 1 of the
    // Hessenberg matrix is then final: s_{j-1} + beta t above, beta alpha below (beta = |y_j|, y_j the vector u_j was
    // normalised from); pass B: q_j = (u_j - Q t) / alpha over row j and y_{j+1} = v - Q p - q_j (q_j^H v), whose norm
    // scales row j + 1 on the device; s_j = ([p ; q_j^H v] - H[0:j+1, 0:j] t) / alpha. The basis is read twice per
    // step (three times with the re-pass). Every slot keeps its own row in the reductions (the set is indexed by
    // slot), so |y|^2 stays on the device from one step to the next.
    void step_delayed(const std::vector<int>& A) {
        std::vector<int> Am;
        for (int qi : A)
            if (R[qi].j < restart) Am.push_back(qi);
        if (!Am.empty()) product(Am);
        KrSetT<TV, TW> k{};
        k.nr = cap;
        k.ld = (int)ld;
        for (int i = 0; i < cap; ++i) k.m[i] = -1;
        for (int qi : A) {
            RhsT& q = R[qi];
            k.V[qi] = q.V.p;
            k.m[qi] = q.j;
            k.w[qi] = q.j < restart ? q.w.p : nullptr;
            k.n[qi] = len(qi);
        }
        kr.dots2(k, s);
        fetch(cap);
        std::vector<int> C, E, Nx;
        for (int qi : A) {
            RhsT& q = R[qi];
            const int j = q.j;
            q.dt.resize(j);
            q.dp.resize(j);
            double tt = 0;
            for (int r = 0; r < j; ++r) {
                q.dt[r] = cd(val(qi, kr.slot_h2() + 2 * r), val(qi, kr.slot_h2() + 2 * r + 1));
                q.dp[r] = cd(val(qi, 2 * r), val(qi, 2 * r + 1));
                tt += std::norm(q.dt[r]);
            }
            const double uu = val(qi, kr.slot_uu());
            q.duv = cd(val(qi, kr.slot_uu() + 1), val(qi, kr.slot_uu() + 2));
            q.dalpha = std::sqrt(std::max(uu - tt, 0.0));
            if (j == 0) {
                q.g[0] *= q.dalpha;                 // r = |r| u_0 = |r| alpha q_0
                Nx.push_back(qi);
                continue;
            }
            const double beta = std::sqrt(val(qi, kr.slot_w_after())) / q.alpha_prev;
            q.h.assign(j, 0);
            for (int r = 0; r < j; ++r) q.h[r] = q.sprev[r] + beta * q.dt[r];
            q.hn = beta * q.dalpha;
            for (int r = 0; r < j; ++r) q.Hraw[(size_t)r * restart + (j - 1)] = q.h[r];
            q.Hraw[(size_t)j * restart + (j - 1)] = q.hn;
            rotate(qi, j - 1, C, E, Nx);
        }
        checks(C, Nx);
        // pass B for the slots going on with rows left; a full basis ends its cycle
        std::vector<int> B;
        for (int qi : Nx) {
            if (R[qi].j < restart && R[qi].dalpha > 0) B.push_back(qi);
            else E.push_back(qi);
        }
        if (!B.empty()) {
            KrSetT<TV, TW> kb{};
            kb.nr = cap;
            kb.ld = (int)ld;
            std::vector<double> ainv(cap, 0.0);
            std::vector<zd> qv(cap);
            VecOpsT<TV, TW> v{};
            v.nr = cap;
            for (int i = 0; i < cap; ++i) kb.m[i] = -1;
            for (int qi : B) {
                RhsT& q = R[qi];
                const int j = q.j;
                cd tp = 0;
                for (int r = 0; r < j; ++r) tp += std::conj(q.dt[r]) * q.dp[r];
                const cd qvc = (q.duv - tp) / q.dalpha;
                ainv[qi] = 1.0 / q.dalpha;
                qv[qi] = zd(qvc.real(), qvc.imag());
                // s_j = ([p ; q^H v] - H[0:j+1, 0:j] t) / alpha, with the unrotated columns
                q.sprev.assign(j + 1, cd(0));
                for (int r = 0; r < j; ++r) q.sprev[r] = q.dp[r];
                q.sprev[j] = qvc;
                for (int c = 0; c < j; ++c) {
                    const cd tc = q.dt[c];
                    if (tc == cd(0)) continue;
                    for (int r = 0; r <= c + 1 && r <= j; ++r) q.sprev[r] -= q.Hraw[(size_t)r * restart + c] * tc;
                }
                for (int r = 0; r <= j; ++r) q.sprev[r] /= q.dalpha;
                q.alpha_prev = q.dalpha;
                kb.V[qi] = q.V.p;
                kb.m[qi] = j;
                kb.w[qi] = q.w.p;
                kb.n[qi] = len(qi);
                v.dst[qi] = row(q, j + 1);
                v.a[qi] = q.w.p;
                v.n[qi] = len(qi);
            }
            kr.update2(kb, ainv.data(), qv.data(), s);
            kr.scale_by_norm(v, s);
            for (int qi : B) R[qi].j++;
        }
        cycle_ends(E);
    }

    void run() {
        init();
        while (step()) {
        }
    }
};

}  // namespace

int default_restart(int n) { return (int)std::min(200.0, std::max(100.0, std::floor(4e8 / (16.0 * n)))); }

namespace {

// TV: the basis storage (double, or float for the single-precision basis); vectors, products and reductions double
template <class TV>
void gmres_batch_t(gt_stack* st, int nr, gt_op* const* ops, const zd* const* b, zd* const* x,
                   const gt_gmres_params& p, gt_gmres_result* res) {
    check_batch(st, nr, ops);
    const int n = st->nb + st->nk;
    int n_max = n;
    for (int i = 0; i < nr; ++i) n_max = std::max(n_max, ops[i]->st->nb + ops[i]->st->nk);
    if (!st->ctx->gmres) st->ctx->gmres.reset(new GmresWork);
    Batch<TV, double> bt(st, nr, p.restart > 0 ? p.restart : default_restart(n), *st->ctx->gmres, n_max);
    for (int i = 0; i < nr; ++i) bt.submit(i, ops[i], b[i], p);
    bt.run();
    // x = M^-1 y (zero when y is)
    std::vector<const zd*> ys;
    for (int i = 0; i < nr; ++i) ys.push_back(bt.R[i].x.p);
    ops_precond<double>(st, nr, ops, ys.data(), x);
    GT_HIP(hipStreamSynchronize(bt.s));
    for (int i = 0; i < nr; ++i) {
        res[i].iterations = bt.R[i].its;
        res[i].converged = bt.R[i].conv ? 1 : 0;
        res[i].products = bt.R[i].products;
        res[i].residual = bt.R[i].resid;
    }
}

}  // namespace

void gmres_batch(gt_stack* st, int nr, gt_op* const* ops, const zd* const* b, zd* const* x, const gt_gmres_params& p,
                 gt_gmres_result* res) {
    gmres_batch_t<double>(st, nr, ops, b, x, p, res);
}

void gmres_batch_fp32_basis(gt_stack* st, int nr, gt_op* const* ops, const zd* const* b, zd* const* x,
                            const gt_gmres_params& p, gt_gmres_result* res) {
    gmres_batch_t<float>(st, nr, ops, b, x, p, res);
}

constexpr double GMRES_PROGRESS_MIXED = 0.7;    // peec.py:99: extra cycles while each falls to this fraction

void gmres_batch_mixed(gt_stack* st, int nr, gt_op* const* ops, const zd* const* b, zd* const* x,
                       const gt_gmres_params& p, double inner_tol, int max_refine, gt_gmres_result* res) {
    if (nr < 1 || nr > KMAX) fail(GT_ERR_ARG, "1 to 16 right-hand sides per batch");
    if (!(inner_tol > 0 && inner_tol < 1)) fail(GT_ERR_ARG, "the inner tolerance must lie in (0, 1)");
    for (int i = 0; i < nr; ++i)
        if (ops[i]->st != st) fail(GT_ERR_ARG, "a batch's operators must share one stack");
    const int n = st->nb + st->nk;
    hipStream_t s = st->ctx->stream;
    if (!st->ctx->gmres) st->ctx->gmres.reset(new GmresWork);
    GmresWork& gw = *st->ctx->gmres;
    const int restart = p.restart > 0 ? p.restart : default_restart(n);
    gw.ensure((size_t)n, restart + 1, nr);
    gw.ensure_ref((size_t)n, nr);
    // per right-hand side, in its FP64 refinement buffer: r (residual), d (correction), u (K x), and r in FP32
    std::vector<zd*> r(nr), d(nr), u(nr);
    std::vector<zf*> r32(nr);
    for (int i = 0; i < nr; ++i) {
        zd* base = gw.ref[i].p;
        r[i] = base;
        d[i] = base + (size_t)n;
        u[i] = base + 2 * (size_t)n;
        r32[i] = reinterpret_cast<zf*>(base + 3 * (size_t)n);
    }
    Krylov& kr = gw.kr;
    auto norms = [&](const std::vector<int>& who, const std::vector<zd*>& v) {
        KrSet k{};
        k.nr = (int)who.size();
        for (int i = 0; i < k.nr; ++i) k.w[i] = v[who[i]];
        kr.norm2(k, s);
        GT_HIP(hipMemcpyAsync(gw.red.p, kr.red.p, sizeof(double) * kr.P * k.nr, hipMemcpyDeviceToHost, s));
        spin_sync(s);
        std::vector<double> out(k.nr);
        for (int i = 0; i < k.nr; ++i) out[i] = std::sqrt(gw.red[(size_t)i * kr.P + kr.slot_norm()]);
        return out;
    };
    auto subset = [&](const std::vector<int>& who, auto&& vec) {
        using E = typename std::decay<decltype(vec[0])>::type;
        std::vector<E> out;
        for (int q : who) out.push_back(vec[q]);
        return out;
    };
    std::vector<int> all(nr);
    for (int i = 0; i < nr; ++i) all[i] = i;
    // x = 0, r = b
    {
        VecOps v{};
        v.nr = nr;
        for (int i = 0; i < nr; ++i) {
            v.dst[i] = x[i];
        }
        k_zero(v, n, s);
        for (int i = 0; i < nr; ++i) {
            v.dst[i] = r[i];
            v.a[i] = b[i];
        }
        k_copy(v, n, s);
    }
    const std::vector<double> b0 = norms(all, r);
    std::vector<double> bnorm(nr), last(nr, std::numeric_limits<double>::infinity());
    std::vector<int> its(nr, 0), prods(nr, 0), conv(nr, 0);
    std::vector<double> resid(nr, 0.0);
    std::vector<int> act;
    for (int i = 0; i < nr; ++i) {
        bnorm[i] = p.bnorm > 0 ? p.bnorm : b0[i];
        resid[i] = bnorm[i] > 0 ? b0[i] / bnorm[i] : 0.0;
        if (bnorm[i] == 0 || b0[i] <= p.tol * bnorm[i]) conv[i] = 1;
        else act.push_back(i);
    }
    last = b0;
    // Each round is one restart cycle of the reference's GMRES (peec.py:3093-3224) run in FP32 on the current FP64
    // residual: (K M^-1) y = r, stopping early only where it meets the outer tolerance (relative to |r|, but never
    // below inner_tol, what an FP32 cycle can resolve); then x += M^-1 y and r = b - K x in FP64. The rounds follow the
    // reference's restart rule: max_cycles of them, then extra ones while each brings |r| to 0.7 of the round before.
    std::vector<int> rounds(nr, 0);
    for (int round = 0; round < max_refine && !act.empty(); ++round) {
        const int na = (int)act.size();
        const std::vector<gt_op*> aop = subset(act, std::vector<gt_op*>(ops, ops + nr));
        k_to_float(na, subset(act, r32).data(), subset(act, r).data(), n, s);
        Batch<float, float> bt(st, na, restart, gw);
        for (int i = 0; i < na; ++i) {
            const int q = act[i];
            gt_gmres_params ip = p;
            ip.tol = std::min(0.5, std::max(inner_tol, p.tol * bnorm[q] / last[q]));
            ip.bnorm = 0;
            ip.restart = restart;
            ip.max_cycles = 1;
            ip.extra_cycles = 0;
            ip.max_its = 0;
            bt.submit(i, aop[i], r32[q], ip);
        }
        bt.run();
        // d = M^-1 y in FP64 (y widened), x += d, u = K x, r = b - u
        {
            VecOps v{};
            v.nr = na;
            for (int i = 0; i < na; ++i) v.dst[i] = u[act[i]];
            k_zero(v, n, s);
            std::vector<const zf*> ys;
            for (int i = 0; i < na; ++i) ys.push_back(bt.R[i].x.p);
            k_add_float(na, subset(act, u).data(), ys.data(), n, s);
            std::vector<const zd*> uin;
            for (int q : act) uin.push_back(u[q]);
            ops_precond<double>(st, na, aop.data(), uin.data(), subset(act, d).data());
            for (int i = 0; i < na; ++i) {
                v.dst[i] = x[act[i]];
                v.a[i] = x[act[i]];
                v.b[i] = d[act[i]];
            }
            // x + d as a - (-d): the correction enters through the sum kernel
            VecOps w{};
            w.nr = na;
            for (int i = 0; i < na; ++i) {
                w.dst[i] = d[act[i]];
                w.a[i] = d[act[i]];
                w.f[i] = -1.0;
            }
            k_scale(w, n, s);
            k_sub(v, n, s);
            std::vector<const zd*> xin;
            for (int q : act) xin.push_back(x[q]);
            ops_matvec<double>(st, na, aop.data(), xin.data(), subset(act, u).data());
            for (int i = 0; i < na; ++i) {
                v.dst[i] = r[act[i]];
                v.a[i] = b[act[i]];
                v.b[i] = u[act[i]];
            }
            k_sub(v, n, s);
        }
        const std::vector<double> rn = norms(act, r);
        std::vector<int> next;
        for (int i = 0; i < na; ++i) {
            const int q = act[i];
            its[q] += bt.R[i].its;
            prods[q] += bt.R[i].products + 2;
            resid[q] = rn[i] / bnorm[q];
            ++rounds[q];
            if (rn[i] <= p.tol * bnorm[q]) {
                conv[q] = 1;
            } else if (rounds[q] >= p.max_cycles &&
                       (rounds[q] >= p.max_cycles + p.extra_cycles || rn[i] > GMRES_PROGRESS_MIXED * last[q])) {
                conv[q] = 0;                   // out of cycles, or no longer falling to 0.7 per cycle
            } else {
                next.push_back(q);
            }
            last[q] = rn[i];
        }
        act = next;
    }
    GT_HIP(hipStreamSynchronize(s));
    for (int i = 0; i < nr; ++i) {
        res[i].iterations = its[i];
        res[i].converged = conv[i];
        res[i].products = prods[i];
        res[i].residual = resid[i];
    }
}

}  // namespace gt

// A serving session: up to 16 slots over one stack and a GMRES workspace of its own, so sessions of several stacks
// interleave on one context. A solve enters a free slot at any time and joins the running batch at its next step;
// every slot runs the same state machine as in a batch, so its result is bit-identical to the same solve alone.
// The slots may hold solves of the stack itself and of what-if stacks derived from it with up to max_extra added
// branches: one batch for a board and its what-ifs. The basis is stored in FP64 or FP32.
namespace gt {
// The context's kept GMRES workspace freed (the next solve makes it again); the caller holds the context's lock.
void ctx_trim_gmres(gt_ctx* ctx) { ctx->gmres.reset(); }
}  // namespace gt

struct gt_session_s {
    gt_stack* st = nullptr;
    int cap = 0, restart = 0, n_max = 0;
    gt::GmresWork gw;
    std::unique_ptr<gt::Batch<double, double>> bt;
    std::unique_ptr<gt::Batch<float, double>> bt32;
    std::vector<gt::zd*> xout;             // the caller's solution vector of each slot
};

namespace gt {

namespace {

template <class F>
auto on_batch(gt_session* ss, F&& f) {
    if (ss->bt32) return f(*ss->bt32);
    return f(*ss->bt);
}

int basis_bytes(int basis) {
    if (basis != GT_BASIS_FP64 && basis != GT_BASIS_FP32) fail(GT_ERR_ARG, "basis is GT_BASIS_FP64 or GT_BASIS_FP32");
    return basis == GT_BASIS_FP32 ? (int)sizeof(zf) : (int)sizeof(zd);
}

}  // namespace

int64_t session_plan_bytes(const gt_stack* st, int cap, int restart, int basis, int max_extra) {
    if (max_extra < 0) fail(GT_ERR_ARG, "negative max_extra");
    const int n = st->nb + max_extra + st->nk;
    const int rows = (restart > 0 ? restart : default_restart(n)) + 1;
    int64_t b = GmresWork::plan_bytes((size_t)n, rows, cap, basis_bytes(basis));
    // the stack's product scratch grows to the session's batch and Woodbury columns, the Schur solve's buffers to
    // the batch
    b += st->work.bytes_after(st, cap, st->k + max_extra) - st->work.bytes();
    if (st->schur) b += st->schur->bytes_after(cap) - st->schur->bytes_after(0);
    return b;
}

gt_session* session_create(gt_stack* st, int cap, int restart, int basis, int max_extra) {
    if (cap < 1 || cap > KMAX) fail(GT_ERR_ARG, "1 to 16 slots per session");
    if (restart < 0) fail(GT_ERR_ARG, "negative restart");
    if (max_extra < 0) fail(GT_ERR_ARG, "negative max_extra");
    const int vb = basis_bytes(basis);
    std::unique_ptr<gt_session_s> ss(new gt_session_s);
    ss->st = st;
    ss->cap = cap;
    ss->n_max = st->nb + max_extra + st->nk;
    ss->restart = restart > 0 ? restart : default_restart(ss->n_max);
    if (vb == (int)sizeof(zf)) ss->bt32.reset(new Batch<float, double>(st, cap, ss->restart, ss->gw, ss->n_max));
    else ss->bt.reset(new Batch<double, double>(st, cap, ss->restart, ss->gw, ss->n_max));
    ss->xout.assign(cap, nullptr);
    st->work.ensure(st, cap, st->k + max_extra);
    if (st->schur) st->schur->ensure(cap);
    ++st->sessions;
    return ss.release();
}

int64_t session_bytes(const gt_session* ss) { return ss->gw.device_bytes(); }

int session_submit(gt_session* ss, gt_op* op, const zd* b, zd* x, const gt_gmres_params& p) {
    check_batch(ss->st, 1, &op);
    if (op->st->nb + op->st->nk > ss->n_max)
        fail(GT_ERR_ARG, "the what-if adds more branches than the session holds");
    if (p.restart > 0 && p.restart != ss->restart) fail(GT_ERR_ARG, "the session's restart differs");
    return on_batch(ss, [&](auto& bt) {
        for (int i = 0; i < ss->cap; ++i) {
            if (bt.R[i].state != FREE) continue;
            bt.submit(i, op, b, p);
            ss->xout[i] = x;
            return i;
        }
        return -1;
    });
}

// Up to max_steps Arnoldi steps of the running slots (the newly submitted set up first); the solves that stopped
// get x = M^-1 y and wait, READY, for their poll. Returns the slots still running.
int session_step(gt_session* ss, int max_steps, int* ready) {
    return on_batch(ss, [&](auto& bt) {
        bt.init();
        for (int k = 0; k < max_steps; ++k)
            if (!bt.step()) break;
        std::vector<int> fin;
        for (int i = 0; i < ss->cap; ++i)
            if (bt.R[i].state == DONE) fin.push_back(i);
        if (!fin.empty()) {
            const int c = (int)fin.size();
            std::vector<gt_op*> ops(c);
            std::vector<const zd*> ys(c);
            std::vector<zd*> xs(c);
            for (int i = 0; i < c; ++i) {
                ops[i] = bt.R[fin[i]].op;
                ys[i] = bt.R[fin[i]].x.p;
                xs[i] = ss->xout[fin[i]];
            }
            ops_precond<double>(ss->st, c, ops.data(), ys.data(), xs.data());
            spin_sync(bt.s);
            for (int i : fin) bt.R[i].state = READY;
        }
        int run = 0, rd = 0;
        for (int i = 0; i < ss->cap; ++i) {
            const State q = bt.R[i].state;
            if (q == READY) rd++;
            else if (q != FREE) run++;
        }
        if (ready) *ready = rd;
        return run;
    });
}

bool session_poll(gt_session* ss, int slot, gt_gmres_result* res) {
    if (slot < 0 || slot >= ss->cap) fail(GT_ERR_ARG, "slot out of range");
    return on_batch(ss, [&](auto& bt) {
        auto& q = bt.R[slot];
        if (q.state != READY) return false;
        res->iterations = q.its;
        res->converged = q.conv ? 1 : 0;
        res->products = q.products;
        res->residual = q.resid;
        q.state = FREE;
        ss->xout[slot] = nullptr;
        return true;
    });
}

// The last session of a stack gives back the scratch the sessions grew (the product workspace and the Schur solve's
// buffers, to one right-hand side): a server that closes an idle session frees what it admitted for it.
void session_destroy(gt_session* ss) {
    gt_stack* st = ss->st;
    delete ss;
    if (--st->sessions == 0) {
        st->work.reset(st);
        if (st->schur) st->schur->shrink(1);
    }
}

gt_stack* session_stack(gt_session* ss) { return ss->st; }

}  // namespace gt

gt_ctx_s::gt_ctx_s() = default;
gt_ctx_s::~gt_ctx_s() {
    for (hipEvent_t e : lp_events) (void)hipEventDestroy(e);
}
