// Restarted, right-preconditioned GMRES on the device (peec.py:3093-3224 `_gmres`, called as peec.py:2581 does:
// (K M^-1) y = b, x = M^-1 y). The basis and all vectors stay in HBM; per iteration the Hessenberg column and the
// norms (|w| before and after the projection) come to the host in one copy, where the Givens rotations, the
// true-residual decision and the restart logic run exactly as the reference's.
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

struct Solver {
    gt_op* op;
    hipStream_t s;
    int n, restart;
    DevBuf<zd> V, w, x, xt, r, t, u, ydev;
    const zd* bp = nullptr;            // the right-hand side (device, caller's)
    Krylov kr;
    std::vector<double> red;
    std::vector<cd> ybuf;             // host source of the last combination's coefficients (kept until copied)
    int products = 0;

    Solver(gt_op* o, int restart_) : op(o), s(o->st->ctx->stream), n(o->st->nb + o->st->nk), restart(restart_) {
        V.alloc((size_t)(restart + 1) * n);
        w.alloc(n);
        x.alloc(n);
        xt.alloc(n);
        r.alloc(n);
        t.alloc(n);
        u.alloc(n);
        ydev.alloc(restart + 1);
        kr.init(n, restart + 1);
        red.resize(2 * (restart + 1) + 8);
    }

    zd* row(int j) { return V.p + (size_t)j * n; }

    // out = K M^-1 v
    void apply(const zd* v, zd* out) {
        op_precond(op, v, t.p);
        op_matvec(op, t.p, out);
        ++products;
    }

    void fetch() {
        GT_HIP(hipMemcpyAsync(red.data(), kr.red.p, red.size() * sizeof(double), hipMemcpyDeviceToHost, s));
        GT_HIP(hipStreamSynchronize(s));
    }

    double norm(const zd* v) {
        kr.norm2(v, 0, s);
        fetch();
        return std::sqrt(red[2 * kr.rows_cap + 2]);
    }

    // x += V[0:k]^T y
    void combine(zd* dst, const std::vector<cd>& y, int k) {
        if (k <= 0) return;
        GT_HIP(hipStreamSynchronize(s));      // the previous copy out of ybuf has finished
        ybuf.assign(y.begin(), y.begin() + k);
        GT_HIP(hipMemcpyAsync(ydev.p, ybuf.data(), sizeof(zd) * k, hipMemcpyHostToDevice, s));
        k_combine(dst, V.p, ydev.p, k, n, s);
    }

    static std::vector<cd> back_substitute(const std::vector<cd>& H, int ld, const std::vector<cd>& g, int k) {
        std::vector<cd> y(k);
        for (int i = k - 1; i >= 0; --i) {
            cd acc = g[i];
            for (int c = i + 1; c < k; ++c) acc -= H[(size_t)i * ld + c] * y[c];
            y[i] = acc / H[(size_t)i * ld + i];
        }
        return y;
    }

    void run(const gt_gmres_params& p, gt_gmres_result& res) {
        const double tol = p.tol;
        const int max_cycles = p.max_cycles, extra = p.extra_cycles;
        const int max_its = p.max_its > 0 ? p.max_its : std::numeric_limits<int>::max();
        const int check_misses = p.check_misses > 0 ? p.check_misses : 4;
        GT_HIP(hipMemsetAsync(x.p, 0, sizeof(zd) * n, s));
        const double b0 = norm(bp);
        const double bnorm = p.bnorm > 0 ? p.bnorm : b0;
        res.iterations = 0;
        res.converged = 0;
        res.products = 0;
        if (bnorm == 0 || b0 <= tol * bnorm) {
            res.converged = 1;
            res.residual = bnorm > 0 ? b0 / bnorm : 0.0;
            return;
        }
        int its = 0;
        GT_HIP(hipMemcpyAsync(r.p, bp, sizeof(zd) * n, hipMemcpyDeviceToDevice, s));
        double last = std::numeric_limits<double>::infinity();
        const int ld = restart;
        std::vector<cd> H((size_t)(restart + 1) * restart), sn(restart), g(restart + 1);
        std::vector<double> cs(restart);
        double rnorm = b0;
        for (int cycle = 0; cycle < max_cycles + extra; ++cycle) {
            const double beta = norm(r.p);
            rnorm = beta;
            if (beta <= tol * bnorm) { finish(res, its, true, beta / bnorm); return; }
            if (its >= max_its) { finish(res, its, false, beta / bnorm); return; }
            if (cycle >= max_cycles && beta > PROGRESS * last) { finish(res, its, false, beta / bnorm); return; }
            last = beta;
            std::fill(H.begin(), H.end(), cd(0));
            std::fill(cs.begin(), cs.end(), 0.0);
            std::fill(sn.begin(), sn.end(), cd(0));
            std::fill(g.begin(), g.end(), cd(0));
            g[0] = beta;
            k_scale(row(0), r.p, 1.0 / beta, n, s);
            int k = 0;
            double aim = tol * bnorm;
            int misses = 0;
            bool restarted = false;
            for (int j = 0; j < restart; ++j) {
                apply(row(j), w.p);
                // one classical Gram-Schmidt pass, |w| before it and after it
                kr.dots(V.p, j + 1, w.p, s);
                kr.project_update(V.p, j + 1, w.p, s);
                fetch();
                const int RC = kr.rows_cap;
                std::vector<cd> h(j + 1);
                for (int i = 0; i <= j; ++i) h[i] = cd(red[2 * i], red[2 * i + 1]);
                const double w_norm = std::sqrt(red[2 * RC]);
                double hn = std::sqrt(red[2 * RC + 1]);
                if (hn < 0.7 * w_norm) {          // DGKS: a second pass only after cancellation
                    kr.dots(V.p, j + 1, w.p, s);
                    kr.project_update(V.p, j + 1, w.p, s);
                    fetch();
                    for (int i = 0; i <= j; ++i) h[i] += cd(red[2 * i], red[2 * i + 1]);
                    hn = std::sqrt(red[2 * RC + 1]);
                }
                std::vector<cd> col(j + 2);
                for (int i = 0; i <= j; ++i) col[i] = h[i];
                col[j + 1] = hn;
                for (int i = 0; i < j; ++i) {
                    const cd a = col[i], c = col[i + 1];
                    col[i] = cs[i] * a + sn[i] * c;
                    col[i + 1] = -std::conj(sn[i]) * a + cs[i] * c;
                }
                const cd a = col[j];
                const double rho = std::hypot(std::abs(a), hn);
                if (rho == 0.0) {
                    cs[j] = 1.0;
                    sn[j] = 0.0;
                } else if (std::abs(a) == 0.0) {
                    cs[j] = 0.0;
                    sn[j] = 1.0;
                    col[j] = hn;
                } else {
                    const cd phase = a / std::abs(a);
                    cs[j] = std::abs(a) / rho;
                    sn[j] = phase * hn / rho;
                    col[j] = phase * rho;
                }
                col[j + 1] = 0.0;
                for (int i = 0; i <= j + 1; ++i) H[(size_t)i * ld + j] = col[i];
                g[j + 1] = -std::conj(sn[j]) * g[j];
                g[j] = cs[j] * g[j];
                ++its;
                k = j + 1;
                double hnorm = 0;
                for (int i = 0; i <= j; ++i) hnorm += std::norm(h[i]);
                hnorm = std::sqrt(hnorm);
                if (hn <= 1e-14 * hnorm || its >= max_its) break;
                if (std::abs(g[j + 1]) <= aim) {
                    // the recurrence meets the aim: check the true residual (one product)
                    std::vector<cd> y = back_substitute(H, ld, g, k);
                    GT_HIP(hipMemcpyAsync(xt.p, x.p, sizeof(zd) * n, hipMemcpyDeviceToDevice, s));
                    combine(xt.p, y, k);
                    apply(xt.p, u.p);
                    k_sub(r.p, bp, u.p, n, s);
                    const double rt = norm(r.p);
                    rnorm = rt;
                    if (rt <= tol * bnorm) {
                        GT_HIP(hipMemcpyAsync(x.p, xt.p, sizeof(zd) * n, hipMemcpyDeviceToDevice, s));
                        finish(res, its, true, rt / bnorm);
                        return;
                    }
                    ++misses;
                    if (misses >= check_misses) {
                        GT_HIP(hipMemcpyAsync(x.p, xt.p, sizeof(zd) * n, hipMemcpyDeviceToDevice, s));
                        restarted = true;
                        break;
                    }
                    aim = 0.5 * std::abs(g[j + 1]);
                }
                k_scale(row(j + 1), w.p, 1.0 / hn, n, s);
            }
            if (restarted) continue;
            while (k > 0 && H[(size_t)(k - 1) * ld + (k - 1)] == cd(0)) --k;
            std::vector<cd> y = k ? back_substitute(H, ld, g, k) : std::vector<cd>();
            combine(x.p, y, k);
            apply(x.p, u.p);
            k_sub(r.p, bp, u.p, n, s);
        }
        rnorm = norm(r.p);
        finish(res, its, rnorm <= tol * bnorm, rnorm / bnorm);
    }

    void finish(gt_gmres_result& res, int its, bool ok, double rel) {
        res.iterations = its;
        res.converged = ok ? 1 : 0;
        res.products = products;
        res.residual = rel;
    }
};

}  // namespace

int default_restart(int n) { return (int)std::min(200.0, std::max(100.0, std::floor(4e8 / (16.0 * n)))); }

void gmres_solve_dev(gt_op* op, const zd* b_dev, zd* x_dev, const gt_gmres_params& p, gt_gmres_result& res) {
    gt_stack* st = op->st;
    const int n = st->nb + st->nk;
    Solver sv(op, p.restart > 0 ? p.restart : default_restart(n));
    sv.bp = b_dev;
    sv.run(p, res);
    // x = M^-1 y (zero when y is)
    op_precond(op, sv.x.p, x_dev);
    GT_HIP(hipStreamSynchronize(sv.s));
}

void gmres_solve(gt_op* op, const double* b_host, double* x_host, const gt_gmres_params& p, gt_gmres_result& res) {
    gt_stack* st = op->st;
    const int n = st->nb + st->nk;
    hipStream_t s = st->ctx->stream;
    DevBuf<zd> b(n), x(n);
    GT_HIP(hipMemcpyAsync(b.p, b_host, sizeof(zd) * n, hipMemcpyHostToDevice, s));
    gmres_solve_dev(op, b.p, x.p, p, res);
    GT_HIP(hipMemcpyAsync(x_host, x.p, sizeof(zd) * n, hipMemcpyDeviceToHost, s));
    GT_HIP(hipStreamSynchronize(s));
}

}  // namespace gt
