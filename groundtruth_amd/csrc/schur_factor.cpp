// Host side of the Schur factor: geometric nested dissection of the kept nodes, the supernodal structure on the
// dissection tree, and the multifrontal Cholesky of B = Ar^T Ar (schur.hpp).
//
// The reference factorises S = Ar^T D^-1 Ar with SuperLU at every frequency (peec.py:2379-2424, MMD ordering,
// symmetric mode). Here, with every strip alike, S = d_s (B + U C U^T) and only the real B is factorised, once per
// stack; the frequency enters through the scalar d_s and the Woodbury matrix (the caller's part).
#include <algorithm>
#include <atomic>
#include <climits>
#include <cmath>
#include <complex>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <mutex>
#include <numeric>
#include <thread>

#include "schur.hpp"

namespace gt {

namespace {

// Worker threads of the host factorisation: the hardware's, at most 16 (one GPU's share of a host), fewer when
// OMP_NUM_THREADS asks for fewer.
int host_threads() {
    int t = (int)std::thread::hardware_concurrency();
    t = std::min(std::max(t, 1), 16);
    if (const char* e = std::getenv("OMP_NUM_THREADS")) {
        const int v = std::atoi(e);
        if (v > 0) t = std::min(t, v);
    }
    return t;
}

// f(i, thread) for every i in [0, n) on up to host_threads() threads, the items taken in order from a shared
// counter; the first exception stops the others and is rethrown once every thread has ended.
template <class F>
void parallel_for(int64_t n, F&& f) {
    const int nt = (int)std::min<int64_t>(host_threads(), n);
    if (nt <= 1) {
        for (int64_t i = 0; i < n; ++i) f(i, 0);
        return;
    }
    std::atomic<int64_t> next(0);
    std::atomic<bool> stop(false);
    std::exception_ptr err;
    std::mutex mu;
    auto work = [&](int tid) {
        try {
            for (int64_t i = next.fetch_add(1); i < n && !stop.load(); i = next.fetch_add(1)) f(i, tid);
        } catch (...) {
            std::lock_guard<std::mutex> lk(mu);
            if (!err) err = std::current_exception();
            stop = true;
        }
    };
    std::vector<std::thread> th;
    th.reserve(nt - 1);
    for (int t = 1; t < nt; ++t) th.emplace_back(work, t);
    work(0);
    for (std::thread& x : th) x.join();
    if (err) std::rethrow_exception(err);
}


struct NDTree {
    const int32_t* xyz;
    int leaf = 32;                          // nodes in a leaf box at most: one dense supernode
    std::vector<int> order;                 // new -> old
    std::vector<int> c0, nc;
    std::vector<std::vector<int>> kids;

    int make(int first, int count, std::vector<int> k) {
        c0.push_back(first);
        nc.push_back(count);
        kids.push_back(std::move(k));
        return (int)c0.size() - 1;
    }

    // Orders `idx` (a box [ilo, ihi] x [jlo, jhi]) and returns the supernodes at its top (more than one when the
    // separator line is empty).
    std::vector<int> rec(std::vector<int>& idx, int ilo, int ihi, int jlo, int jhi) {
        if (idx.empty()) return {};
        auto I = [&](int v) { return xyz[3 * v + 1]; };
        auto J = [&](int v) { return xyz[3 * v + 2]; };
        auto K = [&](int v) { return xyz[3 * v + 0]; };
        if ((int)idx.size() <= leaf || (ilo >= ihi && jlo >= jhi)) {
            std::sort(idx.begin(), idx.end(), [&](int a, int b) {
                if (I(a) != I(b)) return I(a) < I(b);
                if (J(a) != J(b)) return J(a) < J(b);
                return K(a) < K(b);
            });
            const int first = (int)order.size();
            order.insert(order.end(), idx.begin(), idx.end());
            return {make(first, (int)idx.size(), {})};
        }
        const bool split_i = (ihi - ilo) >= (jhi - jlo);
        const int lo = split_i ? ilo : jlo, hi = split_i ? ihi : jhi;
        const int m = lo + (hi - lo) / 2;
        std::vector<int> L, R, Sep;
        for (int v : idx) {
            const int c = split_i ? I(v) : J(v);
            (c < m ? L : c > m ? R : Sep).push_back(v);
        }
        std::vector<int>().swap(idx);
        std::vector<int> kl = split_i ? rec(L, ilo, m - 1, jlo, jhi) : rec(L, ilo, ihi, jlo, m - 1);
        std::vector<int> kr = split_i ? rec(R, m + 1, ihi, jlo, jhi) : rec(R, ilo, ihi, m + 1, jhi);
        kl.insert(kl.end(), kr.begin(), kr.end());
        if (Sep.empty()) return kl;
        std::sort(Sep.begin(), Sep.end(), [&](int a, int b) {
            const int oa = split_i ? J(a) : I(a), ob = split_i ? J(b) : I(b);
            if (oa != ob) return oa < ob;
            return K(a) < K(b);
        });
        const int first = (int)order.size();
        order.insert(order.end(), Sep.begin(), Sep.end());
        return {make(first, (int)Sep.size(), std::move(kl))};
    }
};

}  // namespace

void SchurHost::build(int n_, const std::vector<int>& ea, const std::vector<int>& eb, const std::vector<double>& diag,
                      const int32_t* xyz) {
    n = n_;
    // ---- ordering
    int ilo = 1 << 30, ihi = -1, jlo = 1 << 30, jhi = -1;
    for (int v = 0; v < n; ++v) {
        ilo = std::min(ilo, xyz[3 * v + 1]);
        ihi = std::max(ihi, xyz[3 * v + 1]);
        jlo = std::min(jlo, xyz[3 * v + 2]);
        jhi = std::max(jhi, xyz[3 * v + 2]);
    }
    NDTree t;
    t.xyz = xyz;
    t.leaf = leaf;
    t.order.reserve(n);
    int nsheet = 0;
    for (int v = 0; v < n; ++v) nsheet = std::max(nsheet, xyz[3 * v] + 1);
    std::vector<char> cross(n, 0);
    int n_cross = 0;
    if (sheets_apart && nsheet > 1) {
        for (size_t e = 0; e < ea.size(); ++e)
            if (xyz[3 * ea[e]] != xyz[3 * eb[e]]) cross[ea[e]] = cross[eb[e]] = 1;
        for (int v = 0; v < n; ++v) n_cross += cross[v];
    }
    std::vector<int> roots;
    if (sheets_apart && nsheet > 1 && n_cross <= SHEETS_APART_MAX) {
        std::vector<std::vector<int>> per(nsheet);
        std::vector<int> top;
        for (int v = 0; v < n; ++v) (cross[v] ? top : per[xyz[3 * v]]).push_back(v);
        std::vector<int> kids;
        for (int k = 0; k < nsheet; ++k) {
            if (per[k].empty()) continue;
            int a0 = 1 << 30, a1 = -1, b0 = 1 << 30, b1 = -1;
            for (int v : per[k]) {
                a0 = std::min(a0, xyz[3 * v + 1]);
                a1 = std::max(a1, xyz[3 * v + 1]);
                b0 = std::min(b0, xyz[3 * v + 2]);
                b1 = std::max(b1, xyz[3 * v + 2]);
            }
            std::vector<int> r = t.rec(per[k], a0, a1, b0, b1);
            kids.insert(kids.end(), r.begin(), r.end());
        }
        if (top.empty()) {
            roots = kids;
        } else {
            std::sort(top.begin(), top.end(), [&](int a, int b) {
                for (int c : {1, 2, 0})
                    if (xyz[3 * a + c] != xyz[3 * b + c]) return xyz[3 * a + c] < xyz[3 * b + c];
                return false;
            });
            const int first = (int)t.order.size();
            t.order.insert(t.order.end(), top.begin(), top.end());
            roots = {t.make(first, (int)top.size(), std::move(kids))};
        }
    } else {
        std::vector<int> all(n);
        std::iota(all.begin(), all.end(), 0);
        roots = t.rec(all, ilo, ihi, jlo, jhi);
    }
    perm = t.order;
    if ((int)perm.size() != n) fail(GT_ERR_ARG, "nested dissection lost nodes");
    std::vector<int> iperm(n);
    for (int c = 0; c < n; ++c) iperm[perm[c]] = c;
    const int ns = (int)t.c0.size();
    c0 = t.c0;
    nc = t.nc;
    parent.assign(ns, -1);
    for (int s = 0; s < ns; ++s)
        for (int k : t.kids[s]) parent[k] = s;
    child_ptr.assign(ns + 1, 0);
    for (int s = 0; s < ns; ++s) child_ptr[s + 1] = child_ptr[s] + (int)t.kids[s].size();
    child.clear();
    for (int s = 0; s < ns; ++s) {
        std::vector<int> k = t.kids[s];
        std::sort(k.begin(), k.end());
        child.insert(child.end(), k.begin(), k.end());
    }
    // ---- adjacency (old numbering, both directions, repeated edges kept: each adds -1)
    std::vector<int> aptr(n + 1, 0);
    for (size_t e = 0; e < ea.size(); ++e) { aptr[ea[e] + 1]++; aptr[eb[e] + 1]++; }
    for (int v = 0; v < n; ++v) aptr[v + 1] += aptr[v];
    std::vector<int> adj(aptr[n]);
    {
        std::vector<int> fill(aptr.begin(), aptr.end() - 1);
        for (size_t e = 0; e < ea.size(); ++e) { adj[fill[ea[e]]++] = eb[e]; adj[fill[eb[e]]++] = ea[e]; }
    }
    // ---- symbolic: off-diagonal rows of each supernode = its columns' neighbours and its children's rows above it
    nr.assign(ns, 0);
    roff.assign(ns + 1, 0);
    rows.clear();
    std::vector<int> mark(n, -1);
    for (int s = 0; s < ns; ++s) {
        const int top = c0[s] + nc[s];
        std::vector<int> set;
        for (int q = 0; q < nc[s]; ++q) {
            const int old = perm[c0[s] + q];
            for (int e = aptr[old]; e < aptr[old + 1]; ++e) {
                const int c = iperm[adj[e]];
                if (c < c0[s]) {
                    // a neighbour eliminated earlier must lie in this supernode's subtree (it is a descendant's
                    // column); the tree is checked below through the children's rows
                    continue;
                }
                if (c >= top && mark[c] != s) { mark[c] = s; set.push_back(c); }
            }
        }
        for (int e = child_ptr[s]; e < child_ptr[s + 1]; ++e) {
            const int k = child[e];
            for (int64_t a = roff[k]; a < roff[k] + nr[k]; ++a) {
                const int c = rows[a];
                if (c < c0[s])
                    fail(GT_ERR_UNSUPPORTED, "a branch joins cells the nested dissection separated");
                if (c >= top && mark[c] != s) { mark[c] = s; set.push_back(c); }
            }
        }
        std::sort(set.begin(), set.end());
        nr[s] = (int)set.size();
        rows.insert(rows.end(), set.begin(), set.end());
        roff[s + 1] = roff[s] + nr[s];
    }
    // every edge from a column to one eliminated before it must come from its subtree: check the subtree ranges
    {
        std::vector<int> lo(ns);
        for (int s = 0; s < ns; ++s) {
            lo[s] = c0[s];
            for (int e = child_ptr[s]; e < child_ptr[s + 1]; ++e) lo[s] = std::min(lo[s], lo[child[e]]);
        }
        for (int s = 0; s < ns; ++s)
            for (int q = 0; q < nc[s]; ++q) {
                const int old = perm[c0[s] + q];
                for (int e = aptr[old]; e < aptr[old + 1]; ++e) {
                    const int c = iperm[adj[e]];
                    if (c < lo[s]) fail(GT_ERR_UNSUPPORTED, "a branch joins cells the nested dissection separated");
                }
            }
    }
    for (int s = 0; s < ns; ++s)
        if (parent[s] < 0 && nr[s] > 0) fail(GT_ERR_UNSUPPORTED, "an assembly-tree root has rows above it");
    // ---- which fronts the device takes: the large ones and their ancestors (children come first in postorder)
    on_dev.assign(ns, 0);
    dev_order.clear();
    for (int s = 0; s < ns; ++s) {
        bool d = (int64_t)nc[s] + nr[s] >= (int64_t)dev_min;
        for (int e = child_ptr[s]; e < child_ptr[s + 1]; ++e) d = d || on_dev[child[e]];
        on_dev[s] = d ? 1 : 0;
        if (d) dev_order.push_back(s);
    }
    host_panels = dev_order.empty();
    trip_off.assign(1, 0);
    trip_r.clear();
    trip_c.clear();
    trip_v.clear();
    upd_host.assign(ns, std::vector<double>());
    // ---- numeric: multifrontal partial Cholesky per supernode
    poff.assign(ns + 1, 0);
    for (int s = 0; s < ns; ++s) poff[s + 1] = poff[s] + (int64_t)(nc[s] + nr[s]) * nc[s];
    std::vector<double, UninitAlloc<double>>().swap(panel);
    panel.resize(poff[ns]);
    nnz_L = 0;
    max_cols = 0;
    for (int s = 0; s < ns; ++s) {
        nnz_L += (int64_t)nc[s] * (nc[s] + 1) / 2 + (int64_t)nc[s] * nr[s];
        max_cols = std::max(max_cols, nc[s]);
    }
    std::vector<std::vector<double>> upd(ns);      // update matrices waiting for their parent (nr x nr, lower)
    std::vector<int> pos(n, -1);
    // the device fronts' original entries, in postorder
    for (int s = 0; s < ns; ++s) {
        const int C = nc[s], m = nc[s] + nr[s];
        if (on_dev[s]) {
            // the front's original entries for the device: (front row, front column) in the lower triangle
            for (int q = 0; q < C; ++q) pos[c0[s] + q] = q;
            for (int a = 0; a < nr[s]; ++a) pos[rows[roff[s] + a]] = C + a;
            std::vector<std::pair<int64_t, double>> ent;
            for (int q = 0; q < C; ++q) {
                const int c = c0[s] + q, old = perm[c];
                ent.push_back({(int64_t)q * m + q, diag[old]});
                for (int e = aptr[old]; e < aptr[old + 1]; ++e) {
                    const int c2 = iperm[adj[e]];
                    if (c2 > c) ent.push_back({(int64_t)q * m + pos[c2], -1.0});
                }
            }
            std::sort(ent.begin(), ent.end(),
                      [](const std::pair<int64_t, double>& x, const std::pair<int64_t, double>& y) {
                          return x.first < y.first;
                      });
            for (size_t e = 0; e < ent.size(); ++e) {
                if (e > 0 && ent[e].first == ent[e - 1].first) {
                    trip_v.back() += ent[e].second;
                    continue;
                }
                trip_c.push_back((int32_t)(ent[e].first / m));
                trip_r.push_back((int32_t)(ent[e].first % m));
                trip_v.push_back(ent[e].second);
            }
            trip_off.push_back((int64_t)trip_v.size());
            for (int q = 0; q < C; ++q) pos[c0[s] + q] = -1;
            for (int a = 0; a < nr[s]; ++a) pos[rows[roff[s] + a]] = -1;
        }
    }
    // The host fronts form whole subtrees (a device front's ancestors are device fronts), each a contiguous range
    // [s - size + 1, s] of the postorder: the maximal ones are independent tasks, run on worker threads, largest
    // first, each in postorder with a position map of its own. Every front's arithmetic is the same in any task
    // order, so the factor's bits do not depend on the threads.
    std::vector<int> sub(ns, 1);
    for (int s = 0; s < ns; ++s)
        for (int e = child_ptr[s]; e < child_ptr[s + 1]; ++e) sub[s] += sub[child[e]];
    std::vector<int> tasks;
    for (int s = 0; s < ns; ++s)
        if (!on_dev[s] && (parent[s] < 0 || on_dev[parent[s]])) tasks.push_back(s);
    for (int s = 0; s < ns; ++s)
        for (int e = child_ptr[s]; e < child_ptr[s + 1]; ++e)
            if (child[e] >= s || child[e] < s - sub[s] + 1) fail(GT_ERR_ARG, "the assembly tree is not in postorder");
    std::stable_sort(tasks.begin(), tasks.end(), [&](int a, int b) { return sub[a] > sub[b]; });
    const int nthr = (int)std::min<int64_t>(host_threads(), std::max<size_t>(1, tasks.size()));
    std::vector<std::vector<int>> tpos(nthr);
    tpos[0].swap(pos);
    parallel_for((int64_t)tasks.size(), [&](int64_t ti, int tid) {
        std::vector<int>& pos = tpos[tid];
        if (pos.empty()) pos.assign(n, -1);
        const int root = tasks[ti];
        for (int s = root - sub[root] + 1; s <= root; ++s) {
            const int C = nc[s], m = nc[s] + nr[s];
            std::vector<double> F((size_t)m * m, 0.0);
            for (int q = 0; q < C; ++q) pos[c0[s] + q] = q;
            for (int a = 0; a < nr[s]; ++a) pos[rows[roff[s] + a]] = C + a;
            for (int q = 0; q < C; ++q) {
                const int c = c0[s] + q, old = perm[c];
                F[(size_t)q * m + q] += diag[old];
                for (int e = aptr[old]; e < aptr[old + 1]; ++e) {
                    const int c2 = iperm[adj[e]];
                    if (c2 > c) F[(size_t)q * m + pos[c2]] -= 1.0;
                }
            }
            for (int e = child_ptr[s]; e < child_ptr[s + 1]; ++e) {
                const int k = child[e];
                const int r = nr[k];
                const std::vector<double>& U = upd[k];
                for (int b = 0; b < r; ++b) {
                    const int pb = pos[rows[roff[k] + b]];
                    for (int a = b; a < r; ++a) F[(size_t)pb * m + pos[rows[roff[k] + a]]] += U[(size_t)b * r + a];
                }
                std::vector<double>().swap(upd[k]);
            }
            // right-looking partial Cholesky of the first C columns; the trailing block becomes the update matrix
            for (int j = 0; j < C; ++j) {
                double* cj = &F[(size_t)j * m];
                const double d = cj[j];
                if (!(d > 0.0) || !std::isfinite(d))
                    fail(GT_ERR_NUMERIC, "the Schur complement is not positive definite (a piece without a reference node?)");
                const double sd = std::sqrt(d);
                cj[j] = sd;
                const double inv = 1.0 / sd;
                for (int i = j + 1; i < m; ++i) cj[i] *= inv;
                for (int k = j + 1; k < m; ++k) {
                    const double f = cj[k];
                    if (f == 0.0) continue;
                    double* ck = &F[(size_t)k * m];
                    for (int i = k; i < m; ++i) ck[i] -= cj[i] * f;
                }
            }
            std::memcpy(&panel[poff[s]], F.data(), sizeof(double) * (size_t)m * C);
            if (nr[s] > 0) {
                const int r = nr[s];
                std::vector<double> U((size_t)r * r);
                for (int b = 0; b < r; ++b)
                    for (int a = 0; a < r; ++a) U[(size_t)b * r + a] = F[(size_t)(C + b) * m + C + a];
                if (parent[s] >= 0 && on_dev[parent[s]]) upd_host[s] = std::move(U);
                else upd[s] = std::move(U);
            }
            for (int q = 0; q < C; ++q) pos[c0[s] + q] = -1;
            for (int a = 0; a < nr[s]; ++a) pos[rows[roff[s] + a]] = -1;
        }
    });
    std::vector<std::vector<int>>().swap(tpos);
    // ---- solve panels: Q = [L_CC^-1 ; -L_RC L_CC^-1] in place of each L panel (each supernode on its own)
    std::vector<int> hosts;
    for (int sn = 0; sn < ns; ++sn)
        if (!on_dev[sn]) hosts.push_back(sn);
    parallel_for((int64_t)hosts.size(), [&](int64_t hi, int) {
        const int sn = hosts[hi];
        const int C = nc[sn], m = nc[sn] + nr[sn];
        double* P = &panel[poff[sn]];
        std::vector<double> inv((size_t)C * C, 0.0);      // column-major, lower
        for (int j = 0; j < C; ++j) {
            inv[(size_t)j * C + j] = 1.0 / P[(size_t)j * m + j];
            for (int i = j + 1; i < C; ++i) {
                double acc = 0.0;
                for (int k = j; k < i; ++k) acc += P[(size_t)k * m + i] * inv[(size_t)j * C + k];
                inv[(size_t)j * C + i] = -acc / P[(size_t)i * m + i];
            }
        }
        // -L_RC L_CC^-1, column j = -sum_{k >= j} L_RC[:, k] inv[k, j]
        const int R = nr[sn];
        std::vector<double> M((size_t)R * C, 0.0);
        for (int j = 0; j < C; ++j)
            for (int k = j; k < C; ++k) {
                const double f = inv[(size_t)j * C + k];
                if (f == 0.0) continue;
                const double* Lk = P + (size_t)k * m + C;
                double* Mj = &M[(size_t)j * R];
                for (int a = 0; a < R; ++a) Mj[a] -= Lk[a] * f;
            }
        for (int j = 0; j < C; ++j) {
            for (int i = 0; i < C; ++i) P[(size_t)j * m + i] = i >= j ? inv[(size_t)j * C + i] : 0.0;
            for (int a = 0; a < R; ++a) P[(size_t)j * m + C + a] = M[(size_t)j * R + a];
        }
    });
    // ---- update vectors and the contributors of each front position (children ascending)
    uoff.assign(ns + 1, 0);
    for (int sn = 0; sn < ns; ++sn) uoff[sn + 1] = uoff[sn] + nr[sn];
    upd_total = uoff[ns];
    fpos.assign(ns + 1, 0);
    for (int sn = 0; sn < ns; ++sn) fpos[sn + 1] = fpos[sn] + nc[sn] + nr[sn];
    std::vector<int64_t> at_of(upd_total, 0);          // front position (global) each update entry adds into
    for (int sn = 0; sn < ns; ++sn) {
        const int p = parent[sn];
        if (p < 0) continue;
        for (int a = 0; a < nr[sn]; ++a) {
            const int c = rows[roff[sn] + a];
            int at;
            if (c < c0[p] + nc[p]) {
                at = c - c0[p];
            } else {
                const int* b = rows.data() + roff[p];
                const int* e = b + nr[p];
                const int* f = std::lower_bound(b, e, c);
                if (f == e || *f != c) fail(GT_ERR_ARG, "a child's row is not in its parent's front");
                at = nc[p] + (int)(f - b);
            }
            if (at < 0) fail(GT_ERR_ARG, "a child's row lies below its parent");
            at_of[uoff[sn] + a] = fpos[p] + at;
        }
    }
    cptr.assign(fpos[ns] + 1, 0);
    for (int sn = 0; sn < ns; ++sn)
        if (parent[sn] >= 0)
            for (int a = 0; a < nr[sn]; ++a) cptr[at_of[uoff[sn] + a] + 1]++;
    for (int64_t q = 0; q < fpos[ns]; ++q) cptr[q + 1] += cptr[q];
    cidx.assign(cptr[fpos[ns]], 0);
    {
        std::vector<int64_t> fill(cptr.begin(), cptr.end() - 1);
        for (int sn = 0; sn < ns; ++sn)         // ascending: each position's contributors in child order
            if (parent[sn] >= 0)
                for (int a = 0; a < nr[sn]; ++a) cidx[fill[at_of[uoff[sn] + a]]++] = uoff[sn] + a;
    }
    // ---- levels and their work items
    std::vector<int> h(ns, 0), dep(ns, 0);
    int hmax = 0, dmax = 0;
    for (int sn = 0; sn < ns; ++sn) {
        for (int e = child_ptr[sn]; e < child_ptr[sn + 1]; ++e) h[sn] = std::max(h[sn], h[child[e]] + 1);
        hmax = std::max(hmax, h[sn]);
    }
    for (int sn = ns - 1; sn >= 0; --sn) {
        dep[sn] = parent[sn] < 0 ? 0 : dep[parent[sn]] + 1;
        dmax = std::max(dmax, dep[sn]);
    }
    levels = ns ? hmax + 1 : 0;
    // ---- the dense top: the deepest top_T within the budgets (0 if below SCHUR_TOP_MIN_T)
    top_T = 0;
    {
        std::vector<int64_t> sub(ns), chain(ns);
        for (int T = dmax + 1; T >= SCHUR_TOP_MIN_T && ns; --T) {
            int64_t ncol = 0, fnnz = 0;
            for (int sn = 0; sn < ns; ++sn) {           // postorder: a supernode's top subtree's columns
                if (dep[sn] >= T) continue;
                sub[sn] = nc[sn];
                for (int e = child_ptr[sn]; e < child_ptr[sn + 1]; ++e)
                    if (dep[child[e]] < T) sub[sn] += sub[child[e]];
                ncol += nc[sn];
                fnnz += (int64_t)nc[sn] * (sub[sn] - nc[sn]) + (int64_t)nc[sn] * (nc[sn] + 1) / 2;
            }
            if (ncol <= SCHUR_TOP_COLS && fnnz <= SCHUR_TOP_NNZ) {
                top_T = T;
                break;
            }
        }
        (void)chain;
    }
    ctab.assign(2 * (size_t)fpos[ns], -1);
    for (int64_t q = 0; q < fpos[ns]; ++q) {
        const int64_t k = cptr[q + 1] - cptr[q];
        if (k > 2) {
            if (cptr[q] > INT32_MAX - 2) fail(GT_ERR_UNSUPPORTED, "contributor lists too long");
            ctab[2 * q] = -2 - (int32_t)cptr[q];
            ctab[2 * q + 1] = (int32_t)k;
        } else {
            for (int64_t e = 0; e < k; ++e) {
                if (cidx[cptr[q] + e] > INT32_MAX) fail(GT_ERR_UNSUPPORTED, "update buffer too long");
                ctab[2 * q + e] = (int32_t)cidx[cptr[q] + e];
            }
        }
    }
    slots = 0;
    // the items of one tile: chunks [lo, hi) of `step` over its reduction, split when more than one (then with a
    // finalize item in `fin`)
    auto tile_items = [&](std::vector<SnItem>& out, std::vector<SnItem>& fin, int sn, int t0, int kw, int lo, int hi,
                          int step, bool forward) {
        const int ks = std::max(1, (hi - lo + step - 1) / step);
        for (int k = 0; k < ks; ++k) {
            SnItem it{};
            it.s = sn;
            it.t0 = t0;
            it.kw = kw;
            it.C = nc[sn];
            it.m = nc[sn] + nr[sn];
            it.c0 = c0[sn];
            it.k0 = lo + k * step;
            it.k1 = std::min(hi, lo + (k + 1) * step);
            it.slot0 = ks > 1 ? (int32_t)slots : -1;
            it.kidx = k;
            it.ks = ks;
            it.poff = poff[sn];
            it.aux = forward ? uoff[sn] : roff[sn];
            it.fpos = fpos[sn];
            out.push_back(it);
        }
        if (ks > 1) {
            fin.push_back(out.back());
            slots += ks;
        }
        return std::min(step, hi - lo);
    };
    // forward levels by height; the top supernodes' items apart (tfwd: used only to make M)
    for (int pass = 0; pass < 2; ++pass) {
        const bool top = pass == 1;
        std::vector<int>& ptr = top ? tfwd_ptr : fwd_ptr;
        std::vector<int>& colsv = top ? tfwd_cols : fwd_cols;
        std::vector<int>& fptr = top ? tffin_ptr : ffin_ptr;
        std::vector<SnItem>& items = top ? tfwd : fwd;
        std::vector<SnItem>& fins = top ? tffin : ffin;
        ptr.assign(1, 0);
        items.clear();
        colsv.clear();
        fptr.assign(1, 0);
        fins.clear();
        for (int l = 0; l <= hmax && ns; ++l) {
            int cols = SCHUR_TILE;          // the 16-row cross-wave reduction reuses the staging buffer
            for (int sn = 0; sn < ns; ++sn) {
                if (h[sn] != l || (dep[sn] < top_T) != top) continue;
                const int m = nc[sn] + nr[sn], narrow = nc[sn] <= SCHUR_NARROW;
                const int tile = narrow ? 64 : SCHUR_TILE, step = SCHUR_PV * (narrow ? 4 : 16);
                for (int r0 = 0; r0 < m; r0 += tile) {
                    const int qend = std::min(nc[sn], std::min(m, r0 + tile));  // zeros above L^-1 beyond
                    cols = std::max(cols, tile_items(items, fins, sn, r0, narrow ? 1 : 4, 0, qend, step, true));
                }
            }
            ptr.push_back((int)items.size());
            fptr.push_back((int)fins.size());
            colsv.push_back(cols);
        }
    }
    bwd_ptr.assign(1, 0);
    bwd.clear();
    bwd_rows.clear();
    bfin_ptr.assign(1, 0);
    bfin.clear();
    for (int l = 0; l <= dmax && ns; ++l) {
        int rws = 1;
        for (int sn = 0; sn < ns; ++sn) {
            if (dep[sn] != l || dep[sn] < top_T) continue;
            const int m = nc[sn] + nr[sn];
            const bool wide_tile = nc[sn] > 16 && nc[sn] <= 32 && m <= SCHUR_PV * 8;
            const int rp = wide_tile ? 8 : 16, tile = wide_tile ? 32 : SCHUR_TILE;
            for (int q0 = 0; q0 < nc[sn]; q0 += tile)   // rows above q0 meet the zeros above L^-1
                rws = std::max(rws, tile_items(bwd, bfin, sn, q0, rp, q0, m, SCHUR_PV * rp, false));
        }
        bwd_ptr.push_back((int)bwd.size());
        bfin_ptr.push_back((int)bfin.size());
        bwd_rows.push_back(rws);
    }
    build_top(dep);
    (void)roots;
}

void SchurHost::build_top(const std::vector<int>& dep) {
    const int ns = (int)c0.size();
    top_col.clear();
    top_lo.clear();
    top_foff.clear();
    top_cptr.clear();
    top_clen.clear();
    top_chain.clear();
    top_boff.clear();
    top_gptr.assign(1, 0);
    top_gidx.clear();
    top_fwd.clear();
    top_ffin.clear();
    top_bwd.clear();
    top_bfin.clear();
    top_slots = 0;
    top_fnnz = top_bnnz = 0;
    if (top_T == 0) return;
    // top columns, ascending (supernodes in postorder hold ascending columns); each supernode's first index
    std::vector<int> own(ns, -1), lo(ns, -1);
    std::vector<int> idx_of(n, -1);
    for (int sn = 0; sn < ns; ++sn) {
        if (dep[sn] >= top_T) continue;
        own[sn] = (int)top_col.size();
        for (int q = 0; q < nc[sn]; ++q) {
            idx_of[c0[sn] + q] = (int)top_col.size();
            top_col.push_back(c0[sn] + q);
        }
    }
    const int nt = (int)top_col.size();
    // a supernode's top subtree is contiguous in top-column order (postorder), ending at its own columns
    for (int sn = 0; sn < ns; ++sn) {
        if (dep[sn] >= top_T) continue;
        lo[sn] = own[sn];
        for (int e = child_ptr[sn]; e < child_ptr[sn + 1]; ++e)
            if (dep[child[e]] < top_T) lo[sn] = std::min(lo[sn], lo[child[e]]);
    }
    std::vector<int> sn_of(nt);
    for (int sn = 0; sn < ns; ++sn)
        if (dep[sn] < top_T)
            for (int q = 0; q < nc[sn]; ++q) sn_of[own[sn] + q] = sn;
    top_lo.resize(nt);
    top_foff.resize(nt);
    for (int i = 0; i < nt; ++i) {
        top_lo[i] = lo[sn_of[i]];
        top_foff[i] = top_fnnz;
        top_fnnz += i - top_lo[i] + 1;
    }
    // chains: a supernode's columns, then its top ancestors' (ascending, as postorder puts ancestors later)
    std::vector<int> cptr_sn(ns, -1);
    for (int sn = 0; sn < ns; ++sn) {
        if (dep[sn] >= top_T) continue;
        cptr_sn[sn] = (int)top_chain.size();
        for (int a = sn; a >= 0; a = parent[a])
            for (int q = 0; q < nc[a]; ++q) top_chain.push_back(c0[a] + q);
    }
    top_cptr.resize(nt);
    top_clen.resize(nt);
    top_boff.resize(nt);
    for (int i = 0; i < nt; ++i) {
        const int sn = sn_of[i];
        int len = 0;
        for (int a = sn; a >= 0; a = parent[a]) len += nc[a];
        top_cptr[i] = cptr_sn[sn];
        top_clen[i] = len;
        top_boff[i] = top_bnnz;
        top_bnnz += len;
    }
    // g: the update entries of non-top children at every front position of a top supernode, by the top column the
    // position stands for (its own column, or an ancestor's among its off-diagonal rows)
    std::vector<std::vector<int64_t>> g(nt);
    for (int sn = 0; sn < ns; ++sn) {
        if (dep[sn] >= top_T) continue;
        const int m = nc[sn] + nr[sn];
        for (int q = 0; q < m; ++q) {
            const int col = q < nc[sn] ? c0[sn] + q : rows[roff[sn] + q - nc[sn]];
            const int i = idx_of[col];
            if (i < 0) fail(GT_ERR_NUMERIC, "dense top: a front row outside the top");
            const int64_t pos = fpos[sn] + q;
            for (int64_t e = cptr[pos]; e < cptr[pos + 1]; ++e) {
                const int64_t u = cidx[e];
                // the child owning update entry u: a child's entries are [uoff, uoff + nr)
                int owner = -1;
                for (int c = child_ptr[sn]; c < child_ptr[sn + 1]; ++c) {
                    const int ch = child[c];
                    if (u >= uoff[ch] && u < uoff[ch] + nr[ch]) { owner = ch; break; }
                }
                if (owner < 0) fail(GT_ERR_NUMERIC, "dense top: an update entry of no child");
                if (dep[owner] >= top_T) g[i].push_back(u);
            }
        }
    }
    for (int i = 0; i < nt; ++i) {
        std::sort(g[i].begin(), g[i].end());
        top_gidx.insert(top_gidx.end(), g[i].begin(), g[i].end());
        top_gptr.push_back((int64_t)top_gidx.size());
    }
    // items: tiles of SCHUR_TOP_ROWS rows (forward) or columns (backward) of one supernode, reductions in chunks of
    // SCHUR_TOP_CHUNK
    auto items = [&](std::vector<TopItem>& out, std::vector<TopItem>& fin, int i0, int cnt, int k0, int k1) {
        const int ks = std::max(1, (k1 - k0 + SCHUR_TOP_CHUNK - 1) / SCHUR_TOP_CHUNK);
        for (int k = 0; k < ks; ++k) {
            TopItem it{};
            it.i0 = i0;
            it.cnt = cnt;
            it.k0 = k0 + k * SCHUR_TOP_CHUNK;
            it.k1 = std::min(k1, k0 + (k + 1) * SCHUR_TOP_CHUNK);
            it.slot0 = ks > 1 ? (int32_t)top_slots : -1;
            it.kidx = k;
            it.ks = ks;
            out.push_back(it);
        }
        if (ks > 1) {
            fin.push_back(out.back());
            top_slots += ks;
        }
    };
    for (int sn = 0; sn < ns; ++sn) {
        if (dep[sn] >= top_T) continue;
        for (int r0 = 0; r0 < nc[sn]; r0 += SCHUR_TOP_ROWS) {
            const int i0 = own[sn] + r0, cnt = std::min(SCHUR_TOP_ROWS, nc[sn] - r0);
            items(top_fwd, top_ffin, i0, cnt, lo[sn], i0 + cnt);              // M[i][lo .. i]
            items(top_bwd, top_bfin, i0, cnt, 0, top_clen[i0]);               // M[chain][i]
        }
    }
}

void SchurHost::solve_host(const double* b, double* x) const {
    typedef std::complex<double> cd;
    const int ns = (int)c0.size();
    std::vector<cd> zp(n), xp(n), ub(upd_total);
    for (int sn = 0; sn < ns; ++sn) {
        const int C = nc[sn], m = nc[sn] + nr[sn];
        const double* Q = &panel[poff[sn]];
        std::vector<cd> f(m);
        for (int q = 0; q < m; ++q) {
            cd acc = q < C ? cd(b[2 * perm[c0[sn] + q]], b[2 * perm[c0[sn] + q] + 1]) : cd(0);
            const int64_t pos = fpos[sn] + q;
            for (int64_t e = cptr[pos]; e < cptr[pos + 1]; ++e) acc += ub[cidx[e]];
            f[q] = acc;
        }
        for (int r = 0; r < m; ++r) {
            cd acc = r < C ? cd(0) : f[r];
            const int qmax = r < C ? r + 1 : C;
            for (int q = 0; q < qmax; ++q) acc += Q[(size_t)q * m + r] * f[q];
            if (r < C) zp[c0[sn] + r] = acc;
            else ub[uoff[sn] + r - C] = acc;
        }
    }
    for (int sn = ns - 1; sn >= 0; --sn) {
        const int C = nc[sn], m = nc[sn] + nr[sn];
        const double* Q = &panel[poff[sn]];
        for (int q = 0; q < C; ++q) {
            cd acc = 0;
            for (int r = q; r < m; ++r) {
                const cd v = r < C ? zp[c0[sn] + r] : xp[rows[roff[sn] + r - C]];
                acc += Q[(size_t)q * m + r] * v;
            }
            xp[c0[sn] + q] = acc;
        }
    }
    for (int c = 0; c < n; ++c) { x[2 * perm[c]] = xp[c].real(); x[2 * perm[c] + 1] = xp[c].imag(); }
}

}  // namespace gt
