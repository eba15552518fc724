// Host side of the Schur factor: geometric nested dissection of the kept nodes, the supernodal structure on the
// dissection tree, and the multifrontal Cholesky of B = Ar^T Ar (schur.hpp).
//
// The reference factorises S = Ar^T D^-1 Ar with SuperLU at every frequency (peec.py:2379-2424, MMD ordering,
// symmetric mode). Here, with every strip alike, S = d_s (B + U C U^T) and only the real B is factorised, once per
// stack; the frequency enters through the scalar d_s and the Woodbury matrix (the caller's part).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>

#include "schur.hpp"

namespace gt {

namespace {

constexpr int ND_LEAF = 32;     // nodes in a leaf box: one dense supernode

struct NDTree {
    const int32_t* xyz;
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
        if ((int)idx.size() <= ND_LEAF || (ilo >= ihi && jlo >= jhi)) {
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
    t.order.reserve(n);
    std::vector<int> all(n);
    std::iota(all.begin(), all.end(), 0);
    std::vector<int> roots = t.rec(all, ilo, ihi, jlo, jhi);
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
    // ---- numeric: multifrontal partial Cholesky per supernode
    poff.assign(ns + 1, 0);
    for (int s = 0; s < ns; ++s) poff[s + 1] = poff[s] + (int64_t)(nc[s] + nr[s]) * nc[s];
    panel.assign(poff[ns], 0.0);
    nnz_L = 0;
    max_cols = 0;
    for (int s = 0; s < ns; ++s) {
        nnz_L += (int64_t)nc[s] * (nc[s] + 1) / 2 + (int64_t)nc[s] * nr[s];
        max_cols = std::max(max_cols, nc[s]);
    }
    std::vector<std::vector<double>> upd(ns);      // update matrices waiting for their parent (nr x nr, lower)
    std::vector<int> pos(n, -1);
    for (int s = 0; s < ns; ++s) {
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
            upd[s] = std::move(U);
        }
        for (int q = 0; q < C; ++q) pos[c0[s] + q] = -1;
        for (int a = 0; a < nr[s]; ++a) pos[rows[roff[s] + a]] = -1;
    }
    // ---- each supernode's rows in its parent's front
    rel_off.assign(ns + 1, 0);
    for (int s = 0; s < ns; ++s) rel_off[s + 1] = rel_off[s] + nr[s];
    rel.assign(rel_off[ns], 0);
    for (int s = 0; s < ns; ++s) {
        const int p = parent[s];
        if (p < 0) continue;
        for (int a = 0; a < nr[s]; ++a) {
            const int c = rows[roff[s] + a];
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
            rel[rel_off[s] + a] = at;
        }
    }
    // ---- front vectors and levels
    foff.assign(ns + 1, 0);
    for (int s = 0; s < ns; ++s) foff[s + 1] = foff[s] + nc[s] + nr[s];
    front_total = foff[ns];
    std::vector<int> h(ns, 0), dep(ns, 0);
    int hmax = 0, dmax = 0;
    for (int s = 0; s < ns; ++s) {
        for (int e = child_ptr[s]; e < child_ptr[s + 1]; ++e) h[s] = std::max(h[s], h[child[e]] + 1);
        hmax = std::max(hmax, h[s]);
    }
    for (int s = ns - 1; s >= 0; --s) {
        dep[s] = parent[s] < 0 ? 0 : dep[parent[s]] + 1;
        dmax = std::max(dmax, dep[s]);
    }
    auto group = [&](const std::vector<int>& key, int kmax, std::vector<int>& ptr, std::vector<int>& lst) {
        ptr.assign(kmax + 2, 0);
        for (int s = 0; s < ns; ++s) ptr[key[s] + 1]++;
        for (int k = 0; k <= kmax; ++k) ptr[k + 1] += ptr[k];
        lst.assign(ns, 0);
        std::vector<int> fill(ptr.begin(), ptr.end() - 1);
        for (int s = 0; s < ns; ++s) lst[fill[key[s]]++] = s;
    };
    group(h, hmax, fwd_levels_ptr, fwd_levels);
    group(dep, dmax, bwd_levels_ptr, bwd_levels);
    (void)roots;
}

void SchurHost::solve_host(const double* b, double* x) const {
    std::vector<double> w(2 * (size_t)n);
    for (int c = 0; c < n; ++c) { w[2 * c] = b[2 * perm[c]]; w[2 * c + 1] = b[2 * perm[c] + 1]; }
    const int ns = (int)c0.size();
    for (int s = 0; s < ns; ++s) {
        const int C = nc[s], m = nc[s] + nr[s];
        const double* L = &panel[poff[s]];
        for (int q = 0; q < C; ++q) {
            const int col = c0[s] + q;
            const double d = L[(size_t)q * m + q];
            w[2 * col] /= d;
            w[2 * col + 1] /= d;
            const double xr = w[2 * col], xi = w[2 * col + 1];
            for (int i = q + 1; i < m; ++i) {
                const int row = i < C ? c0[s] + i : rows[roff[s] + i - C];
                const double l = L[(size_t)q * m + i];
                w[2 * row] -= l * xr;
                w[2 * row + 1] -= l * xi;
            }
        }
    }
    for (int s = ns - 1; s >= 0; --s) {
        const int C = nc[s], m = nc[s] + nr[s];
        const double* L = &panel[poff[s]];
        for (int q = C - 1; q >= 0; --q) {
            const int col = c0[s] + q;
            double ar = w[2 * col], ai = w[2 * col + 1];
            for (int i = q + 1; i < m; ++i) {
                const int row = i < C ? c0[s] + i : rows[roff[s] + i - C];
                const double l = L[(size_t)q * m + i];
                ar -= l * w[2 * row];
                ai -= l * w[2 * row + 1];
            }
            const double d = L[(size_t)q * m + q];
            w[2 * col] = ar / d;
            w[2 * col + 1] = ai / d;
        }
    }
    for (int c = 0; c < n; ++c) { x[2 * perm[c]] = w[2 * c]; x[2 * perm[c] + 1] = w[2 * c + 1]; }
}

}  // namespace gt
