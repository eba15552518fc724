// The large fronts of the multifrontal Schur factor on the device (the host's numeric and panel steps of
// schur_factor.cpp, for the fronts it leaves out): per front in postorder,
//   1. assembly: the original entries of B's columns, then each child's update matrix added at its rows' positions
//      (a device child's straight from its front, a host child's uploaded);
//   2. partial Cholesky of the first C of the front's m columns, blocked by 32 columns: the diagonal block factored
//      and inverted in LDS by one workgroup, the panel below it solved against the inverse, the trailing lower
//      triangle updated by 64 x 64 tiles (rank-32 FP64 multiply-adds); the trailing R x R block is then the update
//      matrix the parent adds;
//   3. the solve panel Q = [L_CC^-1 ; -L_RC L_CC^-1] (m x C, column-major) written into the factor's panel buffer:
//      L_CC^-1 by block rows (X_ij = -L_ii^-1 sum_k L_ik X_kj with the diagonal blocks' inverses from step 2), and
//      -L_RC L_CC^-1 as one tiled product.
// Every front is column-major with leading dimension m, as the host's. A non-positive pivot raises a flag the host
// reads after the last front (GT_ERR_NUMERIC, as on the host).
#include <algorithm>
#include <map>

#include "stack.hpp"

namespace gt {

namespace {

constexpr int DB = 32;        // block columns of the factorisation and of L^-1
constexpr int DT = 64;        // output tile of the trailing update and of the panel product

// F[c m + r] += v for the front's original entries (distinct positions)
__global__ __launch_bounds__(256) void dn_scatter(double* __restrict__ F, int m, const int* __restrict__ r,
                                                  const int* __restrict__ c, const double* __restrict__ v, int n) {
    const int e = blockIdx.x * 256 + threadIdx.x;
    if (e < n) F[(size_t)c[e] * m + r[e]] += v[e];
}

// F[map[b] m + map[a]] += U[b ldU + a] for a >= b (the child's update matrix, lower triangle)
__global__ __launch_bounds__(256) void dn_extend(double* __restrict__ F, int m, const double* __restrict__ U, int ldU,
                                                 int r, const int* __restrict__ map) {
    const int a = blockIdx.x * 256 + threadIdx.x, b = blockIdx.y;
    if (a >= r || a < b) return;
    F[(size_t)map[b] * m + map[a]] += U[(size_t)b * ldU + a];
}

// Factor the diagonal block F[j0:j0+nb, j0:j0+nb] in place and write its inverse to dinv (DB x DB, column-major)
__global__ __launch_bounds__(256) void dn_potrf(double* __restrict__ F, int m, int j0, int nb,
                                                double* __restrict__ dinv, int* __restrict__ flag) {
    __shared__ double A[DB][DB + 1];
    __shared__ double X[DB][DB + 1];
    const int tid = threadIdx.x;
    for (int e = tid; e < DB * DB; e += 256) {
        const int i = e % DB, j = e / DB;
        A[i][j] = (i < nb && j < nb && i >= j) ? F[(size_t)(j0 + j) * m + j0 + i] : 0.0;
        X[i][j] = 0.0;
    }
    __syncthreads();
    for (int j = 0; j < nb; ++j) {
        if (tid == 0) {
            double d = A[j][j];
            if (!(d > 0.0) || !isfinite(d)) {
                flag[0] = 1;
                d = 1.0;
            }
            A[j][j] = sqrt(d);
        }
        __syncthreads();
        for (int i = j + 1 + tid; i < nb; i += 256) A[i][j] /= A[j][j];
        __syncthreads();
        const int w = nb - j - 1;
        for (int e = tid; e < w * w; e += 256) {
            const int i = j + 1 + e % w, k = j + 1 + e / w;
            if (i >= k) A[i][k] -= A[i][j] * A[k][j];
        }
        __syncthreads();
    }
    // inverse of the lower triangle, a column per thread, the column in registers (its entries above c are zero, so
    // each row's sum over k < i adds zeros before the sum over c <= k < i) and every A read a broadcast
    if (tid < nb) {
        const int c = tid;
        double x[DB];
#pragma unroll
        for (int i = 0; i < DB; ++i) {
            double s = 0.0;
#pragma unroll
            for (int k = 0; k < i; ++k) s += A[i][k] * x[k];
            x[i] = i < c || i >= nb ? 0.0 : i == c ? 1.0 / A[i][i] : -s / A[i][i];
        }
#pragma unroll
        for (int i = 0; i < DB; ++i) X[i][c] = x[i];
    }
    __syncthreads();
    for (int e = tid; e < DB * DB; e += 256) {
        const int i = e % DB, j = e / DB;
        dinv[(size_t)j * DB + i] = X[i][j];
        if (i < nb && j < nb && i >= j) F[(size_t)(j0 + j) * m + j0 + i] = A[i][j];
    }
}

// The panel below the diagonal block: L21 = F21 L11^-T, a row per thread
__global__ __launch_bounds__(256) void dn_trsm(double* __restrict__ F, int m, int j0, int nb,
                                               const double* __restrict__ dinv) {
    __shared__ double D[DB][DB + 1];          // D[t][u] = (L11^-1)[t][u]
    for (int e = threadIdx.x; e < DB * DB; e += 256) D[e % DB][e / DB] = dinv[e];
    __syncthreads();
    const int i = j0 + nb + blockIdx.x * 256 + threadIdx.x;
    if (i >= m) return;
    double f[DB];
#pragma unroll
    for (int u = 0; u < DB; ++u) f[u] = u < nb ? F[(size_t)(j0 + u) * m + i] : 0.0;
#pragma unroll
    for (int t = 0; t < DB; ++t) {
        if (t >= nb) break;
        double acc = 0.0;
#pragma unroll
        for (int u = 0; u <= t; ++u) acc += f[u] * D[t][u];
        F[(size_t)(j0 + t) * m + i] = acc;
    }
}

// The trailing lower triangle: F[k m + i] -= sum_t L21[i][t] L21[k][t] for b0 <= k <= i < m, b0 = j0 + nb, by
// 64 x 64 tiles (ti >= tk) of one 256-thread workgroup each, 4 x 4 outputs a thread
__global__ __launch_bounds__(256) void dn_syrk(double* __restrict__ F, int m, int j0, int nb) {
    __shared__ double Li[DT][DB + 1];
    __shared__ double Lk[DT][DB + 1];
    const int b0 = j0 + nb;
    const int l = blockIdx.x;
    int ti = (int)((sqrt(8.0 * l + 1.0) - 1.0) * 0.5);
    while ((ti + 1) * (ti + 2) / 2 <= l) ++ti;
    while (ti * (ti + 1) / 2 > l) --ti;
    const int tk = l - ti * (ti + 1) / 2;
    const int i0 = b0 + ti * DT, k0 = b0 + tk * DT;
    const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
    for (int e = tid; e < DT * DB; e += 256) {
        const int rr = e % DT, t = e / DT;
        const int i = i0 + rr, k = k0 + rr;
        Li[rr][t] = (t < nb && i < m) ? F[(size_t)(j0 + t) * m + i] : 0.0;
        Lk[rr][t] = (t < nb && k < m) ? F[(size_t)(j0 + t) * m + k] : 0.0;
    }
    __syncthreads();
    double acc[4][4];
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[r][c] = 0.0;
    for (int t = 0; t < nb; ++t) {
        double a[4], b[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) a[r] = Li[ty * 4 + r][t];
#pragma unroll
        for (int c = 0; c < 4; ++c) b[c] = Lk[tx * 4 + c][t];
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
            for (int c = 0; c < 4; ++c) acc[r][c] = fma(a[r], b[c], acc[r][c]);
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int i = i0 + ty * 4 + r;
        if (i >= m) continue;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const int k = k0 + tx * 4 + c;
            if (k < m && k <= i) F[(size_t)k * m + i] -= acc[r][c];
        }
    }
}

// Block row bi of X = L_CC^-1 (C x C, column-major, leading dimension C): block (bi, bj) for bj = blockIdx.x <= bi
__global__ __launch_bounds__(256) void dn_inv_row(const double* __restrict__ F, int m, int C, int bi,
                                                  const double* __restrict__ dinv_all, double* __restrict__ X) {
    __shared__ double Lb[DB][DB + 1];
    __shared__ double Xb[DB][DB + 1];
    __shared__ double T[DB][DB + 1];
    const int bj = blockIdx.x;
    const int tid = threadIdx.x;
    const int r0 = bi * DB, c0 = bj * DB;
    const double* D = dinv_all + (size_t)bi * DB * DB;      // (L_ii^-1)[t][u] at D[u DB + t]
    if (bj == bi) {
        for (int e = tid; e < DB * DB; e += 256) {
            const int i = e % DB, j = e / DB;
            if (r0 + i < C && c0 + j < C) X[(size_t)(c0 + j) * C + r0 + i] = D[(size_t)j * DB + i];
        }
        return;
    }
    // T = sum_{bk = bj}^{bi - 1} L[bi][bk] X[bk][bj]
    for (int e = tid; e < DB * DB; e += 256) T[e % DB][e / DB] = 0.0;
    __syncthreads();
    for (int bk = bj; bk < bi; ++bk) {
        const int k0 = bk * DB;
        for (int e = tid; e < DB * DB; e += 256) {
            const int i = e % DB, j = e / DB;
            Lb[i][j] = (r0 + i < C && k0 + j < C) ? F[(size_t)(k0 + j) * m + r0 + i] : 0.0;
            Xb[i][j] = (k0 + i < C && c0 + j < C) ? X[(size_t)(c0 + j) * C + k0 + i] : 0.0;
        }
        __syncthreads();
        for (int e = tid; e < DB * DB; e += 256) {
            const int i = e % DB, j = e / DB;
            double acc = 0.0;
            for (int t = 0; t < DB; ++t) acc += Lb[i][t] * Xb[t][j];
            T[i][j] += acc;
        }
        __syncthreads();
    }
    // X[bi][bj] = -L_ii^-1 T
    for (int e = tid; e < DB * DB; e += 256) {
        const int i = e % DB, j = e / DB;
        double acc = 0.0;
        for (int u = 0; u <= i; ++u) acc += D[(size_t)u * DB + i] * T[u][j];
        if (r0 + i < C && c0 + j < C) X[(size_t)(c0 + j) * C + r0 + i] = -acc;
    }
}

// The solve panel P (m x C, column-major) from X = L_CC^-1: rows [0, C) = X (zero above the diagonal), rows
// [C, m) = -L_RC X, by 64 x 64 tiles (rows a of the R part, columns j) with a 16-deep k loop from the tile's first
// column (X is zero above it)
__global__ __launch_bounds__(256) void dn_panel_rc(const double* __restrict__ F, int m, int C,
                                                   const double* __restrict__ X, double* __restrict__ P) {
    __shared__ double As[DT][17];
    __shared__ double Bs[16][DT + 1];
    const int R = m - C;
    const int a0 = blockIdx.y * DT, j0 = blockIdx.x * DT;
    const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
    double acc[4][4];
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[r][c] = 0.0;
    for (int k0 = j0; k0 < C; k0 += 16) {
        for (int e = tid; e < DT * 16; e += 256) {
            const int rr = e % DT, kk = e / DT;
            const int a = a0 + rr, k = k0 + kk;
            As[rr][kk] = (a < R && k < C) ? F[(size_t)k * m + C + a] : 0.0;
        }
        for (int e = tid; e < 16 * DT; e += 256) {
            const int kk = e % 16, cc = e / 16;
            const int k = k0 + kk, j = j0 + cc;
            Bs[kk][cc] = (k < C && j < C) ? X[(size_t)j * C + k] : 0.0;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < 16; ++kk) {
            double av[4], bv[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) av[r] = As[ty * 4 + r][kk];
#pragma unroll
            for (int c = 0; c < 4; ++c) bv[c] = Bs[kk][tx * 4 + c];
#pragma unroll
            for (int r = 0; r < 4; ++r)
#pragma unroll
                for (int c = 0; c < 4; ++c) acc[r][c] = fma(av[r], bv[c], acc[r][c]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int a = a0 + ty * 4 + r;
        if (a >= R) continue;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const int j = j0 + tx * 4 + c;
            if (j < C) P[(size_t)j * m + C + a] = -acc[r][c];
        }
    }
}

__global__ __launch_bounds__(256) void dn_panel_cc(const double* __restrict__ X, int C, int m, double* __restrict__ P) {
    const size_t e = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= (size_t)C * C) return;
    const int i = (int)(e % C), j = (int)(e / C);
    P[(size_t)j * m + i] = i >= j ? X[e] : 0.0;
}

}  // namespace

void Schur::factor_device(hipStream_t st) {
    SchurHost& H = h;
    if (H.dev_order.empty()) return;
    const int nd = (int)H.dev_order.size();
    const int ns = (int)H.nc.size();
    // Everything the fronts read from the host goes up once, before the first front: the original entries, each
    // device front's children's row positions in it, and the update matrices of its host children. The fronts then
    // run back to back on the stream, with no wait between them.
    std::vector<int64_t> map_at(ns, -1), u_at(ns, -1);
    std::vector<int> maps;
    int64_t utot = 0;
    int cmax = 1, bmax = 1;
    for (int di = 0; di < nd; ++di) {
        const int s = H.dev_order[di];
        const int C = H.nc[s], R = H.nr[s];
        cmax = std::max(cmax, C);
        bmax = std::max(bmax, (C + DB - 1) / DB);
        for (int e = H.child_ptr[s]; e < H.child_ptr[s + 1]; ++e) {
            const int k = H.child[e], r = H.nr[k];
            if (r == 0) continue;
            map_at[k] = (int64_t)maps.size();
            for (int a = 0; a < r; ++a) {
                const int c = H.rows[H.roff[k] + a];
                if (c < H.c0[s] + C) {
                    maps.push_back(c - H.c0[s]);
                } else {
                    const int* b0 = H.rows.data() + H.roff[s];
                    const int* f = std::lower_bound(b0, b0 + R, c);
                    if (f == b0 + R || *f != c) fail(GT_ERR_ARG, "a child's row is not in its parent's front");
                    maps.push_back(C + (int)(f - b0));
                }
            }
            if (!H.on_dev[k]) {
                if (H.upd_host[k].size() != (size_t)r * r) fail(GT_ERR_ARG, "a host child's update matrix is missing");
                u_at[k] = utot;
                utot += (int64_t)r * r;
            }
        }
    }
    const int64_t nt = H.trip_off[nd];
    DevBuf<int> flag(1), tr((size_t)std::max<int64_t>(1, nt)), tc((size_t)std::max<int64_t>(1, nt));
    DevBuf<double> tv((size_t)std::max<int64_t>(1, nt));
    DevBuf<int> mp(std::max<size_t>(1, maps.size()));
    DevBuf<double> Uh((size_t)std::max<int64_t>(1, utot));
    GT_HIP(hipMemsetAsync(flag.p, 0, sizeof(int), st));
    if (nt > 0) {
        tr.upload(H.trip_r.data(), (size_t)nt, st);
        tc.upload(H.trip_c.data(), (size_t)nt, st);
        tv.upload(H.trip_v.data(), (size_t)nt, st);
    }
    if (!maps.empty()) mp.upload(maps.data(), maps.size(), st);
    for (int k = 0; k < ns; ++k)
        if (u_at[k] >= 0)
            GT_HIP(hipMemcpyAsync(Uh.p + u_at[k], H.upd_host[k].data(), H.upd_host[k].size() * sizeof(double),
                                  hipMemcpyHostToDevice, st));
    GT_HIP(hipStreamSynchronize(st));
    for (int k = 0; k < ns; ++k)
        if (u_at[k] >= 0) std::vector<double>().swap(H.upd_host[k]);
    // per-front scratch, sized for the largest front and reused in stream order
    DevBuf<double> dinv((size_t)bmax * DB * DB), X((size_t)cmax * cmax);
    // fronts waiting for their parent, and the buffers of consumed ones kept for reuse (stream order makes a
    // buffer's next front start after its last reader)
    std::map<int, DevBuf<double>> fronts;
    std::multimap<size_t, DevBuf<double>> pool;
    auto take = [&](size_t n) {
        auto it = pool.lower_bound(n);
        if (it != pool.end() && it->first <= 2 * n + (size_t(1) << 20)) {
            DevBuf<double> b = std::move(it->second);
            pool.erase(it);
            return b;
        }
        return DevBuf<double>(n);
    };
    for (int di = 0; di < nd; ++di) {
        const int s = H.dev_order[di];
        const int C = H.nc[s], R = H.nr[s], m = C + R;
        DevBuf<double> F = take((size_t)m * m);
        GT_HIP(hipMemsetAsync(F.p, 0, (size_t)m * m * sizeof(double), st));
        // 1. original entries, then the children's update matrices at their rows' positions in this front
        {
            const int64_t e0 = H.trip_off[di], e1 = H.trip_off[di + 1];
            const int n = (int)(e1 - e0);
            if (n > 0) {
                hipLaunchKernelGGL(dn_scatter, dim3(blocks_for(n, 256)), dim3(256), 0, st, F.p, m, tr.p + e0,
                                   tc.p + e0, tv.p + e0, n);
                GT_LAUNCH("dn_scatter");
            }
        }
        for (int e = H.child_ptr[s]; e < H.child_ptr[s + 1]; ++e) {
            const int k = H.child[e], r = H.nr[k];
            if (r == 0) continue;
            const int* map = mp.p + map_at[k];
            if (H.on_dev[k]) {
                auto it = fronts.find(k);
                if (it == fronts.end()) fail(GT_ERR_ARG, "a device child's front is missing");
                const int mk = H.nc[k] + r;
                const double* U = it->second.p + (size_t)H.nc[k] * mk + H.nc[k];
                hipLaunchKernelGGL(dn_extend, dim3(blocks_for(r, 256), r), dim3(256), 0, st, F.p, m, U, mk, r, map);
                GT_LAUNCH("dn_extend");
                DevBuf<double> done = std::move(it->second);
                fronts.erase(it);
                pool.emplace(done.n, std::move(done));
            } else {
                hipLaunchKernelGGL(dn_extend, dim3(blocks_for(r, 256), r), dim3(256), 0, st, F.p, m, Uh.p + u_at[k], r,
                                   r, map);
                GT_LAUNCH("dn_extend");
            }
        }
        // 2. partial Cholesky of the first C columns
        const int nblk = (C + DB - 1) / DB;
        for (int b = 0; b < nblk; ++b) {
            const int j0 = b * DB, nb = std::min(DB, C - j0);
            hipLaunchKernelGGL(dn_potrf, dim3(1), dim3(256), 0, st, F.p, m, j0, nb, dinv.p + (size_t)b * DB * DB,
                               flag.p);
            GT_LAUNCH("dn_potrf");
            const int below = m - (j0 + nb);
            if (below > 0) {
                hipLaunchKernelGGL(dn_trsm, dim3(blocks_for(below, 256)), dim3(256), 0, st, F.p, m, j0, nb,
                                   dinv.p + (size_t)b * DB * DB);
                GT_LAUNCH("dn_trsm");
                const int ntl = (below + DT - 1) / DT;
                hipLaunchKernelGGL(dn_syrk, dim3(ntl * (ntl + 1) / 2), dim3(256), 0, st, F.p, m, j0, nb);
                GT_LAUNCH("dn_syrk");
            }
        }
        // 3. the solve panel into the factor's panel buffer
        // X above its diagonal blocks stays zero: the panel product's tiles read it from their first column on
        GT_HIP(hipMemsetAsync(X.p, 0, (size_t)C * C * sizeof(double), st));
        for (int bi = 0; bi < nblk; ++bi) {
            hipLaunchKernelGGL(dn_inv_row, dim3(bi + 1), dim3(256), 0, st, F.p, m, C, bi, dinv.p, X.p);
            GT_LAUNCH("dn_inv_row");
        }
        double* P = panel.p + H.poff[s];
        hipLaunchKernelGGL(dn_panel_cc, dim3(blocks_for((size_t)C * C, 256)), dim3(256), 0, st, X.p, C, m, P);
        GT_LAUNCH("dn_panel_cc");
        if (R > 0) {
            hipLaunchKernelGGL(dn_panel_rc, dim3((C + DT - 1) / DT, (R + DT - 1) / DT), dim3(256), 0, st, F.p, m, C,
                               X.p, P);
            GT_LAUNCH("dn_panel_rc");
        }
        if (H.parent[s] >= 0 && R > 0) fronts.emplace(s, std::move(F));
        else pool.emplace(F.n, std::move(F));
    }
    GT_HIP(hipStreamSynchronize(st));
    int bad = 0;
    GT_HIP(hipMemcpy(&bad, flag.p, sizeof(int), hipMemcpyDeviceToHost));
    if (bad) fail(GT_ERR_NUMERIC, "the Schur complement is not positive definite (a piece without a reference node?)");
}

void Schur::fetch_panels(hipStream_t st) {
    if (h.host_panels) return;
    for (int s : h.dev_order) {
        const size_t cnt = (size_t)(h.nc[s] + h.nr[s]) * h.nc[s];
        GT_HIP(hipMemcpyAsync(h.panel.data() + h.poff[s], panel.p + h.poff[s], cnt * sizeof(double),
                              hipMemcpyDeviceToHost, st));
    }
    GT_HIP(hipStreamSynchronize(st));
    h.host_panels = true;
}

}  // namespace gt
