// GMRES vector work on the device for a batch of right-hand sides (blockIdx.y = right-hand side). Reductions are
// two-stage and in a fixed order (block partial sums, then one tree per value), so a solve gives the same bits on
// every run and whatever else is in its batch.
#include "krylov.hpp"

namespace gt {

namespace {

constexpr int NTH = 256;

// a stored value as double
__device__ __forceinline__ zd ldz(const zd& v) { return v; }
__device__ __forceinline__ zd ldz(const cpx<float>& v) { return zd(v.re, v.im); }
template <class T>
__device__ __forceinline__ cpx<T> stz(double re, double im) { return cpx<T>((T)re, (T)im); }

__device__ __forceinline__ double wave_sum(double v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

// Sum of `val` over the workgroup; the result is valid in thread 0.
__device__ __forceinline__ double block_sum(double val, double* lds) {
    val = wave_sum(val);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) lds[wave] = val;
    __syncthreads();
    double r = 0;
    if (threadIdx.x == 0)
        for (int w = 0; w < NTH / 64; ++w) r += lds[w];
    return r;
}

// Basis values a lane loads at once: one complex double (16 bytes), or two consecutive complex floats (one 16-byte
// load). The basis rows have an even stride (a multiple of 64 elements), so a pair starting at an even element is
// aligned; the pair past the end of an odd-length row reads the row's padding, which is replaced by zero.
template <class TV>
constexpr int lane_elems() { return sizeof(TV) == 4 ? 2 : 1; }

template <class TV>
__device__ __forceinline__ void ld_basis(const cpx<TV>* p, bool ok, bool ok2, zd (&v)[lane_elems<TV>()]);
template <>
__device__ __forceinline__ void ld_basis<double>(const zd* p, bool ok, bool, zd (&v)[1]) {
    v[0] = ok ? *p : zd();
}
template <>
__device__ __forceinline__ void ld_basis<float>(const cpx<float>* p, bool ok, bool ok2, zd (&v)[2]) {
    float4 f = ok ? *reinterpret_cast<const float4*>(p) : float4{0.f, 0.f, 0.f, 0.f};
    v[0] = zd(f.x, f.y);
    v[1] = ok2 ? zd(f.z, f.w) : zd();
}

// Partial sums of h = V[0:m]^H w (and |w|^2) over one block's chunk of elements and one group of DG basis rows
// (grid: chunks x row groups, right-hand sides): each lane loads its elements of the group's rows for DU element
// steps at once, the sums meet by shuffles in each wave and across the four waves in LDS, in a fixed order.
template <class TV, class TW>
__global__ __launch_bounds__(NTH) void kr_dots(KrSetT<TV, TW> k, int n, int chunk, double* part, int P, int nblk,
                                               int wslot) {
    constexpr int E = lane_elems<TV>();
    constexpr int DG = 4;
    constexpr int DU = E == 1 ? 4 : 1;
    __shared__ double red[NTH / 64][2 * DG + 1];
    const int p = blockIdx.y;
    const int grp = blockIdx.x / nblk, blk = blockIdx.x - grp * nblk;
    const int m = k.m[p];
    const int g0 = grp * DG;
    if (g0 >= m) return;                        // the whole block: this right-hand side has fewer rows
    const int gm = min(DG, m - g0);
    const cpx<TV>* __restrict__ V = k.V[p] + (size_t)g0 * k.ld;
    const cpx<TW>* __restrict__ w = k.w[p];
    if (k.n[p] > 0) n = k.n[p];
    const int i0 = blk * chunk, i1 = min(n, i0 + chunk);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    double ar[DG], ai[DG];
#pragma unroll
    for (int r = 0; r < DG; ++r) ar[r] = ai[r] = 0;
    double ww = 0;
    for (int ib = i0 + E * threadIdx.x; ib < i1; ib += E * NTH * DU) {
        zd wv[DU][E];
        zd vv[DU][DG][E];
#pragma unroll
        for (int u = 0; u < DU; ++u) {
            const int i = ib + u * E * NTH;
            const bool in = i < i1, in2 = i + 1 < i1;
#pragma unroll
            for (int c = 0; c < E; ++c) wv[u][c] = (c == 0 ? in : in2) ? ldz(w[i + c]) : zd();
#pragma unroll
            for (int r = 0; r < DG; ++r) ld_basis<TV>(V + (size_t)r * k.ld + i, in && r < gm, in2, vv[u][r]);
        }
#pragma unroll
        for (int u = 0; u < DU; ++u)
#pragma unroll
            for (int c = 0; c < E; ++c) {
                ww += wv[u][c].re * wv[u][c].re + wv[u][c].im * wv[u][c].im;
#pragma unroll
                for (int r = 0; r < DG; ++r) {
                    ar[r] += vv[u][r][c].re * wv[u][c].re + vv[u][r][c].im * wv[u][c].im;
                    ai[r] += vv[u][r][c].re * wv[u][c].im - vv[u][r][c].im * wv[u][c].re;
                }
            }
    }
#pragma unroll
    for (int r = 0; r < DG; ++r) {
        ar[r] = wave_sum(ar[r]);
        ai[r] = wave_sum(ai[r]);
    }
    ww = wave_sum(ww);
    if (lane == 0) {
#pragma unroll
        for (int r = 0; r < DG; ++r) {
            red[wave][2 * r] = ar[r];
            red[wave][2 * r + 1] = ai[r];
        }
        red[wave][2 * DG] = ww;
    }
    __syncthreads();
    double* out = part + ((size_t)p * nblk + blk) * P;
    const int t = threadIdx.x;
    if (t < 2 * gm) out[2 * g0 + t] = red[0][t] + red[1][t] + red[2][t] + red[3][t];
    if (grp == 0 && t == 2 * DG) out[wslot] = red[0][t] + red[1][t] + red[2][t] + red[3][t];
}

// The update w' = w - V h fused with the products V^H w' (the second Gram-Schmidt pass's coefficients). Lane =
// (element slot e = lane % EL, row group g = lane / EL): a lane holds rows g, g + NG, .. (at most RPL) of its E
// elements' basis columns, so each basis value is loaded once for both; the row groups' parts of V h meet by
// butterfly sums, and the products are summed over the block's elements per row. Launched once per class of m
// (which fixes EL and RPL), its blocks of right-hand sides outside the class return at once: a right-hand side's
// arithmetic depends on its own m only.
template <class TV, class TW, int EL, int RPL>
__global__ __launch_bounds__(NTH) void kr_update_dots(KrSetT<TV, TW> k, int n, int chunk, const double* red, double* part,
                                                      int P, int nblk, int slot_after, int slot_h2, int m_lo) {
    constexpr int E = lane_elems<TV>();
    constexpr int NG = 64 / EL;
    constexpr int R = NG * RPL;
    __shared__ double hs_raw[2 * R];
    __shared__ double wsum[NTH / 64][2 * R + 1];
    const int p = blockIdx.y;
    const int m = k.m[p];
    if (m <= m_lo || m > R) return;
    zd* hs = reinterpret_cast<zd*>(hs_raw);
    const cpx<TV>* __restrict__ V = k.V[p];
    cpx<TW>* w = k.w[p];
    const double* h = red + (size_t)p * P;
    if (k.n[p] > 0) n = k.n[p];
    for (int r = threadIdx.x; r < R; r += NTH) hs[r] = r < m ? zd(h[2 * r], h[2 * r + 1]) : zd();
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, e = lane % EL, g = lane / EL;
    zd hr[RPL];
#pragma unroll
    for (int t = 0; t < RPL; ++t) hr[t] = hs[g + NG * t];
    double dre[RPL], dim[RPL];
#pragma unroll
    for (int t = 0; t < RPL; ++t) dre[t] = dim[t] = 0.0;
    double nn = 0.0;
    const int i0 = blockIdx.x * chunk, i1 = min(n, i0 + chunk);
    for (int ib = i0; ib < i1; ib += 4 * EL * E) {
        const int i = ib + E * (EL * wave + e);
        const bool in = i < i1, in2 = i + 1 < i1;
        zd vv[RPL][E];
#pragma unroll
        for (int t = 0; t < RPL; ++t) {
            const int r = g + NG * t;
            ld_basis<TV>(V + (size_t)r * k.ld + i, in && r < m, in2, vv[t]);
        }
        zd wi[E];
#pragma unroll
        for (int c = 0; c < E; ++c) {
            const zd w0 = (c == 0 ? in : in2) ? ldz(w[i + c]) : zd();
            double sre = 0.0, sim = 0.0;
#pragma unroll
            for (int t = 0; t < RPL; ++t) {
                sre += vv[t][c].re * hr[t].re - vv[t][c].im * hr[t].im;
                sim += vv[t][c].re * hr[t].im + vv[t][c].im * hr[t].re;
            }
#pragma unroll
            for (int off = EL; off < 64; off <<= 1) {
                sre += __shfl_xor(sre, off);
                sim += __shfl_xor(sim, off);
            }
            // the stored value (rounded to TW) is what the products and the norm see
            const cpx<TW> ws = stz<TW>(w0.re - sre, w0.im - sim);
            wi[c] = ldz(ws);
            if ((c == 0 ? in : in2) && g == 0) {
                w[i + c] = ws;
                nn += wi[c].re * wi[c].re + wi[c].im * wi[c].im;
            }
        }
#pragma unroll
        for (int t = 0; t < RPL; ++t)
#pragma unroll
            for (int c = 0; c < E; ++c) {
                dre[t] += vv[t][c].re * wi[c].re + vv[t][c].im * wi[c].im;
                dim[t] += vv[t][c].re * wi[c].im - vv[t][c].im * wi[c].re;
            }
    }
    // the EL element lanes of each row group, then the four waves in order
#pragma unroll
    for (int t = 0; t < RPL; ++t) {
#pragma unroll
        for (int off = 1; off < EL; off <<= 1) {
            dre[t] += __shfl_xor(dre[t], off);
            dim[t] += __shfl_xor(dim[t], off);
        }
    }
    nn = wave_sum(nn);
    if (e == 0) {
#pragma unroll
        for (int t = 0; t < RPL; ++t) {
            wsum[wave][2 * (g + NG * t)] = dre[t];
            wsum[wave][2 * (g + NG * t) + 1] = dim[t];
        }
    }
    if (lane == 0) wsum[wave][2 * R] = nn;
    __syncthreads();
    double* out = part + ((size_t)p * nblk + blockIdx.x) * P;
    for (int c = threadIdx.x; c <= 2 * R; c += NTH) {
        const double v = wsum[0][c] + wsum[1][c] + wsum[2][c] + wsum[3][c];
        if (c == 2 * R) out[slot_after] = v;
        else if (c < 2 * m) out[slot_h2 + c] = v;
    }
}

// w -= V h, |w|^2 after. With test_w >= 0 (the re-pass), only where sqrt(red[test_after]) < 0.7 sqrt(red[test_w]):
// block 0 records the decision at red[flag], and a right-hand side without the re-pass leaves its vector alone. A
// lane takes E elements for UU element steps; the basis rows come UR at a time with their loads issued together,
// and each element's sum runs over the rows in order.
template <class TV, class TW>
__global__ __launch_bounds__(NTH) void kr_update(KrSetT<TV, TW> k, int n, int chunk, const double* red, double* part,
                                                 int P, int nblk, int slot, double* red_out, int test_w,
                                                 int test_after, int flag) {
    constexpr int E = lane_elems<TV>();
    constexpr int UU = E == 1 ? 4 : 1;
    constexpr int UR = E == 1 ? 4 : 8;
    __shared__ double lds[NTH / 64];
    __shared__ double hs_raw[1024];
    zd* hs = reinterpret_cast<zd*>(hs_raw);
    const int p = blockIdx.y;
    if (test_w >= 0) {
        const double* rp = red_out + (size_t)p * P;
        const bool take = sqrt(rp[test_after]) < 0.7 * sqrt(rp[test_w]);
        if (blockIdx.x == 0 && threadIdx.x == 0) red_out[(size_t)p * P + flag] = take ? 1.0 : 0.0;
        if (!take) return;
    }
    const cpx<TV>* __restrict__ V = k.V[p];
    cpx<TW>* w = k.w[p];
    const int m = k.m[p];
    const double* h = red + (size_t)p * P;
    if (k.n[p] > 0) n = k.n[p];
    for (int r = threadIdx.x; r < m; r += NTH) hs[r] = zd(h[2 * r], h[2 * r + 1]);
    __syncthreads();
    const int i0 = blockIdx.x * chunk, i1 = min(n, i0 + chunk);
    double nn = 0;
    for (int ib = i0 + E * threadIdx.x; ib < i1; ib += E * NTH * UU) {
        zd acc[UU][E];
#pragma unroll
        for (int u = 0; u < UU; ++u)
#pragma unroll
            for (int c = 0; c < E; ++c) {
                const int i = ib + u * E * NTH + c;
                acc[u][c] = i < i1 ? ldz(w[i]) : zd();
            }
        for (int r0 = 0; r0 < m; r0 += UR) {
            zd vv[UU][UR][E];
#pragma unroll
            for (int u = 0; u < UU; ++u) {
                const int i = ib + u * E * NTH;
#pragma unroll
                for (int r = 0; r < UR; ++r)
                    ld_basis<TV>(V + (size_t)(r0 + r) * k.ld + i, i < i1 && r0 + r < m, i + 1 < i1, vv[u][r]);
            }
#pragma unroll
            for (int r = 0; r < UR; ++r) {
                const zd hr = r0 + r < m ? hs[r0 + r] : zd();
#pragma unroll
                for (int u = 0; u < UU; ++u)
#pragma unroll
                    for (int c = 0; c < E; ++c) {
                        acc[u][c].re -= vv[u][r][c].re * hr.re - vv[u][r][c].im * hr.im;
                        acc[u][c].im -= vv[u][r][c].re * hr.im + vv[u][r][c].im * hr.re;
                    }
            }
        }
#pragma unroll
        for (int u = 0; u < UU; ++u)
#pragma unroll
            for (int c = 0; c < E; ++c) {
                const int i = ib + u * E * NTH + c;
                if (i < i1) {
                    const cpx<TW> ws = stz<TW>(acc[u][c].re, acc[u][c].im);
                    w[i] = ws;
                    const zd a = ldz(ws);
                    nn += a.re * a.re + a.im * a.im;
                }
            }
    }
    const double s = block_sum(nn, lds);
    if (threadIdx.x == 0) part[((size_t)p * nblk + blockIdx.x) * P + slot] = s;
}


// ---- delayed classical Gram-Schmidt with re-orthogonalisation (DCGS2): per Arnoldi step one multi-dot pass and one
// update pass over the basis. Row m of a basis holds u, the once-projected and normalised vector the product was
// taken of (v = A u); rows 0..m-1 are final.

// Pass A for long FP64 vectors: t = Q^H u and p = Q^H v over the rows [0, m) in groups of DG2 rows, and u^H u,
// u^H v (with group 0).
// v may be absent (a basis's last row: t and u^H u only). A workgroup takes one chunk of elements and gpw
// consecutive groups: its u and v elements are loaded once into registers and serve every group it takes (long
// vectors take every group in one workgroup, so u and v are read once per pass; short ones one group per workgroup,
// for enough workgroups). Each (row, chunk) sum is the same whatever gpw is: the same elements per lane in the same
// order, then the same wave and workgroup trees.
template <class TV, class TW>
__global__ __launch_bounds__(NTH) void kr_dots2_long(KrSetT<TV, TW> k, int n, int chunk, double* part, int P,
                                                     int nblk, int slot_t, int slot_uu, int gpw) {
    constexpr int E = lane_elems<TV>();
    constexpr int DG2 = 8;
    constexpr int ST = 4;                       // element steps per lane: chunk = ST * E * NTH
    __shared__ double red[NTH / 64][4 * DG2 + 3];
    const int p = blockIdx.y;
    const int gs = blockIdx.x / nblk, blk = blockIdx.x - gs * nblk;
    const int m = k.m[p];
    const int ga = gs * gpw;
    if (m < 0 || (ga > 0 && ga * DG2 >= m)) return;     // m < 0: a slot not in this step
    const int gb = max(1, min(ga + gpw, (m + DG2 - 1) / DG2));
    const cpx<TV>* __restrict__ u = k.V[p] + (size_t)m * k.ld;
    const cpx<TW>* __restrict__ v = k.w[p];
    const bool hv = v != nullptr;
    if (k.n[p] > 0) n = k.n[p];
    const int i0 = blk * chunk, i1 = min(n, i0 + chunk);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    zd uv_[ST][E], vv_[ST][E];
#pragma unroll
    for (int st = 0; st < ST; ++st) {
        const int i = i0 + E * threadIdx.x + st * E * NTH;
        const bool in1 = i < i1, in2 = i + 1 < i1;
        if (in1) ld_basis<TV>(u + i, true, in2, uv_[st]);
#pragma unroll
        for (int c = 0; c < E; ++c) {
            if (!in1) uv_[st][c] = zd();
            vv_[st][c] = in1 && hv && (c == 0 || in2) ? ldz(v[i + c]) : zd();
        }
    }
    for (int grp = ga; grp < gb; ++grp) {
        const int g0 = grp * DG2;
        const int gm = max(0, min(DG2, m - g0));
        const cpx<TV>* __restrict__ V = k.V[p] + (size_t)g0 * k.ld;
        double tr[DG2], ti[DG2], pr[DG2], pi[DG2];
#pragma unroll
        for (int r = 0; r < DG2; ++r) tr[r] = ti[r] = pr[r] = pi[r] = 0;
        double uu = 0, uvr = 0, uvi = 0;
#pragma unroll
        for (int st = 0; st < ST; ++st) {
            const int i = i0 + E * threadIdx.x + st * E * NTH;
            if (i >= i1) break;
            const bool in2 = i + 1 < i1;
            zd q[DG2][E];
#pragma unroll
            for (int r = 0; r < DG2; ++r) ld_basis<TV>(V + (size_t)r * k.ld + i, r < gm, in2, q[r]);
#pragma unroll
            for (int c = 0; c < E; ++c) {
                const zd uc = uv_[st][c], vc = vv_[st][c];
                uu += uc.re * uc.re + uc.im * uc.im;
                uvr += uc.re * vc.re + uc.im * vc.im;
                uvi += uc.re * vc.im - uc.im * vc.re;
#pragma unroll
                for (int r = 0; r < DG2; ++r) {
                    tr[r] += q[r][c].re * uc.re + q[r][c].im * uc.im;
                    ti[r] += q[r][c].re * uc.im - q[r][c].im * uc.re;
                    pr[r] += q[r][c].re * vc.re + q[r][c].im * vc.im;
                    pi[r] += q[r][c].re * vc.im - q[r][c].im * vc.re;
                }
            }
        }
#pragma unroll
        for (int r = 0; r < DG2; ++r) {
            tr[r] = wave_sum(tr[r]);
            ti[r] = wave_sum(ti[r]);
            pr[r] = wave_sum(pr[r]);
            pi[r] = wave_sum(pi[r]);
        }
        uu = wave_sum(uu);
        uvr = wave_sum(uvr);
        uvi = wave_sum(uvi);
        if (grp > ga) __syncthreads();              // the previous group's sums are read
        if (lane == 0) {
#pragma unroll
            for (int r = 0; r < DG2; ++r) {
                red[wave][2 * r] = tr[r];
                red[wave][2 * r + 1] = ti[r];
                red[wave][2 * DG2 + 2 * r] = pr[r];
                red[wave][2 * DG2 + 2 * r + 1] = pi[r];
            }
            red[wave][4 * DG2] = uu;
            red[wave][4 * DG2 + 1] = uvr;
            red[wave][4 * DG2 + 2] = uvi;
        }
        __syncthreads();
        double* out = part + ((size_t)p * nblk + blk) * P;
        const int t = threadIdx.x;
        const auto sum = [&](int c) { return red[0][c] + red[1][c] + red[2][c] + red[3][c]; };
        if (t < 2 * gm) out[slot_t + 2 * g0 + t] = sum(t);
        else if (t >= 64 && t < 64 + 2 * gm) out[2 * g0 + t - 64] = sum(2 * DG2 + t - 64);
        else if (grp == 0 && t >= 128 && t < 131) out[slot_uu + t - 128] = sum(4 * DG2 + t - 128);
    }
}

// basis rows per workgroup of the grouped multi-dot: 8 over an FP64 basis, 4 over an FP32 one (whose lanes hold two
// elements each: half the accumulators keep its occupancy)
template <class TV>
constexpr int dots2_group() { return sizeof(TV) == 4 ? 4 : 8; }

// Pass A (short vectors, and FP32 bases): t = Q^H u and p = Q^H v over the rows [0, m) in groups of DG2 rows per
// workgroup, and u^H u, u^H v (group 0).
// v may be absent (a basis's last row: t and u^H u only).
template <class TV, class TW>
__global__ __launch_bounds__(NTH) void kr_dots2(KrSetT<TV, TW> k, int n, int chunk, double* part, int P, int nblk,
                                                int slot_t, int slot_uu) {
    constexpr int E = lane_elems<TV>();
    constexpr int DG2 = dots2_group<TV>();
    __shared__ double red[NTH / 64][4 * DG2 + 3];
    const int p = blockIdx.y;
    const int grp = blockIdx.x / nblk, blk = blockIdx.x - grp * nblk;
    const int m = k.m[p];
    const int g0 = grp * DG2;
    if (m < 0 || (g0 > 0 && g0 >= m)) return;     // m < 0: a slot not in this step
    const int gm = max(0, min(DG2, m - g0));
    const cpx<TV>* __restrict__ V = k.V[p] + (size_t)g0 * k.ld;
    const cpx<TV>* __restrict__ u = k.V[p] + (size_t)m * k.ld;
    const cpx<TW>* __restrict__ v = k.w[p];
    const bool hv = v != nullptr;
    if (k.n[p] > 0) n = k.n[p];
    const int i0 = blk * chunk, i1 = min(n, i0 + chunk);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    double tr[DG2], ti[DG2], pr[DG2], pi[DG2];
#pragma unroll
    for (int r = 0; r < DG2; ++r) tr[r] = ti[r] = pr[r] = pi[r] = 0;
    double uu = 0, uvr = 0, uvi = 0;
    for (int i = i0 + E * threadIdx.x; i < i1; i += E * NTH) {
        const bool in2 = i + 1 < i1;
        zd uv_[E], vv_[E];
        ld_basis<TV>(u + i, true, in2, uv_);
#pragma unroll
        for (int c = 0; c < E; ++c) vv_[c] = hv && (c == 0 || in2) ? ldz(v[i + c]) : zd();
        zd q[DG2][E];
#pragma unroll
        for (int r = 0; r < DG2; ++r) ld_basis<TV>(V + (size_t)r * k.ld + i, r < gm, in2, q[r]);
#pragma unroll
        for (int c = 0; c < E; ++c) {
            uu += uv_[c].re * uv_[c].re + uv_[c].im * uv_[c].im;
            uvr += uv_[c].re * vv_[c].re + uv_[c].im * vv_[c].im;
            uvi += uv_[c].re * vv_[c].im - uv_[c].im * vv_[c].re;
#pragma unroll
            for (int r = 0; r < DG2; ++r) {
                tr[r] += q[r][c].re * uv_[c].re + q[r][c].im * uv_[c].im;
                ti[r] += q[r][c].re * uv_[c].im - q[r][c].im * uv_[c].re;
                pr[r] += q[r][c].re * vv_[c].re + q[r][c].im * vv_[c].im;
                pi[r] += q[r][c].re * vv_[c].im - q[r][c].im * vv_[c].re;
            }
        }
    }
#pragma unroll
    for (int r = 0; r < DG2; ++r) {
        tr[r] = wave_sum(tr[r]);
        ti[r] = wave_sum(ti[r]);
        pr[r] = wave_sum(pr[r]);
        pi[r] = wave_sum(pi[r]);
    }
    uu = wave_sum(uu);
    uvr = wave_sum(uvr);
    uvi = wave_sum(uvi);
    if (lane == 0) {
#pragma unroll
        for (int r = 0; r < DG2; ++r) {
            red[wave][2 * r] = tr[r];
            red[wave][2 * r + 1] = ti[r];
            red[wave][2 * DG2 + 2 * r] = pr[r];
            red[wave][2 * DG2 + 2 * r + 1] = pi[r];
        }
        red[wave][4 * DG2] = uu;
        red[wave][4 * DG2 + 1] = uvr;
        red[wave][4 * DG2 + 2] = uvi;
    }
    __syncthreads();
    double* out = part + ((size_t)p * nblk + blk) * P;
    const int t = threadIdx.x;
    const auto sum = [&](int c) { return red[0][c] + red[1][c] + red[2][c] + red[3][c]; };
    if (t < 2 * gm) out[slot_t + 2 * g0 + t] = sum(t);
    else if (t >= 64 && t < 64 + 2 * gm) out[2 * g0 + t - 64] = sum(2 * DG2 + t - 64);
    else if (grp == 0 && t >= 128 && t < 131) out[slot_uu + t - 128] = sum(4 * DG2 + t - 128);
}

// Pass B: q = (u - Q t) / alpha written over u (row m), and y = v - Q p - q (q^H v) written over v, |y|^2 after; one
// read of the rows [0, m). t and p as reduced by pass A; 1 / alpha and q^H v per right-hand side.
struct DcgsCoef {
    double ainv[KMAX];
    zd qv[KMAX];
};

template <class TV, class TW>
__global__ __launch_bounds__(NTH) void kr_update2(KrSetT<TV, TW> k, int n, int chunk, const double* red, double* part,
                                                  int P, int nblk, int slot_t, int slot_y, DcgsCoef cf) {
    constexpr int E = lane_elems<TV>();
    constexpr int UR = 4;
    __shared__ double lds[NTH / 64];
    __shared__ double tp_raw[2048];
    zd* ts = reinterpret_cast<zd*>(tp_raw);
    zd* ps = ts + 512;
    const int p = blockIdx.y;
    const int m = k.m[p];
    if (m < 0) return;                              // a slot not in this step
    cpx<TV>* u = const_cast<cpx<TV>*>(k.V[p]) + (size_t)m * k.ld;
    const cpx<TV>* __restrict__ V = k.V[p];
    cpx<TW>* w = k.w[p];
    const double* rp = red + (size_t)p * P;
    for (int r = threadIdx.x; r < m; r += NTH) {
        ts[r] = zd(rp[slot_t + 2 * r], rp[slot_t + 2 * r + 1]);
        ps[r] = zd(rp[2 * r], rp[2 * r + 1]);
    }
    __syncthreads();
    if (k.n[p] > 0) n = k.n[p];
    const double ainv = cf.ainv[p];
    const zd qv = cf.qv[p];
    const int i0 = blockIdx.x * chunk, i1 = min(n, i0 + chunk);
    double nn = 0;
    for (int i = i0 + E * threadIdx.x; i < i1; i += E * NTH) {
        const bool in2 = i + 1 < i1;
        zd au[E], av[E];
        ld_basis<TV>(u + i, true, in2, au);
#pragma unroll
        for (int c = 0; c < E; ++c) av[c] = c == 0 || in2 ? ldz(w[i + c]) : zd();
        for (int r0 = 0; r0 < m; r0 += UR) {
            zd vv[UR][E];
#pragma unroll
            for (int r = 0; r < UR; ++r) ld_basis<TV>(V + (size_t)(r0 + r) * k.ld + i, r0 + r < m, in2, vv[r]);
#pragma unroll
            for (int r = 0; r < UR; ++r) {
                const zd tr = r0 + r < m ? ts[r0 + r] : zd(), pr = r0 + r < m ? ps[r0 + r] : zd();
#pragma unroll
                for (int c = 0; c < E; ++c) {
                    au[c].re -= vv[r][c].re * tr.re - vv[r][c].im * tr.im;
                    au[c].im -= vv[r][c].re * tr.im + vv[r][c].im * tr.re;
                    av[c].re -= vv[r][c].re * pr.re - vv[r][c].im * pr.im;
                    av[c].im -= vv[r][c].re * pr.im + vv[r][c].im * pr.re;
                }
            }
        }
        zd qs[E];
#pragma unroll
        for (int c = 0; c < E; ++c) {
            const cpx<TV> qst = stz<TV>(au[c].re * ainv, au[c].im * ainv);
            qs[c] = ldz(qst);
            if (c == 0 || in2) u[i + c] = qst;
        }
#pragma unroll
        for (int c = 0; c < E; ++c) {
            if (!(c == 0 || in2)) continue;
            const cpx<TW> ys = stz<TW>(av[c].re - (qs[c].re * qv.re - qs[c].im * qv.im),
                                       av[c].im - (qs[c].re * qv.im + qs[c].im * qv.re));
            w[i + c] = ys;
            const zd y = ldz(ys);
            nn += y.re * y.re + y.im * y.im;
        }
    }
    const double sb = block_sum(nn, lds);
    if (threadIdx.x == 0) part[((size_t)p * nblk + blockIdx.x) * P + slot_y] = sb;
}

// dst_p = a_p / sqrt(red[p][slot]) (the next basis row, scaled by the norm pass B left on the device)
template <class T, class TA>
__global__ __launch_bounds__(NTH) void kr_scale_red(VecOpsT<T, TA> v, int n, const double* red, int P, int slot) {
    const int i = blockIdx.x * NTH + threadIdx.x, p = blockIdx.y;
    if (v.dst[p] && i < (v.n[p] > 0 ? v.n[p] : n)) {
        const double f = 1.0 / sqrt(red[(size_t)p * P + slot]);
        const zd a = ldz(v.a[p][i]);
        v.dst[p][i] = stz<T>(f * a.re, f * a.im);
    }
}

template <class TV, class TW>
__global__ __launch_bounds__(NTH) void kr_norm(KrSetT<TV, TW> k, int n, int chunk, double* part, int P, int nblk,
                                               int slot) {
    __shared__ double lds[NTH / 64];
    const int p = blockIdx.y;
    const cpx<TW>* __restrict__ v = k.w[p];
    if (k.n[p] > 0) n = k.n[p];
    const int i0 = blockIdx.x * chunk, i1 = min(n, i0 + chunk);
    double nn = 0;
    for (int i = i0 + threadIdx.x; i < i1; i += NTH) {
        const zd a = ldz(v[i]);
        nn += a.re * a.re + a.im * a.im;
    }
    const double s = block_sum(nn, lds);
    if (threadIdx.x == 0) part[((size_t)p * nblk + blockIdx.x) * P + slot] = s;
}

// red[p][c0 + blockIdx.x] = sum over blocks of part[p][b][c0 + blockIdx.x] (one workgroup per value)
__global__ __launch_bounds__(NTH) void kr_reduce(const double* __restrict__ part, int nblk, int P, int c0, double* red) {
    __shared__ double lds[NTH / 64];
    const int p = blockIdx.y, c = c0 + blockIdx.x;
    double a = 0;
    for (int b = threadIdx.x; b < nblk; b += NTH) a += part[((size_t)p * nblk + b) * P + c];
    const double s = block_sum(a, lds);
    if (threadIdx.x == 0) red[(size_t)p * P + c] = s;
}

template <class T, class TA>
__global__ __launch_bounds__(NTH) void kr_scale(VecOpsT<T, TA> v, int n) {
    const int i = blockIdx.x * NTH + threadIdx.x, p = blockIdx.y;
    if (i < (v.n[p] > 0 ? v.n[p] : n)) {
        const zd a = ldz(v.a[p][i]);
        v.dst[p][i] = stz<T>(v.f[p] * a.re, v.f[p] * a.im);
    }
}

template <class T>
__global__ __launch_bounds__(NTH) void kr_sub(VecOpsT<T> v, int n) {
    const int i = blockIdx.x * NTH + threadIdx.x, p = blockIdx.y;
    if (i < (v.n[p] > 0 ? v.n[p] : n)) {
        const zd a = ldz(v.a[p][i]), b = ldz(v.b[p][i]);
        v.dst[p][i] = stz<T>(a.re - b.re, a.im - b.im);
    }
}

template <class T>
__global__ __launch_bounds__(NTH) void kr_copy(VecOpsT<T> v, int n) {
    const int i = blockIdx.x * NTH + threadIdx.x, p = blockIdx.y;
    if (i < (v.n[p] > 0 ? v.n[p] : n)) v.dst[p][i] = v.a[p][i];
}

template <class T>
__global__ __launch_bounds__(NTH) void kr_zero(VecOpsT<T> v, int n) {
    const int i = blockIdx.x * NTH + threadIdx.x, p = blockIdx.y;
    if (i < (v.n[p] > 0 ? v.n[p] : n)) v.dst[p][i] = cpx<T>();
}

template <class T, class TA>
__global__ __launch_bounds__(NTH) void kr_combine(VecOpsT<T, TA> v, const zd* __restrict__ y, int rows, int n, int ld) {
    __shared__ double ys_raw[1024];
    zd* ys = reinterpret_cast<zd*>(ys_raw);
    const int p = blockIdx.y, k = v.k[p];
    for (int r = threadIdx.x; r < k; r += NTH) ys[r] = y[(size_t)p * rows + r];
    __syncthreads();
    const int i = blockIdx.x * NTH + threadIdx.x;
    if (i >= (v.n[p] > 0 ? v.n[p] : n)) return;
    const cpx<TA>* V = v.a[p];
    zd acc = ldz(v.dst[p][i]);
    for (int r = 0; r < k; ++r) acc = acc + ldz(V[(size_t)r * ld + i]) * ys[r];
    v.dst[p][i] = stz<T>(acc.re, acc.im);
}

struct Conv {
    cpx<float>* f[KMAX];
    zd* d[KMAX];
};

__global__ __launch_bounds__(NTH) void kr_to_float(Conv c, int n) {
    const int i = blockIdx.x * NTH + threadIdx.x, p = blockIdx.y;
    if (i < n) c.f[p][i] = cpx<float>((float)c.d[p][i].re, (float)c.d[p][i].im);
}

__global__ __launch_bounds__(NTH) void kr_add_float(Conv c, int n) {
    const int i = blockIdx.x * NTH + threadIdx.x, p = blockIdx.y;
    if (i < n) c.d[p][i] = c.d[p][i] + zd(c.f[p][i].re, c.f[p][i].im);
}

}  // namespace

void Krylov::init(int n_, int rows) {
    if (rows > 512) fail(GT_ERR_ARG, "GMRES restart above 511");
    n = n_;
    rows_cap = rows;
    P = 4 * rows + 8;
    nblk = part_blocks(n);
    nblk_d = std::max(1, (n + chunk_d - 1) / chunk_d);
    part.alloc((size_t)KMAX * part_blocks(n) * P);
    red.alloc((size_t)KMAX * P);
}

template <class TV, class TW>
void Krylov::dots(const KrSetT<TV, TW>& k, hipStream_t s) {
    int mmax = 0;
    for (int p = 0; p < k.nr; ++p) mmax = std::max(mmax, k.m[p]);
    const int groups = (mmax + 3) / 4;
    hipLaunchKernelGGL((kr_dots<TV, TW>), dim3(nblk_d * groups, k.nr), dim3(NTH), 0, s, k, n, chunk_d, part.p, P,
                       nblk_d, slot_w());
    GT_LAUNCH("kr_dots");
    hipLaunchKernelGGL(kr_reduce, dim3(2 * mmax, k.nr), dim3(NTH), 0, s, part.p, nblk_d, P, 0, red.p);
    GT_LAUNCH("kr_reduce");
    hipLaunchKernelGGL(kr_reduce, dim3(1, k.nr), dim3(NTH), 0, s, part.p, nblk_d, P, slot_w(), red.p);
    GT_LAUNCH("kr_reduce");
}

template <class TV, class TW>
void Krylov::project_update(const KrSetT<TV, TW>& k, hipStream_t s, bool second) {
    hipLaunchKernelGGL((kr_update<TV, TW>), dim3(nblk, k.nr), dim3(NTH), 0, s, k, n, chunk,
                       (const double*)(red.p + (second ? slot_h2() : 0)), part.p, P, nblk, slot_w_after(), red.p, -1,
                       -1, 0);
    GT_LAUNCH("kr_update");
    hipLaunchKernelGGL(kr_reduce, dim3(1, k.nr), dim3(NTH), 0, s, part.p, nblk, P, slot_w_after(), red.p);
    GT_LAUNCH("kr_reduce");
}

// The fused pass's classes of m, each with its lane layout: a complex double basis takes 16 element lanes by 4 row
// groups (up to 64 rows) and 8 by 8 above; a complex float basis 8 element lanes (pairs) by 8 row groups.
template <class TV, class TW>
void Krylov::update_dots(const KrSetT<TV, TW>& k, hipStream_t s) {
    int mmax = 0;
    bool cls[4] = {false, false, false, false};
    for (int p = 0; p < k.nr; ++p) {
        mmax = std::max(mmax, k.m[p]);
        cls[k.m[p] <= 32 ? 0 : k.m[p] <= 64 ? 1 : 2] = true;
    }
    if (!fused(mmax)) fail(GT_ERR_ARG, "the fused Gram-Schmidt pass serves at most 128 basis rows");
    const double* rd = red.p;
    const dim3 grid(nblk, k.nr);
    if (sizeof(TV) == 8) {
        if (cls[0])
            hipLaunchKernelGGL((kr_update_dots<TV, TW, 16, 8>), grid, dim3(NTH), 0, s, k, n, chunk, rd, part.p, P,
                               nblk, slot_w_after(), slot_h2(), 0);
        if (cls[1])
            hipLaunchKernelGGL((kr_update_dots<TV, TW, 16, 16>), grid, dim3(NTH), 0, s, k, n, chunk, rd, part.p, P,
                               nblk, slot_w_after(), slot_h2(), 32);
    } else {
        if (cls[0])
            hipLaunchKernelGGL((kr_update_dots<TV, TW, 8, 4>), grid, dim3(NTH), 0, s, k, n, chunk, rd, part.p, P,
                               nblk, slot_w_after(), slot_h2(), 0);
        if (cls[1])
            hipLaunchKernelGGL((kr_update_dots<TV, TW, 8, 8>), grid, dim3(NTH), 0, s, k, n, chunk, rd, part.p, P,
                               nblk, slot_w_after(), slot_h2(), 32);
    }
    if (cls[2])
        hipLaunchKernelGGL((kr_update_dots<TV, TW, 8, 16>), grid, dim3(NTH), 0, s, k, n, chunk, rd, part.p, P,
                           nblk, slot_w_after(), slot_h2(), 64);
    GT_LAUNCH("kr_update_dots");
    hipLaunchKernelGGL(kr_reduce, dim3(2 * mmax, k.nr), dim3(NTH), 0, s, part.p, nblk, P, slot_h2(), red.p);
    GT_LAUNCH("kr_reduce");
    hipLaunchKernelGGL(kr_reduce, dim3(1, k.nr), dim3(NTH), 0, s, part.p, nblk, P, slot_w_after(), red.p);
    GT_LAUNCH("kr_reduce");
}

template <class TV, class TW>
void Krylov::repass(const KrSetT<TV, TW>& k, hipStream_t s) {
    hipLaunchKernelGGL((kr_update<TV, TW>), dim3(nblk, k.nr), dim3(NTH), 0, s, k, n, chunk,
                       (const double*)(red.p + slot_h2()), part.p, P, nblk, slot_w_after2(), red.p, slot_w(),
                       slot_w_after(), slot_repass());
    GT_LAUNCH("kr_update");
    hipLaunchKernelGGL(kr_reduce, dim3(1, k.nr), dim3(NTH), 0, s, part.p, nblk, P, slot_w_after2(), red.p);
    GT_LAUNCH("kr_reduce");
}


template <class TV, class TW>
void Krylov::dots2(const KrSetT<TV, TW>& k, hipStream_t s) {
    int mmax = 0;
    for (int p = 0; p < k.nr; ++p) mmax = std::max(mmax, k.m[p]);
    if (mmax + 1 > rows_cap) fail(GT_ERR_ARG, "a basis longer than the workspace's");
    const int groups = std::max(1, (mmax + 7) / 8);
    const int gsz = dots2_group<TV>(), groups_g = std::max(1, (mmax + gsz - 1) / gsz);
    static_assert(chunk2<TV>() == 4 * lane_elems<TV>() * NTH, "kr_dots2 takes four element steps per lane");
    const int c2 = chunk2<TV>(), nb2 = blocks2<TV>();
    // enough chunks to fill the device: every group in one workgroup per chunk (u and v read once)
    // (FP64 bases only: an FP32 pass's u and v pairs in registers would cost it occupancy). Both kernels give the
    // same bits: the same elements per lane in the same order, the same trees.
    if (sizeof(TV) == 8 && (int64_t)nb2 * k.nr >= 8192)
        hipLaunchKernelGGL((kr_dots2_long<TV, TW>), dim3(nb2, k.nr), dim3(NTH), 0, s, k, n, c2, part.p, P, nb2,
                           slot_h2(), slot_uu(), groups);
    else
        hipLaunchKernelGGL((kr_dots2<TV, TW>), dim3(nb2 * groups_g, k.nr), dim3(NTH), 0, s, k, n, c2, part.p, P, nb2,
                           slot_h2(), slot_uu());
    GT_LAUNCH("kr_dots2");
    if (mmax > 0) {
        hipLaunchKernelGGL(kr_reduce, dim3(2 * mmax, k.nr), dim3(NTH), 0, s, part.p, nb2, P, 0, red.p);
        GT_LAUNCH("kr_reduce");
    }
    // u^H u, u^H v and t, adjacent
    hipLaunchKernelGGL(kr_reduce, dim3(3 + 2 * mmax, k.nr), dim3(NTH), 0, s, part.p, nb2, P, slot_uu(), red.p);
    GT_LAUNCH("kr_reduce");
}

template <class TV, class TW>
void Krylov::update2(const KrSetT<TV, TW>& k, const double* ainv, const zd* qv, hipStream_t s) {
    DcgsCoef cf{};
    for (int p = 0; p < k.nr; ++p) {
        cf.ainv[p] = ainv[p];
        cf.qv[p] = qv[p];
        if (k.m[p] + 1 > rows_cap) fail(GT_ERR_ARG, "a basis longer than the workspace's");
    }
    const int nbu = blocks_u();
    hipLaunchKernelGGL((kr_update2<TV, TW>), dim3(nbu, k.nr), dim3(NTH), 0, s, k, n, (int)chunk, (const double*)red.p,
                       part.p, P, nbu, slot_h2(), slot_w_after(), cf);
    GT_LAUNCH("kr_update2");
    hipLaunchKernelGGL(kr_reduce, dim3(1, k.nr), dim3(NTH), 0, s, part.p, nbu, P, slot_w_after(), red.p);
    GT_LAUNCH("kr_reduce");
}

template <class T, class TA>
void Krylov::scale_by_norm(const VecOpsT<T, TA>& v, hipStream_t s) {
    hipLaunchKernelGGL((kr_scale_red<T, TA>), dim3(blocks_for(n, NTH), v.nr), dim3(NTH), 0, s, v, n, (const double*)red.p,
                       P, slot_w_after());
    GT_LAUNCH("kr_scale_red");
}

template <class TV, class TW>
void Krylov::norm2(const KrSetT<TV, TW>& k, hipStream_t s) {
    hipLaunchKernelGGL((kr_norm<TV, TW>), dim3(nblk, k.nr), dim3(NTH), 0, s, k, n, chunk, part.p, P, nblk, slot_norm());
    GT_LAUNCH("kr_norm");
    hipLaunchKernelGGL(kr_reduce, dim3(1, k.nr), dim3(NTH), 0, s, part.p, nblk, P, slot_norm(), red.p);
    GT_LAUNCH("kr_reduce");
}

template <class T, class TA>
void k_scale(const VecOpsT<T, TA>& v, int n, hipStream_t s) {
    hipLaunchKernelGGL((kr_scale<T, TA>), dim3(blocks_for(n, NTH), v.nr), dim3(NTH), 0, s, v, n);
    GT_LAUNCH("kr_scale");
}

template <class T>
void k_sub(const VecOpsT<T>& v, int n, hipStream_t s) {
    hipLaunchKernelGGL(kr_sub<T>, dim3(blocks_for(n, NTH), v.nr), dim3(NTH), 0, s, v, n);
    GT_LAUNCH("kr_sub");
}

template <class T>
void k_copy(const VecOpsT<T>& v, int n, hipStream_t s) {
    hipLaunchKernelGGL(kr_copy<T>, dim3(blocks_for(n, NTH), v.nr), dim3(NTH), 0, s, v, n);
    GT_LAUNCH("kr_copy");
}

template <class T>
void k_zero(const VecOpsT<T>& v, int n, hipStream_t s) {
    hipLaunchKernelGGL(kr_zero<T>, dim3(blocks_for(n, NTH), v.nr), dim3(NTH), 0, s, v, n);
    GT_LAUNCH("kr_zero");
}

template <class T, class TA>
void k_combine(const VecOpsT<T, TA>& v, const zd* y, int rows, int n, int ld, hipStream_t s) {
    for (int p = 0; p < v.nr; ++p)
        if (v.k[p] > 512) fail(GT_ERR_ARG, "combination of more than 512 basis vectors");
    hipLaunchKernelGGL((kr_combine<T, TA>), dim3(blocks_for(n, NTH), v.nr), dim3(NTH), 0, s, v, y, rows, n, ld);
    GT_LAUNCH("kr_combine");
}

void k_to_float(int nr, cpx<float>* const* dst, const zd* const* a, int n, hipStream_t s) {
    Conv c{};
    for (int p = 0; p < nr; ++p) {
        c.f[p] = dst[p];
        c.d[p] = const_cast<zd*>(a[p]);
    }
    hipLaunchKernelGGL(kr_to_float, dim3(blocks_for(n, NTH), nr), dim3(NTH), 0, s, c, n);
    GT_LAUNCH("kr_to_float");
}

void k_add_float(int nr, zd* const* dst, const cpx<float>* const* a, int n, hipStream_t s) {
    Conv c{};
    for (int p = 0; p < nr; ++p) {
        c.f[p] = const_cast<cpx<float>*>(a[p]);
        c.d[p] = dst[p];
    }
    hipLaunchKernelGGL(kr_add_float, dim3(blocks_for(n, NTH), nr), dim3(NTH), 0, s, c, n);
    GT_LAUNCH("kr_add_float");
}

#define GT_KRYLOV(TV, TW)                                                                         \
    template void Krylov::dots<TV, TW>(const KrSetT<TV, TW>&, hipStream_t);                      \
    template void Krylov::project_update<TV, TW>(const KrSetT<TV, TW>&, hipStream_t, bool);      \
    template void Krylov::update_dots<TV, TW>(const KrSetT<TV, TW>&, hipStream_t);               \
    template void Krylov::repass<TV, TW>(const KrSetT<TV, TW>&, hipStream_t);                    \
    template void Krylov::norm2<TV, TW>(const KrSetT<TV, TW>&, hipStream_t);                      \
    template void Krylov::dots2<TV, TW>(const KrSetT<TV, TW>&, hipStream_t);                      \
    template void Krylov::update2<TV, TW>(const KrSetT<TV, TW>&, const double*, const zd*, hipStream_t); \
    template void Krylov::scale_by_norm<TV, TW>(const VecOpsT<TV, TW>&, hipStream_t);
GT_KRYLOV(double, double)
GT_KRYLOV(float, float)
GT_KRYLOV(float, double)
#undef GT_KRYLOV
#define GT_VECOPS(T)                                                                              \
    template void k_sub<T>(const VecOpsT<T>&, int, hipStream_t);                                 \
    template void k_copy<T>(const VecOpsT<T>&, int, hipStream_t);                                \
    template void k_zero<T>(const VecOpsT<T>&, int, hipStream_t);
GT_VECOPS(double)
GT_VECOPS(float)
#undef GT_VECOPS
// scale within one precision, and between them (a float basis row from a double vector and back)
template void k_scale<double, double>(const VecOpsT<double, double>&, int, hipStream_t);
template void k_scale<float, float>(const VecOpsT<float, float>&, int, hipStream_t);
template void k_scale<float, double>(const VecOpsT<float, double>&, int, hipStream_t);
template void k_scale<double, float>(const VecOpsT<double, float>&, int, hipStream_t);
// combinations: a vector of either precision from a basis of its own, and a double vector from a float basis
template void k_combine<double, double>(const VecOpsT<double, double>&, const zd*, int, int, int, hipStream_t);
template void k_combine<float, float>(const VecOpsT<float, float>&, const zd*, int, int, int, hipStream_t);
template void k_combine<double, float>(const VecOpsT<double, float>&, const zd*, int, int, int, hipStream_t);

}  // namespace gt
