// GMRES vector work on the device for a batch of right-hand sides (blockIdx.y = right-hand side). Reductions are
// two-stage and in a fixed order (block partial sums, then one tree per value), so a solve gives the same bits on
// every run and whatever else is in its batch.
#include "krylov.hpp"

namespace gt {

namespace {

constexpr int NTH = 256;
constexpr int G = 16;             // basis rows per sweep of the multi-dot

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

// Partial sums of h = V[0:m]^H w (and |w|^2) over one block's chunk: G basis rows per sweep over w, every row's sum
// reduced within each wave by shuffles and across the four waves in LDS (two barriers a sweep).
template <class TV, class TW>
__global__ __launch_bounds__(NTH) void kr_dots(KrSetT<TV, TW> k, int n, int chunk, double* part, int P, int nblk,
                                               int wslot) {
    __shared__ double red[NTH / 64][2 * G + 1];
    const int p = blockIdx.y;
    const cpx<TV>* __restrict__ V = k.V[p];
    const cpx<TW>* __restrict__ w = k.w[p];
    const int m = k.m[p];
    const int i0 = blockIdx.x * chunk, i1 = min(n, i0 + chunk);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    double* out = part + ((size_t)p * nblk + blockIdx.x) * P;
    for (int g0 = 0; g0 < m; g0 += G) {
        const int gm = min(G, m - g0);
        double ar[G], ai[G];
#pragma unroll
        for (int r = 0; r < G; ++r) ar[r] = ai[r] = 0;
        double ww = 0;
        for (int i = i0 + threadIdx.x; i < i1; i += NTH) {
            const zd wi = ldz(w[i]);
            ww += wi.re * wi.re + wi.im * wi.im;
#pragma unroll
            for (int r = 0; r < G; ++r) {
                if (r < gm) {
                    const zd v = ldz(V[(size_t)(g0 + r) * n + i]);
                    ar[r] += v.re * wi.re + v.im * wi.im;
                    ai[r] += v.re * wi.im - v.im * wi.re;
                }
            }
        }
#pragma unroll
        for (int r = 0; r < G; ++r) {
            if (r < gm) {
                ar[r] = wave_sum(ar[r]);
                ai[r] = wave_sum(ai[r]);
            }
        }
        ww = wave_sum(ww);
        if (lane == 0) {
#pragma unroll
            for (int r = 0; r < G; ++r) {
                red[wave][2 * r] = ar[r];
                red[wave][2 * r + 1] = ai[r];
            }
            red[wave][2 * G] = ww;
        }
        __syncthreads();
        const int t = threadIdx.x;
        if (t < 2 * gm) out[2 * g0 + t] = red[0][t] + red[1][t] + red[2][t] + red[3][t];
        if (g0 == 0 && t == 2 * G) out[wslot] = red[0][t] + red[1][t] + red[2][t] + red[3][t];
        __syncthreads();
    }
}

// The update w' = w - V h fused with the products V^H w' (the second Gram-Schmidt pass's coefficients). Lane =
// (element e = lane & 7, row group g = lane >> 3): a lane holds rows g, g + 8, .. (at most RPL) of its element's basis
// column, so each basis value is loaded once for both; the row groups' parts of V h meet by butterfly sums, and the
// products are summed over the block's elements per row.
template <class TV, class TW, int RPL>
__global__ __launch_bounds__(NTH) void kr_update_dots(KrSetT<TV, TW> k, int n, int chunk, const double* red, double* part,
                                                      int P, int nblk, int slot_after, int slot_h2) {
    __shared__ double hs_raw[2 * 8 * RPL];
    __shared__ double wsum[NTH / 64][2 * 8 * RPL + 1];
    zd* hs = reinterpret_cast<zd*>(hs_raw);
    const int p = blockIdx.y;
    const cpx<TV>* __restrict__ V = k.V[p];
    cpx<TW>* w = k.w[p];
    const int m = k.m[p];
    const double* h = red + (size_t)p * P;
    for (int r = threadIdx.x; r < 8 * RPL; r += NTH) hs[r] = r < m ? zd(h[2 * r], h[2 * r + 1]) : zd();
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, e = lane & 7, g = lane >> 3;
    zd hr[RPL];
#pragma unroll
    for (int t = 0; t < RPL; ++t) hr[t] = hs[g + 8 * t];
    double dre[RPL], dim[RPL];
#pragma unroll
    for (int t = 0; t < RPL; ++t) dre[t] = dim[t] = 0.0;
    double nn = 0.0;
    const int i0 = blockIdx.x * chunk, i1 = min(n, i0 + chunk);
    for (int ib = i0; ib < i1; ib += 32) {            // 4 waves x 8 elements per step
        const int i = ib + 8 * wave + e;
        const bool in = i < i1;
        zd vv[RPL];
#pragma unroll
        for (int t = 0; t < RPL; ++t) {
            const int r = g + 8 * t;
            vv[t] = in && r < m ? ldz(V[(size_t)r * n + i]) : zd();
        }
        const zd wi0 = in ? ldz(w[i]) : zd();
        double sre = 0.0, sim = 0.0;
#pragma unroll
        for (int t = 0; t < RPL; ++t) {
            sre += vv[t].re * hr[t].re - vv[t].im * hr[t].im;
            sim += vv[t].re * hr[t].im + vv[t].im * hr[t].re;
        }
#pragma unroll
        for (int off = 8; off < 64; off <<= 1) {
            sre += __shfl_xor(sre, off);
            sim += __shfl_xor(sim, off);
        }
        // the stored value (rounded to T) is what the products and the norm see
        const cpx<TW> ws = stz<TW>(wi0.re - sre, wi0.im - sim);
        const zd wi = ldz(ws);
        if (in && g == 0) {
            w[i] = ws;
            nn += wi.re * wi.re + wi.im * wi.im;
        }
#pragma unroll
        for (int t = 0; t < RPL; ++t) {
            dre[t] += vv[t].re * wi.re + vv[t].im * wi.im;
            dim[t] += vv[t].re * wi.im - vv[t].im * wi.re;
        }
    }
    // the eight element lanes of each row group, then the four waves in order
#pragma unroll
    for (int t = 0; t < RPL; ++t) {
#pragma unroll
        for (int off = 1; off < 8; off <<= 1) {
            dre[t] += __shfl_xor(dre[t], off);
            dim[t] += __shfl_xor(dim[t], off);
        }
    }
    nn = wave_sum(nn);
    if (e == 0) {
#pragma unroll
        for (int t = 0; t < RPL; ++t) {
            wsum[wave][2 * (g + 8 * t)] = dre[t];
            wsum[wave][2 * (g + 8 * t) + 1] = dim[t];
        }
    }
    if (lane == 0) wsum[wave][2 * 8 * RPL] = nn;
    __syncthreads();
    double* out = part + ((size_t)p * nblk + blockIdx.x) * P;
    for (int c = threadIdx.x; c <= 2 * 8 * RPL; c += NTH) {
        const double v = wsum[0][c] + wsum[1][c] + wsum[2][c] + wsum[3][c];
        if (c == 2 * 8 * RPL) out[slot_after] = v;
        else if (c < 2 * m) out[slot_h2 + c] = v;
    }
}

// w -= V h, |w|^2 after. With test_w >= 0 (the re-pass), only where sqrt(red[test_after]) < 0.7 sqrt(red[test_w]):
// block 0 records the decision at red[flag], and a right-hand side without the re-pass leaves its vector alone.
template <class TV, class TW>
__global__ __launch_bounds__(NTH) void kr_update(KrSetT<TV, TW> k, int n, int chunk, const double* red, double* part,
                                                 int P, int nblk, int slot, double* red_out, int test_w,
                                                 int test_after, int flag) {
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
    for (int r = threadIdx.x; r < m; r += NTH) hs[r] = zd(h[2 * r], h[2 * r + 1]);
    __syncthreads();
    const int i0 = blockIdx.x * chunk, i1 = min(n, i0 + chunk);
    double nn = 0;
    for (int i = i0 + threadIdx.x; i < i1; i += NTH) {
        zd acc = ldz(w[i]);
        for (int r = 0; r < m; ++r) {
            const zd v = ldz(V[(size_t)r * n + i]);
            const zd hr = hs[r];
            acc.re -= v.re * hr.re - v.im * hr.im;
            acc.im -= v.re * hr.im + v.im * hr.re;
        }
        const cpx<TW> ws = stz<TW>(acc.re, acc.im);
        w[i] = ws;
        const zd a = ldz(ws);
        nn += a.re * a.re + a.im * a.im;
    }
    const double s = block_sum(nn, lds);
    if (threadIdx.x == 0) part[((size_t)p * nblk + blockIdx.x) * P + slot] = s;
}

template <class TV, class TW>
__global__ __launch_bounds__(NTH) void kr_norm(KrSetT<TV, TW> k, int n, int chunk, double* part, int P, int nblk,
                                               int slot) {
    __shared__ double lds[NTH / 64];
    const int p = blockIdx.y;
    const cpx<TW>* __restrict__ v = k.w[p];
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
    if (i < n) {
        const zd a = ldz(v.a[p][i]);
        v.dst[p][i] = stz<T>(v.f[p] * a.re, v.f[p] * a.im);
    }
}

template <class T>
__global__ __launch_bounds__(NTH) void kr_sub(VecOpsT<T> v, int n) {
    const int i = blockIdx.x * NTH + threadIdx.x, p = blockIdx.y;
    if (i < n) {
        const zd a = ldz(v.a[p][i]), b = ldz(v.b[p][i]);
        v.dst[p][i] = stz<T>(a.re - b.re, a.im - b.im);
    }
}

template <class T>
__global__ __launch_bounds__(NTH) void kr_copy(VecOpsT<T> v, int n) {
    const int i = blockIdx.x * NTH + threadIdx.x, p = blockIdx.y;
    if (i < n) v.dst[p][i] = v.a[p][i];
}

template <class T>
__global__ __launch_bounds__(NTH) void kr_zero(VecOpsT<T> v, int n) {
    const int i = blockIdx.x * NTH + threadIdx.x, p = blockIdx.y;
    if (i < n) v.dst[p][i] = cpx<T>();
}

template <class T, class TA>
__global__ __launch_bounds__(NTH) void kr_combine(VecOpsT<T, TA> v, const zd* __restrict__ y, int rows, int n) {
    __shared__ double ys_raw[1024];
    zd* ys = reinterpret_cast<zd*>(ys_raw);
    const int p = blockIdx.y, k = v.k[p];
    for (int r = threadIdx.x; r < k; r += NTH) ys[r] = y[(size_t)p * rows + r];
    __syncthreads();
    const int i = blockIdx.x * NTH + threadIdx.x;
    if (i >= n) return;
    const cpx<TA>* V = v.a[p];
    zd acc = ldz(v.dst[p][i]);
    for (int r = 0; r < k; ++r) acc = acc + ldz(V[(size_t)r * n + i]) * ys[r];
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
    nblk = (int)std::min<long>(2048, std::max<long>(1, (n + 1023) / 1024));
    part.alloc((size_t)KMAX * nblk * P);
    red.alloc((size_t)KMAX * P);
}

template <class TV, class TW>
void Krylov::dots(const KrSetT<TV, TW>& k, hipStream_t s) {
    const int chunk = (n + nblk - 1) / nblk;
    int mmax = 0;
    for (int p = 0; p < k.nr; ++p) mmax = std::max(mmax, k.m[p]);
    hipLaunchKernelGGL((kr_dots<TV, TW>), dim3(nblk, k.nr), dim3(NTH), 0, s, k, n, chunk, part.p, P, nblk, slot_w());
    GT_LAUNCH("kr_dots");
    hipLaunchKernelGGL(kr_reduce, dim3(2 * mmax, k.nr), dim3(NTH), 0, s, part.p, nblk, P, 0, red.p);
    GT_LAUNCH("kr_reduce");
    hipLaunchKernelGGL(kr_reduce, dim3(1, k.nr), dim3(NTH), 0, s, part.p, nblk, P, slot_w(), red.p);
    GT_LAUNCH("kr_reduce");
}

template <class TV, class TW>
void Krylov::project_update(const KrSetT<TV, TW>& k, hipStream_t s, bool second) {
    const int chunk = (n + nblk - 1) / nblk;
    hipLaunchKernelGGL((kr_update<TV, TW>), dim3(nblk, k.nr), dim3(NTH), 0, s, k, n, chunk,
                       (const double*)(red.p + (second ? slot_h2() : 0)), part.p, P, nblk, slot_w_after(), red.p, -1,
                       -1, 0);
    GT_LAUNCH("kr_update");
    hipLaunchKernelGGL(kr_reduce, dim3(1, k.nr), dim3(NTH), 0, s, part.p, nblk, P, slot_w_after(), red.p);
    GT_LAUNCH("kr_reduce");
}

template <class TV, class TW>
void Krylov::update_dots(const KrSetT<TV, TW>& k, hipStream_t s) {
    const int chunk = (n + nblk - 1) / nblk;
    int mmax = 0;
    for (int p = 0; p < k.nr; ++p) mmax = std::max(mmax, k.m[p]);
    if (!fused(mmax)) fail(GT_ERR_ARG, "the fused Gram-Schmidt pass serves at most 128 basis rows");
    const double* rd = red.p;
    if (mmax <= 32)
        hipLaunchKernelGGL((kr_update_dots<TV, TW, 4>), dim3(nblk, k.nr), dim3(NTH), 0, s, k, n, chunk, rd, part.p, P, nblk,
                           slot_w_after(), slot_h2());
    else if (mmax <= 64)
        hipLaunchKernelGGL((kr_update_dots<TV, TW, 8>), dim3(nblk, k.nr), dim3(NTH), 0, s, k, n, chunk, rd, part.p, P, nblk,
                           slot_w_after(), slot_h2());
    else
        hipLaunchKernelGGL((kr_update_dots<TV, TW, 16>), dim3(nblk, k.nr), dim3(NTH), 0, s, k, n, chunk, rd, part.p, P,
                           nblk, slot_w_after(), slot_h2());
    GT_LAUNCH("kr_update_dots");
    hipLaunchKernelGGL(kr_reduce, dim3(2 * mmax, k.nr), dim3(NTH), 0, s, part.p, nblk, P, slot_h2(), red.p);
    GT_LAUNCH("kr_reduce");
    hipLaunchKernelGGL(kr_reduce, dim3(1, k.nr), dim3(NTH), 0, s, part.p, nblk, P, slot_w_after(), red.p);
    GT_LAUNCH("kr_reduce");
}

template <class TV, class TW>
void Krylov::repass(const KrSetT<TV, TW>& k, hipStream_t s) {
    const int chunk = (n + nblk - 1) / nblk;
    hipLaunchKernelGGL((kr_update<TV, TW>), dim3(nblk, k.nr), dim3(NTH), 0, s, k, n, chunk,
                       (const double*)(red.p + slot_h2()), part.p, P, nblk, slot_w_after2(), red.p, slot_w(),
                       slot_w_after(), slot_repass());
    GT_LAUNCH("kr_update");
    hipLaunchKernelGGL(kr_reduce, dim3(1, k.nr), dim3(NTH), 0, s, part.p, nblk, P, slot_w_after2(), red.p);
    GT_LAUNCH("kr_reduce");
}

template <class TV, class TW>
void Krylov::norm2(const KrSetT<TV, TW>& k, hipStream_t s) {
    const int chunk = (n + nblk - 1) / nblk;
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
void k_combine(const VecOpsT<T, TA>& v, const zd* y, int rows, int n, hipStream_t s) {
    for (int p = 0; p < v.nr; ++p)
        if (v.k[p] > 512) fail(GT_ERR_ARG, "combination of more than 512 basis vectors");
    hipLaunchKernelGGL((kr_combine<T, TA>), dim3(blocks_for(n, NTH), v.nr), dim3(NTH), 0, s, v, y, rows, n);
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
    template void Krylov::norm2<TV, TW>(const KrSetT<TV, TW>&, hipStream_t);
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
template void k_combine<double, double>(const VecOpsT<double, double>&, const zd*, int, int, hipStream_t);
template void k_combine<float, float>(const VecOpsT<float, float>&, const zd*, int, int, hipStream_t);
template void k_combine<double, float>(const VecOpsT<double, float>&, const zd*, int, int, hipStream_t);

}  // namespace gt
