// GMRES vector work on the device. Reductions are two-stage and in a fixed order (block partial sums, then one
// tree per value), so a solve gives the same bits on every run.
#include "krylov.hpp"

namespace gt {

namespace {

constexpr int NTH = 256;
constexpr int G = 16;             // basis rows per pass of the multi-dot

// Sum of `val` over the workgroup; the result is valid in thread 0.
__device__ __forceinline__ double block_sum(double val, double* lds) {
    for (int off = 32; off > 0; off >>= 1) val += __shfl_down(val, off, 64);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) lds[wave] = val;
    __syncthreads();
    double r = 0;
    if (threadIdx.x == 0)
        for (int w = 0; w < NTH / 64; ++w) r += lds[w];
    return r;
}

__device__ __forceinline__ double wave_sum(double v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

// Partial sums of h = V[0:m]^H w (and |w|^2) over one block's chunk of the vectors: G basis rows per sweep over
// w, every row's sum reduced within each wave by shuffles and across the four waves in LDS (two barriers a sweep).
__global__ __launch_bounds__(NTH) void kr_dots(const zd* __restrict__ V, int m, const zd* __restrict__ w, int n,
                                               int chunk, double* part, int P, int wslot) {
    __shared__ double red[NTH / 64][2 * G + 1];
    const int i0 = blockIdx.x * chunk, i1 = min(n, i0 + chunk);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    double* out = part + (size_t)blockIdx.x * P;
    for (int g0 = 0; g0 < m; g0 += G) {
        const int gm = min(G, m - g0);
        double ar[G], ai[G];
#pragma unroll
        for (int r = 0; r < G; ++r) ar[r] = ai[r] = 0;
        double ww = 0;
        for (int i = i0 + threadIdx.x; i < i1; i += NTH) {
            const zd wi = w[i];
            ww += wi.re * wi.re + wi.im * wi.im;
#pragma unroll
            for (int r = 0; r < G; ++r) {
                if (r < gm) {
                    const zd v = V[(size_t)(g0 + r) * n + i];
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

__global__ __launch_bounds__(NTH) void kr_update(const zd* __restrict__ V, int m, zd* w, int n, int chunk,
                                                 const double* __restrict__ h, double* part, int P, int slot) {
    __shared__ double lds[NTH / 64];
    __shared__ double hs_raw[1024];
    zd* hs = reinterpret_cast<zd*>(hs_raw);
    for (int r = threadIdx.x; r < m; r += NTH) hs[r] = zd(h[2 * r], h[2 * r + 1]);
    __syncthreads();
    const int i0 = blockIdx.x * chunk, i1 = min(n, i0 + chunk);
    double nn = 0;
    for (int i = i0 + threadIdx.x; i < i1; i += NTH) {
        zd acc = w[i];
        for (int r = 0; r < m; ++r) {
            const zd v = V[(size_t)r * n + i];
            const zd hr = hs[r];
            acc.re -= v.re * hr.re - v.im * hr.im;
            acc.im -= v.re * hr.im + v.im * hr.re;
        }
        w[i] = acc;
        nn += acc.re * acc.re + acc.im * acc.im;
    }
    const double s = block_sum(nn, lds);
    if (threadIdx.x == 0) part[(size_t)blockIdx.x * P + slot] = s;
}

__global__ __launch_bounds__(NTH) void kr_norm(const zd* __restrict__ v, int n, int chunk, double* part, int P,
                                               int slot) {
    __shared__ double lds[NTH / 64];
    const int i0 = blockIdx.x * chunk, i1 = min(n, i0 + chunk);
    double nn = 0;
    for (int i = i0 + threadIdx.x; i < i1; i += NTH) {
        const zd a = v[i];
        nn += a.re * a.re + a.im * a.im;
    }
    const double s = block_sum(nn, lds);
    if (threadIdx.x == 0) part[(size_t)blockIdx.x * P + slot] = s;
}

// red[c] = sum over blocks of part[b][c] for c in [c0, c1): one workgroup per value
__global__ __launch_bounds__(NTH) void kr_reduce(const double* __restrict__ part, int nblk, int P, int c0,
                                                 double* red) {
    __shared__ double lds[NTH / 64];
    const int c = c0 + blockIdx.x;
    double a = 0;
    for (int b = threadIdx.x; b < nblk; b += NTH) a += part[(size_t)b * P + c];
    const double s = block_sum(a, lds);
    if (threadIdx.x == 0) red[c] = s;
}

__global__ __launch_bounds__(NTH) void kr_scale(zd* dst, const zd* src, double f, int n) {
    const int i = blockIdx.x * NTH + threadIdx.x;
    if (i < n) dst[i] = f * src[i];
}

__global__ __launch_bounds__(NTH) void kr_combine(zd* x, const zd* __restrict__ V, const zd* __restrict__ y, int k,
                                                  int n) {
    __shared__ double ys_raw[1024];
    zd* ys = reinterpret_cast<zd*>(ys_raw);
    for (int r = threadIdx.x; r < k; r += NTH) ys[r] = y[r];
    __syncthreads();
    const int i = blockIdx.x * NTH + threadIdx.x;
    if (i >= n) return;
    zd acc = x[i];
    for (int r = 0; r < k; ++r) acc = acc + V[(size_t)r * n + i] * ys[r];
    x[i] = acc;
}

__global__ __launch_bounds__(NTH) void kr_sub(zd* r, const zd* b, const zd* t, int n) {
    const int i = blockIdx.x * NTH + threadIdx.x;
    if (i < n) r[i] = b[i] - t[i];
}

__global__ __launch_bounds__(NTH) void kr_add(zd* x, const zd* t, int n) {
    const int i = blockIdx.x * NTH + threadIdx.x;
    if (i < n) x[i] = x[i] + t[i];
}

}  // namespace

void Krylov::init(int n_, int rows) {
    if (rows > 512) fail(GT_ERR_ARG, "GMRES restart above 511");
    n = n_;
    rows_cap = rows;
    nblk = (int)std::min<long>(2048, std::max<long>(1, (n + 1023) / 1024));
    part.alloc((size_t)nblk * (2 * rows + 8));
    red.alloc(2 * rows + 8);
}

void Krylov::dots(const zd* V, int m, const zd* w, hipStream_t s) {
    const int P = 2 * rows_cap + 8, chunk = (n + nblk - 1) / nblk;
    hipLaunchKernelGGL(kr_dots, dim3(nblk), dim3(NTH), 0, s, V, m, w, n, chunk, part.p, P, 2 * rows_cap);
    GT_LAUNCH("kr_dots");
    hipLaunchKernelGGL(kr_reduce, dim3(2 * m), dim3(NTH), 0, s, part.p, nblk, P, 0, red.p);
    GT_LAUNCH("kr_reduce");
    hipLaunchKernelGGL(kr_reduce, dim3(1), dim3(NTH), 0, s, part.p, nblk, P, 2 * rows_cap, red.p);
    GT_LAUNCH("kr_reduce");
}

void Krylov::project_update(const zd* V, int m, zd* w, hipStream_t s) {
    const int P = 2 * rows_cap + 8, chunk = (n + nblk - 1) / nblk;
    hipLaunchKernelGGL(kr_update, dim3(nblk), dim3(NTH), 0, s, V, m, w, n, chunk, red.p, part.p, P,
                       2 * rows_cap + 1);
    GT_LAUNCH("kr_update");
    hipLaunchKernelGGL(kr_reduce, dim3(1), dim3(NTH), 0, s, part.p, nblk, P, 2 * rows_cap + 1, red.p);
    GT_LAUNCH("kr_reduce");
}

void Krylov::norm2(const zd* v, int slot, hipStream_t s) {
    const int P = 2 * rows_cap + 8, chunk = (n + nblk - 1) / nblk;
    const int c = 2 * rows_cap + 2 + slot;
    hipLaunchKernelGGL(kr_norm, dim3(nblk), dim3(NTH), 0, s, v, n, chunk, part.p, P, c);
    GT_LAUNCH("kr_norm");
    hipLaunchKernelGGL(kr_reduce, dim3(1), dim3(NTH), 0, s, part.p, nblk, P, c, red.p);
    GT_LAUNCH("kr_reduce");
}

void k_scale(zd* dst, const zd* src, double f, int n, hipStream_t s) {
    hipLaunchKernelGGL(kr_scale, dim3(blocks_for(n, NTH)), dim3(NTH), 0, s, dst, src, f, n);
    GT_LAUNCH("kr_scale");
}

void k_combine(zd* x, const zd* V, const zd* y, int k, int n, hipStream_t s) {
    if (k <= 0) return;
    if (k > 512) fail(GT_ERR_ARG, "combination of more than 512 basis vectors");
    hipLaunchKernelGGL(kr_combine, dim3(blocks_for(n, NTH)), dim3(NTH), 0, s, x, V, y, k, n);
    GT_LAUNCH("kr_combine");
}

void k_sub(zd* r, const zd* b, const zd* t, int n, hipStream_t s) {
    hipLaunchKernelGGL(kr_sub, dim3(blocks_for(n, NTH)), dim3(NTH), 0, s, r, b, t, n);
    GT_LAUNCH("kr_sub");
}

void k_add(zd* x, const zd* t, int n, hipStream_t s) {
    hipLaunchKernelGGL(kr_add, dim3(blocks_for(n, NTH)), dim3(NTH), 0, s, x, t, n);
    GT_LAUNCH("kr_add");
}

}  // namespace gt
