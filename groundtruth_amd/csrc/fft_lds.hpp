// Power-of-two complex FFTs in LDS for one workgroup: Stockham autosort passes of radix 8 (then one of 4 or 2),
// each pass reading its butterflies' inputs into registers, a barrier, and writing them back in place in the next
// pass's order [Stockham 1966; the radix-R form of Bainville's OpenCL FFT]. Several transforms of the same length
// share one buffer (rows of a tile, or the sheets of a column); their butterflies are spread over all threads.
//
// Forward: X[k] = sum_n x[n] e^(-2 pi i n k / N); inverse: the same with +i and no 1/N (the caller folds the
// scale into the kernel spectrum). Twiddles come from a table tw[m] = e^(-2 pi i m / N), m < N, computed on the
// host in double precision.
#pragma once

#include "gt_common.hpp"

namespace gt {

// x * (-i) (forward) or x * (+i) (inverse)
template <class T, bool INV>
__device__ __forceinline__ cpx<T> rot90(cpx<T> x) {
    return INV ? cpx<T>(-x.im, x.re) : cpx<T>(x.im, -x.re);
}

template <class T, bool INV>
__device__ __forceinline__ void dft2(cpx<T>& a, cpx<T>& b) {
    cpx<T> t = a;
    a = t + b;
    b = t - b;
}

template <class T, bool INV>
__device__ __forceinline__ void dft4(cpx<T>& a0, cpx<T>& a1, cpx<T>& a2, cpx<T>& a3) {
    cpx<T> t0 = a0 + a2, t1 = a0 - a2, t2 = a1 + a3, t3 = rot90<T, INV>(a1 - a3);
    a0 = t0 + t2;
    a2 = t0 - t2;
    a1 = t1 + t3;
    a3 = t1 - t3;
}

template <class T, bool INV>
__device__ __forceinline__ void dft8(cpx<T>* a) {
    // even / odd halves, then the W8 twiddles on the odd half
    dft4<T, INV>(a[0], a[2], a[4], a[6]);
    dft4<T, INV>(a[1], a[3], a[5], a[7]);
    const T h = (T)0.70710678118654752440;
    cpx<T> o1 = a[3], o2 = a[5], o3 = a[7];
    // W8^1 = (1 -/+ i) / sqrt2, W8^2 = -/+ i, W8^3 = (-1 -/+ i) / sqrt2
    cpx<T> w1 = INV ? cpx<T>(h * (o1.re - o1.im), h * (o1.re + o1.im)) : cpx<T>(h * (o1.re + o1.im), h * (o1.im - o1.re));
    cpx<T> w2 = rot90<T, INV>(o2);
    cpx<T> w3 = INV ? cpx<T>(-h * (o3.re + o3.im), h * (o3.re - o3.im)) : cpx<T>(h * (o3.im - o3.re), -h * (o3.re + o3.im));
    cpx<T> e0 = a[0], e1 = a[2], e2 = a[4], e3 = a[6], o0 = a[1];
    a[0] = e0 + o0;
    a[4] = e0 - o0;
    a[1] = e1 + w1;
    a[5] = e1 - w1;
    a[2] = e2 + w2;
    a[6] = e2 - w2;
    a[3] = e3 + w3;
    a[7] = e3 - w3;
}

template <class T, int R, bool INV>
__device__ __forceinline__ void dft(cpx<T>* a) {
    if constexpr (R == 2) dft2<T, INV>(a[0], a[1]);
    else if constexpr (R == 4) dft4<T, INV>(a[0], a[1], a[2], a[3]);
    else dft8<T, INV>(a);
}

// A transform's element p at p + p / 8 (PAD): the 16 lanes of a 16-byte access that the first pass's stride-R
// writes spread over meet distinct banks.
__device__ __forceinline__ int xpad(int p) { return p + (p >> 3); }

template <bool PAD>
__device__ __forceinline__ int lds_at(int p) { return PAD ? xpad(p) : p; }

// One Stockham pass of radix R over `ntrans` transforms of length N at buf + t * tstride. Ns: the product of the
// radices of the passes before. MAXB: butterflies per thread (ceil(ntrans N / R / NTH)). Twiddles: W_N^(k r N / (Ns R))
// from the full table tw, or (TWP) from the pass's own table at tw + twoff, laid out [k][r - 1] so that consecutive
// butterflies read consecutive entries (no bank conflicts when tw is in LDS).
template <class T, int R, bool INV, int NTH, int MAXB, bool PAD = false, bool TWP = false>
__device__ __forceinline__ void fft_pass(cpx<T>* buf, int N, int ntrans, int tstride, int Ns,
                                         const cpx<T>* __restrict__ tw, int twoff = 0) {
    const int nbf = N / R;
    const int total = nbf * ntrans;
    const int twstep = N / (Ns * R);
    cpx<T> v[MAXB][R];
#pragma unroll
    for (int c = 0; c < MAXB; ++c) {
        const int b = threadIdx.x + c * NTH;
        if (b < total) {
            const int t = b / nbf, j = b - t * nbf;
            const cpx<T>* src = buf + t * tstride;
#pragma unroll
            for (int r = 0; r < R; ++r) v[c][r] = src[lds_at<PAD>(j + r * nbf)];
            const int k = j & (Ns - 1);
            if (Ns > 1) {
#pragma unroll
                for (int r = 1; r < R; ++r) {
                    cpx<T> w = TWP ? tw[twoff + k * (R - 1) + r - 1] : tw[k * r * twstep];
                    if (INV) w.im = -w.im;
                    v[c][r] = v[c][r] * w;
                }
            }
            dft<T, R, INV>(v[c]);
        }
    }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < MAXB; ++c) {
        const int b = threadIdx.x + c * NTH;
        if (b < total) {
            const int t = b / nbf, j = b - t * nbf;
            const int k = j & (Ns - 1);
            cpx<T>* dst = buf + t * tstride;
#pragma unroll
            for (int r = 0; r < R; ++r) dst[lds_at<PAD>((j - k) * R + k + r * Ns)] = v[c][r];
        }
    }
    __syncthreads();
}

// All passes of ntrans length-N transforms (N = 2^logN, ntrans * N <= MAXPTS). Radix 8 passes, then one radix-4
// pass (logN = 2 mod 3) or two (logN = 1 mod 3, N >= 16); radix 2 only for N = 2. The buffer must be complete
// (written and synchronised) before the call; it is synchronised after.
// the passes of a length-2^logN transform: n8 of radix 8, then n4 of radix 4, then n2 of radix 2
__host__ __device__ __forceinline__ void fft_plan(int logN, int& n8, int& n4, int& n2) {
    n8 = logN / 3;
    n4 = n2 = 0;
    const int r = logN - 3 * n8;
    if (r == 2) n4 = 1;
    else if (r == 1 && n8 > 0) { n8 -= 1; n4 = 2; }
    else if (r == 1) n2 = 1;
}

template <class T, bool INV, int NTH, int MAXPTS, bool PAD = false, bool TWP = false>
__device__ __forceinline__ void fft_lds(cpx<T>* buf, int N, int logN, int ntrans, int tstride,
                                        const cpx<T>* __restrict__ tw) {
    int n8, n4, n2;
    fft_plan(logN, n8, n4, n2);
    int Ns = 1, off = 0;
    for (int p = 0; p < n8; ++p, Ns *= 8) {
        fft_pass<T, 8, INV, NTH, (MAXPTS + 8 * NTH - 1) / (8 * NTH), PAD, TWP>(buf, N, ntrans, tstride, Ns, tw, off);
        if (Ns > 1) off += Ns * 7;
    }
    for (int p = 0; p < n4; ++p, Ns *= 4) {
        fft_pass<T, 4, INV, NTH, (MAXPTS + 4 * NTH - 1) / (4 * NTH), PAD, TWP>(buf, N, ntrans, tstride, Ns, tw, off);
        if (Ns > 1) off += Ns * 3;
    }
    if (n2) fft_pass<T, 2, INV, NTH, (MAXPTS + 2 * NTH - 1) / (2 * NTH), PAD, TWP>(buf, N, ntrans, tstride, Ns, tw, off);
}

// The per-pass twiddle tables fft_lds<..., TWP = true> reads, from the full table tw of W_N^m (m < N), written to dst
// by threads tid < nth (the caller synchronises): pass by pass, entry [k][r - 1] = W_N^(k r N / (Ns R)) for every
// k < Ns and 1 <= r < R of the passes with Ns > 1. Returns nothing; fft_twiddle_count gives the entries.
__host__ __device__ __forceinline__ int fft_twiddle_count(int logN) {
    int n8, n4, n2, Ns = 1, cnt = 0;
    fft_plan(logN, n8, n4, n2);
    for (int p = 0; p < n8; ++p, Ns *= 8) cnt += Ns > 1 ? Ns * 7 : 0;
    for (int p = 0; p < n4; ++p, Ns *= 4) cnt += Ns > 1 ? Ns * 3 : 0;
    if (n2) cnt += Ns > 1 ? Ns : 0;
    return cnt;
}

template <class T>
__device__ __forceinline__ void fft_twiddle_table(cpx<T>* dst, int N, int logN, const cpx<T>* __restrict__ tw, int tid,
                                                  int nth) {
    int n8, n4, n2;
    fft_plan(logN, n8, n4, n2);
    int Ns = 1, off = 0;
    auto fill = [&](int R) {
        if (Ns > 1) {
            const int step = N / (Ns * R), cnt = Ns * (R - 1);
            for (int e = tid; e < cnt; e += nth) {
                const int k = e / (R - 1), r = e - k * (R - 1) + 1;
                dst[off + e] = tw[k * r * step];
            }
            off += cnt;
        }
        Ns *= R;
    };
    for (int p = 0; p < n8; ++p) fill(8);
    for (int p = 0; p < n4; ++p) fill(4);
    if (n2) fill(2);
}

}  // namespace gt
