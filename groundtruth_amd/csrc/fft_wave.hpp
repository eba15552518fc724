// Wave-level FFTs in registers: S transforms of length N = 64 E at once, lane l holding element l + 64 m of
// transform s in v[s][m] (the natural order at stride 64), forward or inverse, the result in the same layout.
//
// Stockham passes of radix R (R divides E): in a pass, lane l does the butterflies j = l + 64 c (c < E / R) in
// registers, its inputs x[j + r N / R] in v[s][c + r E / R]. That is the stride-64 layout itself before the first
// pass, and the last pass's outputs land there again, so only the passes in between go through LDS: the outputs at
// (j - k) R + k + r Ns (k = j mod Ns, Ns = product of the radices before) are written to the wave's exchange buffer
// and the next pass's inputs read back. The buffer holds one transform (N + N / 8 entries, padded one in eight so
// that 16 lanes of a 16-byte access meet distinct banks, two at worst) and is reused sheet by sheet. A workgroup
// is one wave, so its barriers cost no waiting on other waves.
//
// Twiddles W_N^m (m < N) come from the host table tw; the inverse conjugates them (no 1/N: the caller folds it).
#pragma once

#include "fft_lds.hpp"

namespace gt {

template <int E>
struct WavePlan;                          // radices of the passes for N = 64 E
template <>
struct WavePlan<4> {                      // 256 = 4^4
    static constexpr int n = 4;
    static constexpr int r[4] = {4, 4, 4, 4};
};
template <>
struct WavePlan<8> {                      // 512 = 8^3
    static constexpr int n = 3;
    static constexpr int r[4] = {8, 8, 8, 1};
};
template <>
struct WavePlan<16> {                     // 1024 = 8^3 * 2
    static constexpr int n = 4;
    static constexpr int r[4] = {8, 8, 8, 2};
};

// The butterflies of one pass (radix R, Ns) on every sheet, in registers.
template <class T, int E, int R, int S, bool INV>
__device__ __forceinline__ void wave_butterflies(cpx<T> (&v)[S][E], int N, int Ns, const cpx<T>* __restrict__ tw,
                                                 int lane) {
    constexpr int NB = E / R;
    const int step = N / (Ns * R);
#pragma unroll
    for (int c = 0; c < NB; ++c) {
        const int j = lane + 64 * c;
        const int k = j & (Ns - 1);
        cpx<T> w[R];
        if (Ns > 1) {
#pragma unroll
            for (int r = 1; r < R; ++r) {
                w[r] = tw[k * r * step];
                if (INV) w[r].im = -w[r].im;
            }
        }
#pragma unroll
        for (int s = 0; s < S; ++s) {
            cpx<T> u[R];
#pragma unroll
            for (int r = 0; r < R; ++r) u[r] = v[s][c + r * NB];
            if (Ns > 1) {
#pragma unroll
                for (int r = 1; r < R; ++r) u[r] = u[r] * w[r];
            }
            dft<T, R, INV>(u);
#pragma unroll
            for (int r = 0; r < R; ++r) v[s][c + r * NB] = u[r];
        }
    }
}

// After a pass (radix R, Ns): move sheet s's outputs into the next pass's input order (radix R2) through ex.
template <class T, int E, int R, int R2, int S>
__device__ __forceinline__ void wave_exchange(cpx<T> (&v)[S][E], int N, int Ns, cpx<T>* ex, int lane) {
    constexpr int NB = E / R, NB2 = E / R2;
#pragma unroll
    for (int s = 0; s < S; ++s) {
#pragma unroll
        for (int c = 0; c < NB; ++c) {
            const int j = lane + 64 * c;
            const int k = j & (Ns - 1);
#pragma unroll
            for (int r = 0; r < R; ++r) ex[xpad((j - k) * R + k + r * Ns)] = v[s][c + r * NB];
        }
        __syncthreads();
#pragma unroll
        for (int c = 0; c < NB2; ++c)
#pragma unroll
            for (int r = 0; r < R2; ++r) v[s][c + r * NB2] = ex[xpad(lane + 64 * c + r * (N / R2))];
        __syncthreads();
    }
}

template <class T, int E, int S, bool INV>
__device__ __forceinline__ void wave_fft(cpx<T> (&v)[S][E], cpx<T>* ex, const cpx<T>* __restrict__ tw, int lane) {
    using P = WavePlan<E>;
    constexpr int N = 64 * E;
    constexpr int R0 = P::r[0], R1 = P::r[1], R2 = P::r[2], R3 = P::r[3];
    wave_butterflies<T, E, R0, S, INV>(v, N, 1, tw, lane);
    wave_exchange<T, E, R0, R1, S>(v, N, 1, ex, lane);
    wave_butterflies<T, E, R1, S, INV>(v, N, R0, tw, lane);
    wave_exchange<T, E, R1, R2, S>(v, N, R0, ex, lane);
    wave_butterflies<T, E, R2, S, INV>(v, N, R0 * R1, tw, lane);
    if constexpr (P::n == 4) {
        wave_exchange<T, E, R2, R3, S>(v, N, R0 * R1, ex, lane);
        wave_butterflies<T, E, R3, S, INV>(v, N, R0 * R1 * R2, tw, lane);
    }
}

}  // namespace gt
