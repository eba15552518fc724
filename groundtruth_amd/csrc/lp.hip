// The FFT-convolved partial-inductance product L_p I (peec.py:2054-2135 `_conv` / `Lp_matvec`), fused with the
// system product's strip rows (peec.py:3043-3062) and its incidence rows.
//
// Per strip orientation the currents live on a dense (S, n0, n1) grid (n0 along the strips' axis i, n1 across);
// the kernel spectrum is real and even, stored as its quadrant. Three passes, each one HBM round trip of the
// padded spectrum's valid rows (rows i >= n0 of the padded array are zero on the way in and not needed on the way
// out, so they never touch HBM):
//   A  rows_fwd : per (sheet, tile of rows): scatter the strips into LDS rows, FFT along j (P1), write T1[s][j][i]
//                 (transposed, so that B reads whole columns contiguously).
//   B  cols_mac : per column j: read T1[*][j][0:n0] for every sheet, FFT along i (P0), multiply by the kernel
//                 quadrant and sum over source sheets per pixel (peec.py:2066-2074, _mac.py:20-34), inverse FFT,
//                 write rows 0:n0 to T2[t][j][i].
//   C  rows_inv : per (sheet, tile of rows): read T2, inverse FFT along j, gather at the strips and finish the
//                 product's strip rows: R I + jw (L_p I) - Ar V.
// The incidence rows (branch tail and Ar^T I) run in `tail_bottom`.
#include <mutex>
#include <set>

#include "fft_lds.hpp"
#include "fft_wave.hpp"
#include "lp.hpp"

namespace gt {

namespace {

template <class T>
__device__ __forceinline__ const OrientArgs<T>& pick(const LpArgs<T>& a, int oi) {
    return oi ? a.o[1] : a.o[0];
}

__device__ __forceinline__ int ilog2_dev(int v) { return 31 - __clz(v); }

// Workgroups go to the 8 XCDs round-robin (workgroup w to XCD w mod 8), each XCD with its own L2. Row tiles i0 and
// i0 + RW share the 128-byte lines of the transposed intermediates ([t][j][i]: 8 complex doubles of consecutive i),
// so the tile order is remapped for consecutive tiles to run on one XCD: tile = (w mod 8) (G / 8) + w / 8 over a
// grid of G = a multiple of 8 workgroups (those past the last tile exit).
__device__ __forceinline__ int xcd_tile(int bx, int G) { return (bx & 7) * (G >> 3) + (bx >> 3); }

// Loads a thread of the generic row passes keeps in flight before their LDS stores: a 4096-point row (one per
// workgroup) reads its column-strided T2 entries with every load issued at once, not one round trip per entry.
template <int NTH, int MAXPTS>
constexpr int row_depth() { return MAXPTS / NTH >= 8 ? 8 : MAXPTS / NTH >= 1 ? MAXPTS / NTH : 1; }
constexpr int STRIP_DEPTH = 2;      // strips per thread in flight in the scatter and gather

template <class T, int NTH, int MAXPTS>
__global__ __launch_bounds__(NTH) __attribute__((amdgpu_waves_per_eu(NTH >= 1024 ? 8 : 1))) void lp_rows_fwd(LpArgs<T> a) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cpx<T>* buf = reinterpret_cast<cpx<T>*>(smem_raw);
    int blk = xcd_tile(blockIdx.x, gridDim.x), oi = 0;      // consecutive tiles on one XCD (grid: a multiple of 8)
    if (blk >= a.nblk_rows[0] + a.nblk_rows[1]) return;
    if (blk >= a.nblk_rows[0]) { blk -= a.nblk_rows[0]; oi = 1; }
    const OrientArgs<T> o = pick(a, oi);
    const int ntiles = (o.n0 + o.TR - 1) / o.TR;
    const int s = blk / ntiles, i0 = (blk - s * ntiles) * o.TR;
    const int nr = min(o.TR, o.n0 - i0);
    const int P1s = o.P1 + 1;
    const int p = blockIdx.y;
    for (int e = threadIdx.x; e < nr * P1s; e += NTH) buf[e] = cpx<T>();
    __syncthreads();
    const int kb = o.rowptr[s * o.n0 + i0], ke = o.rowptr[s * o.n0 + i0 + nr];
    const cpx<T>* xin = a.rhs.x[p] + o.lo;
    for (int k0 = kb + threadIdx.x; k0 < ke; k0 += NTH * STRIP_DEPTH) {
        int ij[STRIP_DEPTH];
        cpx<T> v[STRIP_DEPTH];
#pragma unroll
        for (int u = 0; u < STRIP_DEPTH; ++u) {
            const int k = k0 + u * NTH;
            if (k < ke) {
                ij[u] = o.rowcol[k];
                v[u] = xin[k];
            }
        }
#pragma unroll
        for (int u = 0; u < STRIP_DEPTH; ++u)
            if (k0 + u * NTH < ke) buf[((ij[u] >> 16) - i0) * P1s + (ij[u] & 0xffff)] = v[u];
    }
    __syncthreads();
    fft_lds<T, false, NTH, MAXPTS>(buf, o.P1, o.log1, nr, P1s, o.tw1);
    cpx<T>* dst = o.T1 + p * a.t_stride + (size_t)s * o.P1 * o.n0 + i0;
    for (int e = threadIdx.x; e < nr * o.P1; e += NTH) {
        const int r = e % nr, j = e / nr;
        dst[(size_t)j * o.n0 + r] = buf[r * P1s + j];
    }
}

template <class T, int NTH, int MAXPTS, int MAXS>
__global__ __launch_bounds__(NTH) void lp_cols_mac(LpArgs<T> a) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cpx<T>* buf = reinterpret_cast<cpx<T>*>(smem_raw);
    int j = blockIdx.x, oi = 0;
    if (j >= a.nblk_cols[0]) { j -= a.nblk_cols[0]; oi = 1; }
    const OrientArgs<T> o = pick(a, oi);
    const int S = a.S, P0 = o.P0, n0 = o.n0;
    const cpx<T>* T1 = o.T1 + blockIdx.y * a.t_stride;
    cpx<T>* T2 = o.T2 + blockIdx.y * a.t_stride;
    for (int e = threadIdx.x; e < S * P0; e += NTH) {
        const int s = e / P0, i = e - s * P0;
        buf[e] = i < n0 ? T1[((size_t)s * o.P1 + j) * n0 + i] : cpx<T>();
    }
    __syncthreads();
    fft_lds<T, false, NTH, MAXPTS>(buf, P0, o.log0, S, P0, o.tw0);
    // pixel (i, j) of the full spectrum reads quadrant (min(i, P0 - i), min(j, P1 - j)): the spectrum is even
    const int qj = min(j, o.P1 - j);
    const int h0 = P0 / 2 + 1;
    __shared__ int gm[MAXS * MAXS];
    for (int e = threadIdx.x; e < S * S; e += NTH) gm[e] = a.gmap[e];
    __syncthreads();
    for (int i = threadIdx.x; i < P0; i += NTH) {
        const int qi = min(i, P0 - i);
        const T* Kp = o.K + ((size_t)qj * h0 + qi) * a.G;
        cpx<T> v[MAXS];
#pragma unroll
        for (int s = 0; s < MAXS; ++s)
            if (s < S) v[s] = buf[s * P0 + i];
#pragma unroll
        for (int t = 0; t < MAXS; ++t) {
            if (t < S) {
                cpx<T> acc;
#pragma unroll
                for (int s = 0; s < MAXS; ++s) {
                    if (s < S) {
                        const T k = Kp[gm[t * S + s]];
                        acc.re += k * v[s].re;
                        acc.im += k * v[s].im;
                    }
                }
                buf[t * P0 + i] = acc;
            }
        }
    }
    __syncthreads();
    fft_lds<T, true, NTH, MAXPTS>(buf, P0, o.log0, S, P0, o.tw0);
    for (int e = threadIdx.x; e < S * n0; e += NTH) {
        const int t = e / n0, i = e - t * n0;
        T2[((size_t)t * o.P1 + j) * n0 + i] = buf[t * P0 + i];
    }
}


// Pass B for columns whose S sheets exceed the LDS budget (S * P0 > 8192 points), in one launch, by a split of the
// spectrum into its four residue classes of bins k = 4 k' + r [the decimation-in-frequency split of Cooley and Tukey
// 1965]. A column's input u_s[i] is zero for i >= n0, and n0 <= P0 / 2 = 2 M (M = P0 / 4), so with i = m + M q:
//   U_s[4 k' + r] = FFT_M(z_{s,r})[k'],   z_{s,r}[m] = w^(m r) (u_s[m] + (-i)^r u_s[m + M]),   w = e^(-2 pi i / P0),
// and the inverse transform's rows i = m + M q (q = 0, 1) of output sheet t are
//   y_t[m + M q] = sum_r i^(q r) conj(w^(m r)) IFFT_M(Y_{t,r})[m],   Y_{t,r}[k'] = sum_s K_ts(4 k' + r) U_s[4 k' + r].
// One workgroup per column takes the residues in turn: the S transforms of one residue class (S * M points, 96 KB of
// FP64 LDS for six sheets of 4096-point columns), the kernel multiply-add per bin, the S inverse transforms, and the
// rows accumulated in registers (thread m holds rows m and m + M of every sheet). Nothing goes through HBM between
// the forward transform and the inverse one; the column's input is read once per residue class (from the L2 or the
// last-level cache after the first). The sums are the same DFT as the one-workgroup pass, grouped differently.
constexpr int RES4_FG = 3;       // sheets per group of transforms in the residue-class pass
constexpr int RES4_KQ = 8;       // distinct kernels per pixel the residue-class pass stages in LDS (more: read direct)

template <class T, int NTH, int SMAX, int MPT, bool TOEP>
__global__ __launch_bounds__(NTH) void lp_cols_res4(LpArgs<T> a) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cpx<T>* buf = reinterpret_cast<cpx<T>*>(smem_raw);
    constexpr int M = NTH * MPT;                              // thread x holds rows x + NTH c (c < MPT)
    constexpr int MS = M + M / 8;                             // a sheet's transform in LDS, padded one in eight
    const int ncol = a.nblk_cols[0] + a.nblk_cols[1];
    int k = xcd_tile(blockIdx.x, gridDim.x), oi = 0;        // the columns j and P1 - j (one kernel row) on one XCD
    if (k >= ncol) return;
    if (k >= a.nblk_cols[0]) { k -= a.nblk_cols[0]; oi = 1; }
    const OrientArgs<T> o = pick(a, oi);
    const int S = a.S, P0 = o.P0, P1 = o.P1, n0 = o.n0;
    const int j = k < 2 ? (k ? P1 / 2 : 0) : ((k & 1) ? P1 - (k >> 1) : (k >> 1));
    const size_t sheet = (size_t)P1 * n0;
    const cpx<T>* T1 = o.T1 + blockIdx.y * a.t_stride + (size_t)j * n0;
    cpx<T>* T2 = o.T2 + blockIdx.y * a.t_stride + (size_t)j * n0;
    // the M-point transforms' twiddles, pass by pass (conflict-free reads), in LDS after the sheets
    const int logM = ilog2_dev(M);
    cpx<T>* twm = buf + S * MS;
    fft_twiddle_table<T>(twm, M, logM, o.twq, threadIdx.x, NTH);
    // a bin's kernel values, one column of RES4_KQ per thread, after the twiddles (not for eight FP64 sheets of
    // 4096-point columns: their LDS is full)
    constexpr bool KQ = !(sizeof(T) == 8 && SMAX == 8 && MPT == 2);
    T* kq = reinterpret_cast<T*>(twm + fft_twiddle_count(logM));
    const int qj = min(j, P1 - j), h0 = P0 / 2 + 1;
    const T* Kcol = o.K + (size_t)qj * h0 * a.G;
    cpx<T> acc[MPT][SMAX][2];
    for (int r = 0; r < 4; ++r) {
        // z_{s,r}[m] = w^(m r) (u_s[m] + (-i)^r u_s[m + M]) for every sheet: a row's loads all in flight at once
#pragma unroll 1
        for (int c = 0; c < MPT; ++c) {
            const int m = threadIdx.x + NTH * c;
            const cpx<T> wr = o.tw0[m * r];                   // w^(m r), m r < 3 M < P0
            cpx<T> u0[SMAX], u1[SMAX];
#pragma unroll
            for (int s = 0; s < SMAX; ++s) {
                if (s < S) {
                    u0[s] = m < n0 ? T1[s * sheet + m] : cpx<T>();
                    u1[s] = m + M < n0 ? T1[s * sheet + m + M] : cpx<T>();
                }
            }
#pragma unroll
            for (int s = 0; s < SMAX; ++s) {
                if (s < S) {
                    const cpx<T> p = u0[s], q = u1[s];
                    const cpx<T> z = r == 0 ? p + q : r == 1 ? cpx<T>(p.re + q.im, p.im - q.re)
                                   : r == 2 ? p - q : cpx<T>(p.re - q.im, p.im + q.re);
                    buf[s * MS + xpad(m)] = z * wr;
                }
            }
        }
        __syncthreads();
        // the sheets' transforms in groups of RES4_FG: fewer butterflies, so fewer registers, per thread and pass
        for (int h = 0; h < S; h += RES4_FG)
            fft_lds<T, false, NTH, RES4_FG * M, true, true>(buf + h * MS, M, logM, min(RES4_FG, S - h), MS, twm);
        // bins k = 4 k' + r: the kernel multiply-add over source sheets. Sheets equally spaced (TOEP: the kernel of
        // sheets t, s is kernel |t - s|): every bin's kernel values loaded at once into registers. Otherwise a bin's G
        // kernel values (one per distinct face gap) are loaded together into the thread's column of kq, then read per
        // sheet pair through the map.
        T kl[MPT][SMAX];
        if constexpr (TOEP) {
#pragma unroll
            for (int c = 0; c < MPT; ++c) {
                const int kk = 4 * (threadIdx.x + NTH * c) + r;
                const T* Kp = Kcol + (size_t)min(kk, P0 - kk) * a.G;
#pragma unroll
                for (int g = 0; g < SMAX; ++g)
                    if (g < S) kl[c][g] = Kp[g];
            }
        }
#pragma unroll
        for (int c = 0; c < MPT; ++c) {
            const int m = threadIdx.x + NTH * c;
            const int kk = 4 * m + r;
            const T* Kp = Kcol + (size_t)min(kk, P0 - kk) * a.G;
            cpx<T> v[SMAX];
#pragma unroll
            for (int s = 0; s < SMAX; ++s)
                if (s < S) v[s] = buf[s * MS + xpad(m)];
            if constexpr (TOEP) {
#pragma unroll
                for (int t = 0; t < SMAX; ++t) {
                    if (t < S) {
                        cpx<T> y;
#pragma unroll
                        for (int s = 0; s < SMAX; ++s) {
                            if (s < S) {
                                const T kv = kl[c][t > s ? t - s : s - t];
                                y.re += kv * v[s].re;
                                y.im += kv * v[s].im;
                            }
                        }
                        buf[t * MS + xpad(m)] = y;
                    }
                }
            } else if (KQ && a.G <= RES4_KQ) {
                T kl[RES4_KQ];
#pragma unroll
                for (int g = 0; g < RES4_KQ; ++g)
                    if (g < a.G) kl[g] = Kp[g];
#pragma unroll
                for (int g = 0; g < RES4_KQ; ++g)
                    if (g < a.G) kq[g * NTH + threadIdx.x] = kl[g];
#pragma unroll 1
                for (int t = 0; t < S; ++t) {
                    cpx<T> y;
#pragma unroll
                    for (int s = 0; s < SMAX; ++s) {
                        if (s < S) {
                            const T kv = kq[a.gpad[t * LP_MAX_SHEETS + s] * NTH + threadIdx.x];
                            y.re += kv * v[s].re;
                            y.im += kv * v[s].im;
                        }
                    }
                    buf[t * MS + xpad(m)] = y;
                }
            } else {
#pragma unroll 1
                for (int t = 0; t < S; ++t) {
                    cpx<T> y;
#pragma unroll
                    for (int s = 0; s < SMAX; ++s) {
                        if (s < S) {
                            const T kv = Kp[a.gpad[t * LP_MAX_SHEETS + s]];
                            y.re += kv * v[s].re;
                            y.im += kv * v[s].im;
                        }
                    }
                    buf[t * MS + xpad(m)] = y;
                }
            }
        }
        __syncthreads();
        for (int h = 0; h < S; h += RES4_FG)
            fft_lds<T, true, NTH, RES4_FG * M, true, true>(buf + h * MS, M, logM, min(RES4_FG, S - h), MS, twm);
        // rows m and m + M: conj(w^(m r)) Z, and i^r times it
#pragma unroll
        for (int c = 0; c < MPT; ++c) {
            const int m = threadIdx.x + NTH * c;
            const cpx<T> wr = o.tw0[m * r];
#pragma unroll
            for (int t = 0; t < SMAX; ++t) {
                if (t < S) {
                    const cpx<T> z = buf[t * MS + xpad(m)];
                    const cpx<T> A(z.re * wr.re + z.im * wr.im, z.im * wr.re - z.re * wr.im);
                    acc[c][t][0] = acc[c][t][0] + A;
                    acc[c][t][1] = r == 0 ? acc[c][t][1] + A : r == 1 ? acc[c][t][1] + cpx<T>(-A.im, A.re)
                                 : r == 2 ? acc[c][t][1] - A : acc[c][t][1] + cpx<T>(A.im, -A.re);
                }
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int c = 0; c < MPT; ++c) {
        const int m = threadIdx.x + NTH * c;
#pragma unroll
        for (int t = 0; t < SMAX; ++t) {
            if (t < S) {
                if (m < n0) T2[t * sheet + m] = acc[c][t][0];
                if (m + M < n0) T2[t * sheet + m + M] = acc[c][t][1];
            }
        }
    }
}

// Pass B with the transforms in registers (column length P0 = 64 E, S <= 4 sheets): one wave per column j. Lane l
// holds rows l + 64 m of every sheet, loaded straight from T1 (whole contiguous columns); the forward FFT, the
// kernel multiply-add over source sheets per pixel (both sheets of a pixel sit in the same lane) and the inverse
// FFT run in registers with the wave's LDS exchanges between passes; rows 0:n0 go straight to T2.
template <class T, int E, int S, int CW>
__global__ __launch_bounds__(64 * CW) void lp_cols_mac_w(LpArgs<T> a) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    constexpr int N = 64 * E;
    const int wave = threadIdx.x >> 6;
    cpx<T>* ex = reinterpret_cast<cpx<T>*>(smem_raw) + wave * (N + N / 8);
    // one (column, right-hand side) per wave, CW consecutive ones per workgroup, in an order that keeps a column's
    // right-hand sides, and the columns j and P1 - j that read the same kernel row, next to each other on one XCD
    // (xcd_tile). A wave past the last item runs item 0 without storing it, so every wave meets every barrier.
    const int nr = a.rhs.nr;
    const int items = (a.nblk_cols[0] + a.nblk_cols[1]) * nr;
    int item = xcd_tile(blockIdx.x, gridDim.x) * CW + wave;
    const bool live = item < items;
    if (!live) item = 0;
    int k = item / nr, oi = 0;
    const int p = item - k * nr;
    if (k >= a.nblk_cols[0]) { k -= a.nblk_cols[0]; oi = 1; }
    const OrientArgs<T> o = pick(a, oi);
    const int lane = threadIdx.x & 63, n0 = o.n0, P1 = o.P1;
    const int j = (k & 1) ? P1 - 1 - (k >> 1) : (k >> 1);
    const cpx<T>* T1 = o.T1 + p * a.t_stride;
    cpx<T>* T2 = o.T2 + p * a.t_stride;
    cpx<T> v[S][E];
#pragma unroll
    for (int s = 0; s < S; ++s)
#pragma unroll
        for (int m = 0; m < E; ++m) {
            const int i = lane + 64 * m;
            v[s][m] = i < n0 ? T1[((size_t)s * P1 + j) * n0 + i] : cpx<T>();
        }
    wave_fft<T, E, S, false>(v, ex, o.tw0, lane);
    // pixel (i, j) of the full spectrum reads quadrant (min(i, P0 - i), min(j, P1 - j)): the spectrum is even
    const int qj = min(j, P1 - j);
    constexpr int h0 = N / 2 + 1;
    int gi[S * S];
#pragma unroll
    for (int e = 0; e < S * S; ++e) gi[e] = a.gmap[e];
#pragma unroll
    for (int m = 0; m < E; ++m) {
        const int i = lane + 64 * m;
        const T* Kp = o.K + ((size_t)qj * h0 + min(i, N - i)) * a.G;
        cpx<T> out[S];
#pragma unroll
        for (int t = 0; t < S; ++t) {
#pragma unroll
            for (int s = 0; s < S; ++s) {
                const T k = Kp[gi[t * S + s]];
                out[t].re += k * v[s][m].re;
                out[t].im += k * v[s][m].im;
            }
        }
#pragma unroll
        for (int t = 0; t < S; ++t) v[t][m] = out[t];
    }
    wave_fft<T, E, S, true>(v, ex, o.tw0, lane);
    if (!live) return;
#pragma unroll
    for (int t = 0; t < S; ++t)
#pragma unroll
        for (int m = 0; m < E; ++m) {
            const int i = lane + 64 * m;
            if (i < n0) T2[((size_t)t * P1 + j) * n0 + i] = v[t][m];
        }
}

// Passes A and C with the row transforms in registers (row length P1 = 64 E): a workgroup is RW waves, one row
// each, of one sheet's tile of rows [i0, i0 + RW). Each row lives in LDS as a padded strip (xpad) that is also its
// wave's exchange buffer. Between LDS and HBM the tile moves transposed (T1 / T2 are [sheet][j][i]): RW-row segments.

template <class T, int E, int RW>
__global__ __launch_bounds__(64 * RW) void lp_rows_fwd_w(LpArgs<T> a, int nb0, int nblk) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    constexpr int N = 64 * E, NS = N + N / 8;
    int blk = xcd_tile(blockIdx.x, gridDim.x), oi = 0;
    if (blk >= nblk) return;
    if (blk >= nb0) { blk -= nb0; oi = 1; }
    const OrientArgs<T> o = pick(a, oi);
    const int ntiles = (o.n0 + RW - 1) / RW;
    const int s = blk / ntiles, i0 = (blk - s * ntiles) * RW;
    const int nr = min(RW, o.n0 - i0);
    const int p = blockIdx.y, lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    cpx<T>* row = reinterpret_cast<cpx<T>*>(smem_raw) + w * NS;
    for (int e = lane; e < NS; e += 64) row[e] = cpx<T>();
    __syncthreads();
    if (w < nr) {
        const int kb = o.rowptr[s * o.n0 + i0 + w], ke = o.rowptr[s * o.n0 + i0 + w + 1];
        const cpx<T>* xin = a.rhs.x[p] + o.lo;
        for (int k = kb + lane; k < ke; k += 64) row[xpad(o.rowcol[k] & 0xffff)] = xin[k];
    }
    __syncthreads();
    cpx<T> v[1][E];
#pragma unroll
    for (int m = 0; m < E; ++m) v[0][m] = row[xpad(lane + 64 * m)];
    __syncthreads();
    wave_fft<T, E, 1, false>(v, row, o.tw1, lane);
#pragma unroll
    for (int m = 0; m < E; ++m) row[xpad(lane + 64 * m)] = v[0][m];
    __syncthreads();
    cpx<T>* dst = o.T1 + p * a.t_stride + (size_t)s * N * o.n0 + i0;
    const cpx<T>* tile = reinterpret_cast<const cpx<T>*>(smem_raw);
    for (int e = threadIdx.x; e < nr * N; e += 64 * RW) {
        const int r = e % nr, j = e / nr;
        dst[(size_t)j * o.n0 + r] = tile[r * NS + xpad(j)];
    }
}

template <class T, int E, int RW>
__global__ __launch_bounds__(64 * RW) void lp_rows_inv_w(LpArgs<T> a, int nb0, int nblk) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    constexpr int N = 64 * E, NS = N + N / 8;
    int blk = xcd_tile(blockIdx.x, gridDim.x), oi = 0;
    if (blk >= nblk) return;
    if (blk >= nb0) { blk -= nb0; oi = 1; }
    const OrientArgs<T> o = pick(a, oi);
    const int ntiles = (o.n0 + RW - 1) / RW;
    const int t = blk / ntiles, i0 = (blk - t * ntiles) * RW;
    const int nr = min(RW, o.n0 - i0);
    const int p = blockIdx.y, lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    cpx<T>* tile = reinterpret_cast<cpx<T>*>(smem_raw);
    const cpx<T>* src = o.T2 + p * a.t_stride + (size_t)t * N * o.n0 + i0;
    for (int e = threadIdx.x; e < nr * N; e += 64 * RW) {
        const int r = e % nr, j = e / nr;
        tile[r * NS + xpad(j)] = src[(size_t)j * o.n0 + r];
    }
    __syncthreads();
    cpx<T>* row = tile + w * NS;
    cpx<T> v[1][E];
#pragma unroll
    for (int m = 0; m < E; ++m) v[0][m] = w < nr ? row[xpad(lane + 64 * m)] : cpx<T>();
    __syncthreads();
    wave_fft<T, E, 1, true>(v, row, o.tw1, lane);
#pragma unroll
    for (int m = 0; m < E; ++m) row[xpad(lane + 64 * m)] = v[0][m];
    __syncthreads();
    if (w >= nr) return;
    const int i = i0 + w;
    const int kb = o.rowptr[t * o.n0 + i], ke = o.rowptr[t * o.n0 + i + 1];
    cpx<T>* y = a.rhs.y[p];
    const cpx<T>* x = a.rhs.x[p];
    for (int k = kb + lane; k < ke; k += 64) {
        const cpx<T> val = row[xpad(o.rowcol[k] & 0xffff)];
        const int b = o.lo + k;
        if (a.mode == 0) {
            y[b] = val;
        } else {
            const cpx<T> xb = x[b];
            const cpx<T> R = a.rhs.r_sheet[p][t];
            const T wf = a.rhs.w[p];
            cpx<T> vv = R * xb;
            vv.re -= wf * val.im;
            vv.im += wf * val.re;
            const int ka = a.a_keep[b], kk = a.b_keep[b];
            if (ka >= 0) vv = vv - x[a.rhs.nb[p] + ka];
            if (kk >= 0) vv = vv + x[a.rhs.nb[p] + kk];
            y[b] = vv;
        }
    }
}

// The register forms of passes A and C when both orientations share a row length they have (256, 512, 1024), with
// RW rows per workgroup.
template <class T, int E, int RW>
bool launch_rows_w_rw(const LpArgs<T>& a, int nr_rhs, bool inverse, hipStream_t st, int P1) {
    int nb[2] = {0, 0};
    for (int oi = 0; oi < 2; ++oi)
        if (a.nblk_cols[oi] != 0) nb[oi] = a.S * ((a.o[oi].n0 + RW - 1) / RW);
    const int blocks = nb[0] + nb[1];
    if (blocks == 0) return false;
    const size_t lds = (size_t)RW * (P1 + P1 / 8) * sizeof(cpx<T>);
    if (lds > 65536) return false;            // (P1 = 1024 in FP64 with 4 rows) a smaller tile takes it
    const int grid = (blocks + 7) / 8 * 8;
    if (inverse)
        hipLaunchKernelGGL((lp_rows_inv_w<T, E, RW>), dim3(grid, nr_rhs), dim3(64 * RW), lds, st, a, nb[0], blocks);
    else
        hipLaunchKernelGGL((lp_rows_fwd_w<T, E, RW>), dim3(grid, nr_rhs), dim3(64 * RW), lds, st, a, nb[0], blocks);
    return true;
}

// rw rows per workgroup, or fewer where rw rows do not fit 64 KB of LDS: a row length that has the register form
// takes it for every plan and batch size (the LDS passes split the transform into other radices, so they would round
// differently and a right-hand side would not get the same bits alone and batched)
template <class T, int E>
bool launch_rows_w_e(const LpArgs<T>& a, int nr_rhs, bool inverse, hipStream_t st, int P1, int rw) {
    switch (rw) {
    case 1: return launch_rows_w_rw<T, E, 1>(a, nr_rhs, inverse, st, P1);
    case 2:
        return launch_rows_w_rw<T, E, 2>(a, nr_rhs, inverse, st, P1) ||
               launch_rows_w_rw<T, E, 1>(a, nr_rhs, inverse, st, P1);
    default:
        return launch_rows_w_rw<T, E, 4>(a, nr_rhs, inverse, st, P1) ||
               launch_rows_w_rw<T, E, 2>(a, nr_rhs, inverse, st, P1) ||
               launch_rows_w_rw<T, E, 1>(a, nr_rhs, inverse, st, P1);
    }
}

template <class T>
bool launch_rows_w(const LpArgs<T>& a, int nr_rhs, bool inverse, hipStream_t st, int rw) {
    int P1 = 0;
    for (int oi = 0; oi < 2; ++oi) {
        if (a.nblk_cols[oi] == 0) continue;
        if (P1 != 0 && a.o[oi].P1 != P1) return false;
        P1 = a.o[oi].P1;
    }
    switch (P1) {
    case 256: return launch_rows_w_e<T, 4>(a, nr_rhs, inverse, st, P1, rw);
    case 512: return launch_rows_w_e<T, 8>(a, nr_rhs, inverse, st, P1, rw);
    case 1024: return launch_rows_w_e<T, 16>(a, nr_rhs, inverse, st, P1, rw);
    default: return false;
    }
}

template <class T, int E, int CW>
bool launch_cols_w_cw(const LpArgs<T>& a, int ncol, int nr, hipStream_t st) {
    const size_t lds = (size_t)CW * (64 * E + 8 * E) * sizeof(cpx<T>);
    if (lds > 65536) return false;
    const int groups = (ncol * nr + CW - 1) / CW;
    const int grid = (groups + 7) / 8 * 8;
    switch (a.S) {
    case 1: hipLaunchKernelGGL((lp_cols_mac_w<T, E, 1, CW>), dim3(grid), dim3(64 * CW), lds, st, a); return true;
    case 2: hipLaunchKernelGGL((lp_cols_mac_w<T, E, 2, CW>), dim3(grid), dim3(64 * CW), lds, st, a); return true;
    case 3: hipLaunchKernelGGL((lp_cols_mac_w<T, E, 3, CW>), dim3(grid), dim3(64 * CW), lds, st, a); return true;
    case 4: hipLaunchKernelGGL((lp_cols_mac_w<T, E, 4, CW>), dim3(grid), dim3(64 * CW), lds, st, a); return true;
    default: return false;
    }
}

template <class T, int E>
bool launch_cols_w_e(const LpArgs<T>& a, int ncol, int nr, hipStream_t st, int cw) {
    switch (cw) {
    case 2: return launch_cols_w_cw<T, E, 2>(a, ncol, nr, st) || launch_cols_w_cw<T, E, 1>(a, ncol, nr, st);
    case 4: return launch_cols_w_cw<T, E, 4>(a, ncol, nr, st) || launch_cols_w_cw<T, E, 1>(a, ncol, nr, st);
    default: return launch_cols_w_cw<T, E, 1>(a, ncol, nr, st);
    }
}

// The register form of pass B when both orientations share a column length it has (256, 512, 1024) and S <= 4.
template <class T>
bool launch_cols_w(const LpArgs<T>& a, int ncol, int nr, hipStream_t st, int cw) {
    int P0 = 0;
    for (int oi = 0; oi < 2; ++oi) {
        if (a.nblk_cols[oi] == 0) continue;
        if (P0 != 0 && a.o[oi].P0 != P0) return false;
        P0 = a.o[oi].P0;
    }
    switch (P0) {
    case 256: return launch_cols_w_e<T, 4>(a, ncol, nr, st, cw);
    case 512: return launch_cols_w_e<T, 8>(a, ncol, nr, st, cw);
    case 1024: return launch_cols_w_e<T, 16>(a, ncol, nr, st, cw);
    default: return false;
    }
}

template <class T, int NTH, int MAXPTS>
__global__ __launch_bounds__(NTH) __attribute__((amdgpu_waves_per_eu(NTH >= 1024 ? 8 : 1))) void lp_rows_inv(LpArgs<T> a) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cpx<T>* buf = reinterpret_cast<cpx<T>*>(smem_raw);
    int blk = xcd_tile(blockIdx.x, gridDim.x), oi = 0;      // consecutive tiles on one XCD (grid: a multiple of 8)
    if (blk >= a.nblk_rows[0] + a.nblk_rows[1]) return;
    if (blk >= a.nblk_rows[0]) { blk -= a.nblk_rows[0]; oi = 1; }
    const OrientArgs<T> o = pick(a, oi);
    const int ntiles = (o.n0 + o.TR - 1) / o.TR;
    const int t = blk / ntiles, i0 = (blk - t * ntiles) * o.TR;
    const int nr = min(o.TR, o.n0 - i0);
    const int P1s = o.P1 + 1;
    const int p = blockIdx.y;
    const cpx<T>* src = o.T2 + p * a.t_stride + (size_t)t * o.P1 * o.n0 + i0;
    {
        constexpr int U = row_depth<NTH, MAXPTS>();
        const int tot = nr * o.P1;
        for (int e0 = threadIdx.x; e0 < tot; e0 += NTH * U) {
            cpx<T> v[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int e = e0 + u * NTH;
                if (e < tot) {
                    const int r = e % nr, j = e / nr;
                    v[u] = src[(size_t)j * o.n0 + r];
                }
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int e = e0 + u * NTH;
                if (e < tot) {
                    const int r = e % nr, j = e / nr;
                    buf[r * P1s + j] = v[u];
                }
            }
        }
    }
    __syncthreads();
    fft_lds<T, true, NTH, MAXPTS>(buf, o.P1, o.log1, nr, P1s, o.tw1);
    const int kb = o.rowptr[t * o.n0 + i0], ke = o.rowptr[t * o.n0 + i0 + nr];
    cpx<T>* y = a.rhs.y[p];
    const cpx<T>* x = a.rhs.x[p];
    constexpr int V = STRIP_DEPTH;
    for (int k0 = kb + threadIdx.x; k0 < ke; k0 += NTH * V) {
        // each strip's loads (its column, x, its end nodes, then their potentials) issued for all V strips at once
        int ij[V], ka[V], kk[V];
        cpx<T> xb[V], xa[V], xk[V];
#pragma unroll
        for (int u = 0; u < V; ++u) {
            const int k = k0 + u * NTH;
            if (k < ke) {
                ij[u] = o.rowcol[k];
                if (a.mode != 0) {
                    const int b = o.lo + k;
                    xb[u] = x[b];
                    ka[u] = a.a_keep[b];
                    kk[u] = a.b_keep[b];
                }
            }
        }
        if (a.mode != 0) {
#pragma unroll
            for (int u = 0; u < V; ++u) {
                if (k0 + u * NTH < ke) {
                    if (ka[u] >= 0) xa[u] = x[a.rhs.nb[p] + ka[u]];
                    if (kk[u] >= 0) xk[u] = x[a.rhs.nb[p] + kk[u]];
                }
            }
        }
#pragma unroll
        for (int u = 0; u < V; ++u) {
            const int k = k0 + u * NTH;
            if (k >= ke) continue;
            const cpx<T> val = buf[((ij[u] >> 16) - i0) * P1s + (ij[u] & 0xffff)];
            const int b = o.lo + k;
            if (a.mode == 0) {
                y[b] = val;
            } else {
                const cpx<T> R = a.rhs.r_sheet[p][t];
                const T w = a.rhs.w[p];
                cpx<T> v = R * xb[u];
                v.re -= w * val.im;
                v.im += w * val.re;
                if (ka[u] >= 0) v = v - xa[u];
                if (kk[u] >= 0) v = v + xk[u];
                y[b] = v;
            }
        }
    }
}

// Branch rows (z I - Ar V) and node rows (Ar^T Ib) of the system product, one right-hand side per blockIdx.y with
// its own stack's branches.
template <class T>
__global__ __launch_bounds__(256) void tail_bottom(TailArgs<T> a) {
    const int idx = blockIdx.x * 256 + threadIdx.x;
    const int p = blockIdx.y;
    const cpx<T>* x = a.rhs.x[p];
    cpx<T>* y = a.rhs.y[p];
    const int nbr = a.nbr[p], nb = a.rhs.nb[p];
    if (idx < nbr) {
        const int b = a.ns + idx;
        cpx<T> v = a.rhs.z_tail[p][idx] * x[b];
        const int ka = a.a_keep[p][b], kk = a.b_keep[p][b];
        if (ka >= 0) v = v - x[nb + ka];
        if (kk >= 0) v = v + x[nb + kk];
        y[b] = v;
    } else if (idx < nbr + a.nk) {
        const int n = idx - nbr;
        const int* adj = a.adj[p];
        cpx<T> acc;
        for (int e = a.adj_ptr[p][n]; e < a.adj_ptr[p][n + 1]; ++e) {
            const int code = adj[e];
            const cpx<T> v = x[code >> 1];
            if (code & 1) acc = acc - v;
            else acc = acc + v;
        }
        y[nb + n] = acc;
    }
}

}  // namespace

// the residue-class column pass's dynamic LDS at most: the CU's 160 KB (it has no static LDS)
constexpr size_t lds_res4_max() { return 160 * 1024; }

static int ilog2_host(int v) {
    int l = 0;
    while ((1 << l) < v) ++l;
    return l;
}

static int dev_of() {
    int d = 0;
    GT_HIP(hipGetDevice(&d));
    return d;
}

// kernels whose LDS above the default 64 KB was asked for, per device (a function's attributes are the current
// device's); contexts on several devices may launch from several threads
static std::mutex lds_mu;
static std::set<int> lds_set;
static std::set<std::pair<int, const void*>> res4_set;


// The residue-class column pass over the columns of length P0 (2048 or 4096) that `a` names.
template <class T>
void cols_res4(const LpArgs<T>& a, int P0, int nr, hipStream_t st) {
    const int ncol = a.nblk_cols[0] + a.nblk_cols[1];
    if (ncol == 0) return;
    if (P0 != 2048 && P0 != 4096) fail(GT_ERR_UNSUPPORTED, "residue-class column pass: P0 is 2048 or 4096");
    if (a.S > LP_MAX_SHEETS) fail(GT_ERR_UNSUPPORTED, "residue-class column pass: more sheets than it holds");
    const int grid = (ncol + 7) / 8 * 8;              // xcd_tile maps a multiple of 8
    // columns of 4096 points: 512 threads of two rows each (FP64 accumulators take 96 registers), of 2048 points one
    // row each; the twiddles in LDS beside up to six FP64 sheets
    constexpr bool dbl = sizeof(T) == 8;
    constexpr int NT4 = 512, MP4 = 2;
    const bool wide = a.S > 6;
    const int M = P0 / 4, MS = M + M / 8;
    const bool kq = !(dbl && wide && P0 == 4096);
    const size_t lds = ((size_t)a.S * MS + fft_twiddle_count(ilog2_host(M))) * sizeof(cpx<T>) +
                       (kq ? (size_t)RES4_KQ * 512 * sizeof(T) : 0);
    const bool toep = a.toeplitz != 0;
    const void* fn;
#define GT_RES4(NT, SM, MP, TP) (const void*)&lp_cols_res4<T, NT, SM, MP, TP>
    if (P0 == 4096)
        fn = wide ? (toep ? GT_RES4(NT4, 8, MP4, true) : GT_RES4(NT4, 8, MP4, false))
                  : (toep ? GT_RES4(NT4, 6, MP4, true) : GT_RES4(NT4, 6, MP4, false));
    else
        fn = wide ? (toep ? GT_RES4(512, 8, 1, true) : GT_RES4(512, 8, 1, false))
                  : (toep ? GT_RES4(512, 6, 1, true) : GT_RES4(512, 6, 1, false));
#undef GT_RES4
    {
        std::lock_guard<std::mutex> lk(lds_mu);
        const int d = dev_of();
        if (!res4_set.count({d, fn})) {
            GT_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_res4_max()));
            res4_set.insert({d, fn});
        }
    }
    if (lds > lds_res4_max()) fail(GT_ERR_UNSUPPORTED, "residue-class column pass: LDS above its budget");
#define GT_RES4_GO(NT, SM, MP, TP) hipLaunchKernelGGL((lp_cols_res4<T, NT, SM, MP, TP>), dim3(grid, nr), dim3(NT), lds, st, a)
    if (P0 == 4096) {
        if (!wide) { if (toep) GT_RES4_GO(NT4, 6, MP4, true); else GT_RES4_GO(NT4, 6, MP4, false); }
        else { if (toep) GT_RES4_GO(NT4, 8, MP4, true); else GT_RES4_GO(NT4, 8, MP4, false); }
    } else {
        if (!wide) { if (toep) GT_RES4_GO(512, 6, 1, true); else GT_RES4_GO(512, 6, 1, false); }
        else { if (toep) GT_RES4_GO(512, 8, 1, true); else GT_RES4_GO(512, 8, 1, false); }
    }
#undef GT_RES4_GO
    GT_LAUNCH("lp_cols_res4");
}

template <class T>
void launch_lp(const LpArgs<T>& a, const LpShape& sh, hipStream_t st) {
    const int nrow = a.nblk_rows[0] + a.nblk_rows[1];
    const int ncol = a.nblk_cols[0] + a.nblk_cols[1];
    constexpr int NTH = 256;
    // the 128 KB LDS passes (one workgroup per CU) and the split pass's 64 KB column transforms take wider
    // workgroups, so a CU holds 16 waves of them; a butterfly's arithmetic does not depend on its thread
    constexpr int NTH_BIG = 1024;
    const int nrow8 = (nrow + 7) / 8 * 8;       // the generic row passes' grid: xcd_tile maps a multiple of 8
    const int nr = a.rhs.nr;
    const int bucket = nr == 1 ? 0 : nr <= 4 ? 1 : 2;
    const int rw = sh.rw[bucket], cw = sh.cw[bucket];
    // the LDS passes of long transforms (up to 8192 points of a workgroup: 128 KB of FP64) ask for more than the
    // default 64 KB of dynamic LDS, once per kernel
    constexpr int BIG = LpPts<T>::big;
    const bool big_rows = sh.lds_rows > (size_t)LpPts<T>::rows * sizeof(cpx<T>);
    const bool big_cols = sh.lds_cols > (size_t)LpPts<T>::cols * sizeof(cpx<T>);
    // once per device (a function's attributes are the current device's), under a lock: contexts on several
    // devices may launch from several threads
    int dev = 0;
    if (big_rows || big_cols) GT_HIP(hipGetDevice(&dev));
    std::unique_lock<std::mutex> lds_lk(lds_mu, std::defer_lock);
    if (big_rows || big_cols) lds_lk.lock();
    if ((big_rows || big_cols) && !lds_set.count(dev)) {
        const int bytes = (int)(BIG * sizeof(cpx<T>));
        GT_HIP(hipFuncSetAttribute((const void*)&lp_rows_fwd<T, NTH_BIG, BIG>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, bytes));
        GT_HIP(hipFuncSetAttribute((const void*)&lp_rows_inv<T, NTH_BIG, BIG>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, bytes));
        GT_HIP(hipFuncSetAttribute((const void*)&lp_cols_mac<T, NTH_BIG, BIG, LP_MAX_SHEETS>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, bytes));
        lds_set.insert(dev);
    }
    if (lds_lk.owns_lock()) lds_lk.unlock();
    if (!launch_rows_w(a, nr, false, st, rw)) {
        if (big_rows)
            hipLaunchKernelGGL((lp_rows_fwd<T, NTH_BIG, BIG>), dim3(nrow8, nr), dim3(NTH_BIG), sh.lds_rows, st, a);
        else
            hipLaunchKernelGGL((lp_rows_fwd<T, NTH, LpPts<T>::rows>), dim3(nrow8, nr), dim3(NTH), sh.lds_rows, st, a);
    }
    GT_LAUNCH("lp_rows_fwd");
    if (sh.split) {
        const int P0a = a.nblk_cols[0] ? a.o[0].P0 : 0, P0b = a.nblk_cols[1] ? a.o[1].P0 : 0;
        if (P0a == P0b || P0a == 0 || P0b == 0) {
            cols_res4(a, std::max(P0a, P0b), nr, st);
        } else {
            // the orientations' column lengths differ (nx = 2^k + 1): each its own launch, the shorter one (S x P0
            // <= 8192 points) by the one-workgroup pass
            for (int oi = 0; oi < 2; ++oi) {
                LpArgs<T> b = a;
                b.nblk_cols[1 - oi] = 0;
                const int P0 = a.o[oi].P0;
                if (P0 >= 2048) {
                    cols_res4(b, P0, nr, st);
                } else {
                    const size_t lds = (size_t)a.S * P0 * sizeof(cpx<T>);
                    if (lds > (size_t)LpPts<T>::cols * sizeof(cpx<T>))
                        hipLaunchKernelGGL((lp_cols_mac<T, NTH_BIG, BIG, LP_MAX_SHEETS>), dim3(b.nblk_cols[oi], nr),
                                           dim3(NTH_BIG), lds, st, b);
                    else
                        hipLaunchKernelGGL((lp_cols_mac<T, NTH, LpPts<T>::cols, LP_MAX_SHEETS>), dim3(b.nblk_cols[oi], nr),
                                           dim3(NTH), lds, st, b);
                    GT_LAUNCH("lp_cols_mac");
                }
            }
        }
    } else if (!launch_cols_w(a, ncol, nr, st, cw)) {
        if (big_cols)
            hipLaunchKernelGGL((lp_cols_mac<T, NTH_BIG, BIG, LP_MAX_SHEETS>), dim3(ncol, nr), dim3(NTH_BIG), sh.lds_cols,
                               st, a);
        else
            hipLaunchKernelGGL((lp_cols_mac<T, NTH, LpPts<T>::cols, LP_MAX_SHEETS>), dim3(ncol, nr), dim3(NTH),
                               sh.lds_cols, st, a);
    }
    GT_LAUNCH("lp_cols_mac");
    if (!launch_rows_w(a, nr, true, st, rw)) {
        if (big_rows)
            hipLaunchKernelGGL((lp_rows_inv<T, NTH_BIG, BIG>), dim3(nrow8, nr), dim3(NTH_BIG), sh.lds_rows, st, a);
        else
            hipLaunchKernelGGL((lp_rows_inv<T, NTH, LpPts<T>::rows>), dim3(nrow8, nr), dim3(NTH), sh.lds_rows, st, a);
    }
    GT_LAUNCH("lp_rows_inv");
}

template <class T>
void launch_tail_bottom(const TailArgs<T>& a, hipStream_t st) {
    int nbr = 0;
    for (int p = 0; p < a.rhs.nr; ++p) nbr = std::max(nbr, a.nbr[p]);
    const int n = nbr + a.nk;
    if (n == 0) return;
    hipLaunchKernelGGL((tail_bottom<T>), dim3(blocks_for(n, 256), a.rhs.nr), dim3(256), 0, st, a);
    GT_LAUNCH("tail_bottom");
}

template void launch_lp<double>(const LpArgs<double>&, const LpShape&, hipStream_t);
template void launch_lp<float>(const LpArgs<float>&, const LpShape&, hipStream_t);
template void launch_tail_bottom<double>(const TailArgs<double>&, hipStream_t);
template void launch_tail_bottom<float>(const TailArgs<float>&, hipStream_t);

}  // namespace gt
