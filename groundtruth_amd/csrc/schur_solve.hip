// Level-scheduled supernodal solves of B x = b on the device (the Schur block of the preconditioner,
// peec.py:3064-3083 `op.solve_S`), with each supernode's panel stored as Q = [L_CC^-1 ; -L_RC L_CC^-1]
// (schur_factor.cpp), so that neither direction has a serial chain:
//   forward (leaves first):  f = [b_C ; 0] + the children's update entries (pulled in a fixed order),
//                            [z_C ; u_R] = Q f_C + [0 ; f_R]          -- a dense product over the panel's rows
//   backward (roots first):  x_C = Q^T [z_C ; x_R]                     -- a dense product over its columns
// One launch per level. The real panel multiplies up to 16 complex right-hand sides at once: every panel value is
// loaded once and feeds 2 nr fused multiply-adds. A work item is one chunk of one tile: each lane loads its at most
// SCHUR_PV panel values at once, together with the chunk of right-hand sides staged in LDS (rows padded by one entry
// so that 16 rows read at once fall in distinct banks), so an item waits on memory about once. Lanes take 16
// consecutive panel rows x 4 columns: each panel load is four 128-byte segments of the column-major panel. A tile
// split over several chunks writes partial sums, which a finalize launch after the level adds in chunk order. Every
// sum runs in a fixed order that does not depend on the batch: a right-hand side gets the same
// bits alone or in any batch.
#include <cstdint>
#include <type_traits>

#include "lp.hpp"
#include "schur.hpp"

namespace gt {

namespace {

constexpr int NTH = 256;
constexpr int PV = SCHUR_PV;
constexpr int SLOT = SCHUR_TILE * KMAX;          // entries of one partial-sum slot: [tile row or column][rhs]

struct SolveArgs {
    const SnItem* item;
    const int* perm;
    const int* rows;
    const int2* ctab;
    const int64_t* cidx;
    const double* panel;
    zd* ubuf;
    zd* zp;
    zd* xp;
    const zd* b;
    zd* x;
    zd* part;
};

// staged entries per thread with their loads in flight together (fewer for wide batches, whose accumulators take
// the registers)
template <int KB>
constexpr int stage_depth() { return KB >= 8 ? 4 : 8; }
// the forward pass at one right-hand side stages four, so that it fits 64 VGPRs and a CU holds 32 of its waves
template <int KB>
constexpr int fwd_stage_depth() { return KB == 1 ? 4 : stage_depth<KB>(); }

// acc + the children's update entries at a front position (right-hand side p of nr), in the contributors' fixed order
__device__ __forceinline__ zd add_contrib(const SolveArgs& a, int2 ct, int p, int nr, zd acc) {
    if (ct.x >= -1) {
        if (ct.x >= 0) acc = acc + a.ubuf[(size_t)ct.x * nr + p];
        if (ct.y >= 0) acc = acc + a.ubuf[(size_t)ct.y * nr + p];
    } else {
        const int64_t e0 = -2 - (int64_t)ct.x;
        for (int64_t e = e0; e < e0 + ct.y; ++e) acc = acc + a.ubuf[(size_t)a.cidx[e] * nr + p];
    }
    return acc;
}

// acc[p] += w * u[p] over the batch: the LDS reads of each half first, then its multiply-adds (padded right-hand
// sides are staged as zeros)
template <int KB>
__device__ __forceinline__ void mac_row(double (&are)[KB], double (&aim)[KB], double w, const zd* u) {
    constexpr int H = KB > 1 ? KB / 2 : 1;
#pragma unroll
    for (int h0 = 0; h0 < KB; h0 += H) {
        zd t[H];
#pragma unroll
        for (int p = 0; p < H; ++p) t[p] = u[h0 + p];
#pragma unroll
        for (int p = 0; p < H; ++p) {
            are[h0 + p] = fma(w, t[p].re, are[h0 + p]);
            aim[h0 + p] = fma(w, t[p].im, aim[h0 + p]);
        }
    }
}

// Forward item: rows [t0, t0 + 64) of supernode s with its columns split 4 ways in each wave (kw = 1; wave w takes
// rows t0 + 16 w ..), or rows [t0, t0 + 16) with the columns split 16 ways over the four waves (kw = 4), whose sums
// meet in LDS in wave order; columns [k0, k1). Lane = (row rho = lane & 15, column phase lane >> 4).
template <int KB>
__global__ __launch_bounds__(NTH) __attribute__((amdgpu_waves_per_eu(KB == 1 ? 8 : 4))) void schur_fwd(SolveArgs a, int item0,
                                                                                          int nr) {
    extern __shared__ double lds_raw[];
    zd* f = reinterpret_cast<zd*>(lds_raw);
    constexpr int LK = KB + 1;
    const SnItem it = a.item[item0 + blockIdx.x];
    const int C = it.C, m = it.m, c0 = it.c0, kw = it.kw, k0 = it.k0, k1 = it.k1;
    const double* __restrict__ Q = a.panel + it.poff;
    const int2* __restrict__ ct = a.ctab + it.fpos;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int rho = lane & 15, kq = lane >> 4;
    const int r = kw == 1 ? it.t0 + 16 * w + rho : it.t0 + rho;
    const int ph = kw == 1 ? kq : 4 * w + kq;
    const int nph = 4 * kw;
    const bool direct = it.slot0 < 0;
    // this lane's panel values (columns k0 + ph + nph i), loaded with the staging
    double pv[PV];
#pragma unroll
    for (int i = 0; i < PV; ++i) {
        const int q = k0 + ph + nph * i;
        pv[i] = r < m && q < k1 ? Q[(size_t)q * m + r] : 0.0;
    }
    const int2 cr = direct && (kw == 1 || w == 0) && r < m && r >= C ? ct[r] : int2{-1, -1};
    {
        constexpr int U = fwd_stage_depth<KB>();
        const int tot = (k1 - k0) * KB;
        for (int e0 = threadIdx.x; e0 < tot; e0 += NTH * U) {
            int pm[U];
            int2 cc[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int e = e0 + u * NTH, q = k0 + e / KB;
                if (e < tot) {
                    pm[u] = a.perm[c0 + q];
                    cc[u] = ct[q];
                }
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int e = e0 + u * NTH, dq = e / KB, p = e - dq * KB;
                if (e < tot) f[dq * LK + p] = p < nr ? add_contrib(a, cc[u], p, nr, a.b[(size_t)pm[u] * nr + p]) : zd();
            }
        }
    }
    __syncthreads();
    double are[KB], aim[KB];
#pragma unroll
    for (int p = 0; p < KB; ++p) are[p] = aim[p] = 0.0;
    if (r < m) {
#pragma unroll
        for (int i = 0; i < PV; ++i) {
            const int q = k0 + ph + nph * i;
            if (q < k1) mac_row<KB>(are, aim, pv[i], f + (q - k0) * LK);
        }
    }
    // the four column phases of a wave: lanes rho, rho + 16, rho + 32, rho + 48
#pragma unroll
    for (int p = 0; p < KB; ++p) {
        if (p < nr) {
            are[p] += __shfl_xor(are[p], 16);
            aim[p] += __shfl_xor(aim[p], 16);
            are[p] += __shfl_xor(are[p], 32);
            aim[p] += __shfl_xor(aim[p], 32);
        }
    }
    if (kw == 4) {                                  // waves 0, 1, 2, 3 in turn, into the staging buffer
        __syncthreads();
        for (int ww = 0; ww < 4; ++ww) {
            if (w == ww) {
#pragma unroll
                for (int p = 0; p < KB; ++p) {
                    if (p < nr && (p & 3) == kq) {
                        zd* d = f + rho * LK + p;
                        *d = ww == 0 ? zd(are[p], aim[p]) : zd(d->re + are[p], d->im + aim[p]);
                    }
                }
            }
            __syncthreads();
        }
    }
    if (direct) {
        if ((kw == 1 || w == 0) && r < m) {
#pragma unroll
            for (int p = 0; p < KB; ++p) {
                if (p < nr && (p & 3) == kq) {
                    const zd t = kw == 1 ? zd(are[p], aim[p]) : f[rho * LK + p];
                    if (r < C) a.zp[(size_t)(c0 + r) * nr + p] = t;
                    else a.ubuf[(size_t)(it.aux + r - C) * nr + p] = add_contrib(a, cr, p, nr, t);
                }
            }
        }
        return;
    }
    // a split tile (kw = 4): partial sums of its 16 rows
    zd* slot = a.part + (size_t)(it.slot0 + it.kidx) * SLOT;
    if (w == 0 && r < m) {
#pragma unroll
        for (int p = 0; p < KB; ++p)
            if (p < nr && (p & 3) == kq) slot[rho * KMAX + p] = f[rho * LK + p];
    }
}

// Backward item: a tile of 4 * 64 / RP columns of supernode s starting at t0 (RP = 16: 16 columns; RP = 8: 32), rows
// [k0, k1); wave w takes columns t0 + w * 64 / RP + lane / RP, the RP lanes of a column split the rows (r = k0 + rho,
// k0 + rho + RP, ..) and meet by butterfly sums.
template <int KB, int RP>
__device__ __forceinline__ void bwd_item(const SolveArgs& a, const SnItem& it, int nr, zd* v) {
    constexpr int LK = KB + 1, CW = 64 / RP;
    const int C = it.C, m = it.m, c0 = it.c0, q0 = it.t0, k0 = it.k0, k1 = it.k1;
    const double* __restrict__ Q = a.panel + it.poff;
    const int* __restrict__ rows = a.rows + it.aux;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int rho = lane % RP;
    const int q = q0 + CW * w + lane / RP;
    const bool live = q < C, direct = it.slot0 < 0;
    double pv[PV];
#pragma unroll
    for (int i = 0; i < PV; ++i) {
        const int r = k0 + rho + RP * i;
        pv[i] = live && r < k1 ? Q[(size_t)q * m + r] : 0.0;
    }
    const int pq = live && direct ? a.perm[c0 + q] : 0;
    {
        constexpr int U = stage_depth<KB>();
        const int tot = (k1 - k0) * KB;
        for (int e0 = threadIdx.x; e0 < tot; e0 += NTH * U) {
            int src[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int e = e0 + u * NTH, r = k0 + e / KB;
                if (e < tot) src[u] = r < C ? -1 : rows[r - C];
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int e = e0 + u * NTH, dr = e / KB, p = e - dr * KB, r = k0 + dr;
                if (e < tot)
                    v[dr * LK + p] = p >= nr ? zd() : src[u] < 0 ? a.zp[(size_t)(c0 + r) * nr + p]
                                                                 : a.xp[(size_t)src[u] * nr + p];
            }
        }
    }
    __syncthreads();
    double are[KB], aim[KB];
#pragma unroll
    for (int p = 0; p < KB; ++p) are[p] = aim[p] = 0.0;
    if (live) {
#pragma unroll
        for (int i = 0; i < PV; ++i) {
            const int r = k0 + rho + RP * i;
            if (r < k1) mac_row<KB>(are, aim, pv[i], v + (r - k0) * LK);
        }
    }
#pragma unroll
    for (int p = 0; p < KB; ++p) {
        if (p < nr) {
#pragma unroll
            for (int off = 1; off < RP; off <<= 1) {
                are[p] += __shfl_xor(are[p], off);
                aim[p] += __shfl_xor(aim[p], off);
            }
        }
    }
    if (direct) {
        if (live) {
#pragma unroll
            for (int p = 0; p < KB; ++p) {
                if (p < nr && p % RP == rho) {
                    const zd t(are[p], aim[p]);
                    a.xp[(size_t)(c0 + q) * nr + p] = t;
                    a.x[(size_t)pq * nr + p] = t;
                }
            }
        }
        return;
    }
    // a split tile (RP = 16, 16 columns): partial sums
    zd* slot = a.part + (size_t)(it.slot0 + it.kidx) * SLOT;
    if (live) {
#pragma unroll
        for (int p = 0; p < KB; ++p)
            if (p < nr && p % RP == rho) slot[(q - q0) * KMAX + p] = zd(are[p], aim[p]);
    }
}

template <int KB>
__global__ __launch_bounds__(NTH) __attribute__((amdgpu_waves_per_eu(4))) void schur_bwd(SolveArgs a, int item0,
                                                                                          int nr) {
    extern __shared__ double lds_raw[];
    zd* v = reinterpret_cast<zd*>(lds_raw);
    const SnItem it = a.item[item0 + blockIdx.x];
    if (it.kw == 8) bwd_item<KB, 8>(a, it, nr, v);
    else bwd_item<KB, 16>(a, it, nr, v);
}

// The split tiles of a level: the partial sums of each tile's chunks added in chunk order, then written as the
// forward (rows: z_C, or u_R with the passthrough) or backward (columns: x_C) result. Thread = (tile row or column
// i, right-hand side p).
__global__ __launch_bounds__(NTH) void schur_fin(SolveArgs a, int item0, int nr, int forward) {
    const SnItem it = a.item[item0 + blockIdx.x];
    const int i = threadIdx.x >> 4, p = threadIdx.x & 15, t = it.t0 + i;
    if (p >= nr || t >= (forward ? it.m : it.C)) return;
    const zd* s0 = a.part + (size_t)it.slot0 * SLOT + i * KMAX + p;
    zd acc = s0[0];
    for (int k = 1; k < it.ks; ++k) acc = acc + s0[(size_t)k * SLOT];
    if (forward) {
        if (t < it.C) a.zp[(size_t)(it.c0 + t) * nr + p] = acc;
        else a.ubuf[(size_t)(it.aux + t - it.C) * nr + p] = add_contrib(a, a.ctab[it.fpos + t], p, nr, acc);
    } else {
        a.xp[(size_t)(it.c0 + t) * nr + p] = acc;
        a.x[(size_t)a.perm[it.c0 + t] * nr + p] = acc;
    }
}

// ---- the dense top of the tree
struct TopArgs {
    const TopItem* item;
    const int* perm;
    const int* col;
    const int* lo;
    const int* cptr;
    const int* clen;
    const int* chain;
    const int64_t* foff;
    const int64_t* boff;
    const int64_t* gptr;
    const int64_t* gidx;
    const double* Mf;
    const double* Mb;
    const zd* b;
    zd* x;
    zd* zp;
    zd* xp;
    const zd* ubuf;
    zd* g;
    zd* part;
};

constexpr int TOP_K = 32;            // reduction entries staged in LDS at a time
constexpr int TOP_SLOT = SCHUR_TOP_ROWS * KMAX;      // entries of one partial-sum slot of a dense tile

// g[i] = b at top column i plus the update entries its non-top children send up, in ascending order
__global__ __launch_bounds__(NTH) void schur_top_g(TopArgs a, int nt, int nr) {
    const int e = blockIdx.x * NTH + threadIdx.x;
    const int i = e / nr, p = e - i * nr;
    if (i >= nt) return;
    zd acc = a.b[(size_t)a.perm[a.col[i]] * nr + p];
    for (int64_t q = a.gptr[i]; q < a.gptr[i + 1]; ++q) acc = acc + a.ubuf[(size_t)a.gidx[q] * nr + p];
    a.g[(size_t)i * nr + p] = acc;
}

// Forward: z at top rows [i0, i0 + cnt) (cnt <= SCHUR_TOP_ROWS) = M[rows][k0 .. k1) g[k0 .. k1). Thread (row
// group rg, right-hand side p) holds rows 4 rg .. 4 rg + 3, so each staged g value feeds 8 multiply-adds. g and M go
// through LDS TOP_K entries at a time, double-buffered: the next chunk's loads are in flight while this one's
// multiply-adds run. The sum over k runs in one fixed order for every batch.
constexpr int TOP_R = 4;                      // rows (columns) per thread
__global__ __launch_bounds__(NTH) void schur_top_fwd(TopArgs a, int nr) {
    __shared__ double gs[2][TOP_K][KMAX][2];
    __shared__ double ms[2][TOP_K][SCHUR_TOP_ROWS + 1];
    __shared__ int64_t offs[SCHUR_TOP_ROWS];
    const TopItem it = a.item[blockIdx.x];
    const int rg = threadIdx.x >> 4, p = threadIdx.x & 15;
    const int lo = a.lo[it.i0];
    if (threadIdx.x < SCHUR_TOP_ROWS) offs[threadIdx.x] = threadIdx.x < it.cnt ? a.foff[it.i0 + threadIdx.x] : 0;
    __syncthreads();
    constexpr int GPT = TOP_K * KMAX / NTH, MPT = SCHUR_TOP_ROWS * TOP_K / NTH;   // staged entries per thread
    zd gl[GPT];
    double ml[MPT];
    auto load = [&](int k0) {
        const int len = min(TOP_K, it.k1 - k0);
#pragma unroll
        for (int u = 0; u < GPT; ++u) {
            const int e = threadIdx.x + u * NTH, kk = e >> 4, q = e & 15;
            gl[u] = q < nr && kk < len ? a.g[(size_t)(k0 + kk) * nr + q] : zd();
        }
#pragma unroll
        for (int u = 0; u < MPT; ++u) {         // row rr's run of k: consecutive threads, consecutive k
            const int e = threadIdx.x + u * NTH, rr = e / TOP_K, kk = e - rr * TOP_K, k = k0 + kk;
            ml[u] = rr < it.cnt && kk < len && k <= it.i0 + rr ? a.Mf[offs[rr] + (k - lo)] : 0.0;
        }
    };
    auto store = [&](int b) {
#pragma unroll
        for (int u = 0; u < GPT; ++u) {
            const int e = threadIdx.x + u * NTH, kk = e >> 4, q = e & 15;
            gs[b][kk][q][0] = gl[u].re;
            gs[b][kk][q][1] = gl[u].im;
        }
#pragma unroll
        for (int u = 0; u < MPT; ++u) {
            const int e = threadIdx.x + u * NTH, rr = e / TOP_K, kk = e - rr * TOP_K;
            ms[b][kk][rr] = ml[u];
        }
    };
    double are[TOP_R] = {}, aim[TOP_R] = {};
    load(it.k0);
    store(0);
    __syncthreads();
    int b = 0;
    for (int k0 = it.k0; k0 < it.k1; k0 += TOP_K) {
        const bool more = k0 + TOP_K < it.k1;
        if (more) load(k0 + TOP_K);
        const int len = min(TOP_K, it.k1 - k0);
        for (int kk = 0; kk < len; ++kk) {
            const double gr = gs[b][kk][p][0], gi = gs[b][kk][p][1];
#pragma unroll
            for (int u = 0; u < TOP_R; ++u) {
                const double m = ms[b][kk][TOP_R * rg + u];
                are[u] = fma(m, gr, are[u]);
                aim[u] = fma(m, gi, aim[u]);
            }
        }
        if (more) store(b ^ 1);
        __syncthreads();
        b ^= 1;
    }
    if (p >= nr) return;
#pragma unroll
    for (int u = 0; u < TOP_R; ++u) {
        const int r = TOP_R * rg + u;
        if (r >= it.cnt) break;
        if (it.slot0 < 0) a.zp[(size_t)a.col[it.i0 + r] * nr + p] = zd(are[u], aim[u]);
        else a.part[(size_t)(it.slot0 + it.kidx) * TOP_SLOT + r * KMAX + p] = zd(are[u], aim[u]);
    }
}

// Backward: x at top columns [i0, i0 + cnt) = M[chain[k0 .. k1)][cols]^T z[chain[k0 .. k1)], the columns of one
// supernode (one chain); thread (column group, p) holds 4 columns. The chain's rows as columns of zp (ccol).
__global__ __launch_bounds__(NTH) void schur_top_bwd(TopArgs a, int nr) {
    __shared__ double zs[2][TOP_K][KMAX][2];
    __shared__ double ms[2][TOP_K][SCHUR_TOP_ROWS + 1];
    __shared__ int64_t offs[SCHUR_TOP_ROWS];
    __shared__ int ccol[SCHUR_TOP_CHUNK];       // the item's chain rows [k0, k1), as columns of zp
    const TopItem it = a.item[blockIdx.x];
    const int cg = threadIdx.x >> 4, p = threadIdx.x & 15;
    const int* chain = a.chain + a.cptr[it.i0];
    if (threadIdx.x < SCHUR_TOP_ROWS) offs[threadIdx.x] = threadIdx.x < it.cnt ? a.boff[it.i0 + threadIdx.x] : 0;
    for (int c = threadIdx.x; c < it.k1 - it.k0; c += NTH) ccol[c] = chain[it.k0 + c];
    __syncthreads();
    constexpr int GPT = TOP_K * KMAX / NTH, MPT = SCHUR_TOP_ROWS * TOP_K / NTH;
    zd zl[GPT];
    double ml[MPT];
    auto load = [&](int k0) {
        const int len = min(TOP_K, it.k1 - k0);
#pragma unroll
        for (int u = 0; u < GPT; ++u) {
            const int e = threadIdx.x + u * NTH, kk = e >> 4, q = e & 15;
            zl[u] = q < nr && kk < len ? a.zp[(size_t)ccol[k0 - it.k0 + kk] * nr + q] : zd();
        }
#pragma unroll
        for (int u = 0; u < MPT; ++u) {
            const int e = threadIdx.x + u * NTH, cc = e / TOP_K, kk = e - cc * TOP_K;
            ml[u] = cc < it.cnt && kk < len ? a.Mb[offs[cc] + k0 + kk] : 0.0;
        }
    };
    auto store = [&](int b) {
#pragma unroll
        for (int u = 0; u < GPT; ++u) {
            const int e = threadIdx.x + u * NTH, kk = e >> 4, q = e & 15;
            zs[b][kk][q][0] = zl[u].re;
            zs[b][kk][q][1] = zl[u].im;
        }
#pragma unroll
        for (int u = 0; u < MPT; ++u) {
            const int e = threadIdx.x + u * NTH, cc = e / TOP_K, kk = e - cc * TOP_K;
            ms[b][kk][cc] = ml[u];
        }
    };
    double are[TOP_R] = {}, aim[TOP_R] = {};
    load(it.k0);
    store(0);
    __syncthreads();
    int b = 0;
    for (int k0 = it.k0; k0 < it.k1; k0 += TOP_K) {
        const bool more = k0 + TOP_K < it.k1;
        if (more) load(k0 + TOP_K);
        const int len = min(TOP_K, it.k1 - k0);
        for (int kk = 0; kk < len; ++kk) {
            const double zr = zs[b][kk][p][0], zi = zs[b][kk][p][1];
#pragma unroll
            for (int u = 0; u < TOP_R; ++u) {
                const double m = ms[b][kk][TOP_R * cg + u];
                are[u] = fma(m, zr, are[u]);
                aim[u] = fma(m, zi, aim[u]);
            }
        }
        if (more) store(b ^ 1);
        __syncthreads();
        b ^= 1;
    }
    if (p >= nr) return;
#pragma unroll
    for (int u = 0; u < TOP_R; ++u) {
        const int c = TOP_R * cg + u;
        if (c >= it.cnt) break;
        const int col = a.col[it.i0 + c];
        if (it.slot0 < 0) {
            a.xp[(size_t)col * nr + p] = zd(are[u], aim[u]);
            a.x[(size_t)a.perm[col] * nr + p] = zd(are[u], aim[u]);
        } else {
            a.part[(size_t)(it.slot0 + it.kidx) * TOP_SLOT + c * KMAX + p] = zd(are[u], aim[u]);
        }
    }
}

// The split dense tiles: partial sums added in chunk order, then written as the forward (z) or backward (x) result;
// thread (row or column r, right-hand side p)
__global__ __launch_bounds__(SCHUR_TOP_ROWS * KMAX) void schur_top_fin(TopArgs a, int nr, int forward) {
    const TopItem it = a.item[blockIdx.x];
    const int r = threadIdx.x >> 4, p = threadIdx.x & 15;
    if (r >= it.cnt || p >= nr) return;
    const zd* s0 = a.part + (size_t)it.slot0 * TOP_SLOT + r * KMAX + p;
    zd acc = s0[0];
    for (int k0 = 1; k0 < it.ks; k0 += 16) {    // sixteen loads in flight, added in chunk order
        zd t[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) t[u] = k0 + u < it.ks ? s0[(size_t)(k0 + u) * TOP_SLOT] : zd();
#pragma unroll
        for (int u = 0; u < 16; ++u)
            if (k0 + u < it.ks) acc = acc + t[u];
    }
    const int col = a.col[it.i0 + r];
    if (forward) {
        a.zp[(size_t)col * nr + p] = acc;
    } else {
        a.xp[(size_t)col * nr + p] = acc;
        a.x[(size_t)a.perm[col] * nr + p] = acc;
    }
}

TopArgs top_args(const Schur& S, const zd* b, zd* x) {
    return TopArgs{nullptr, S.perm.p, S.top_col.p, S.top_lo.p, S.top_cptr.p, S.top_clen.p, S.top_chain.p,
                   S.top_foff.p, S.top_boff.p, S.top_gptr.p, S.top_gidx.p, S.top_Mf.p, S.top_Mb.p, b, x, S.zp.p,
                   S.xp.p, S.ubuf.p, S.gbuf.p, S.top_part.p};
}

void run_top_fwd(const Schur& S, const zd* b, int nr, hipStream_t st) {
    const int nt = (int)S.h.top_col.size();
    TopArgs a = top_args(S, b, nullptr);
    hipLaunchKernelGGL(schur_top_g, dim3(blocks_for((size_t)nt * nr, NTH)), dim3(NTH), 0, st, a, nt, nr);
    GT_LAUNCH("schur_top_g");
    a.item = S.top_fwd.p;
    hipLaunchKernelGGL(schur_top_fwd, dim3((unsigned)S.h.top_fwd.size()), dim3(NTH), 0, st, a, nr);
    GT_LAUNCH("schur_top_fwd");
    if (!S.h.top_ffin.empty()) {
        a.item = S.top_ffin.p;
        hipLaunchKernelGGL(schur_top_fin, dim3((unsigned)S.h.top_ffin.size()), dim3(SCHUR_TOP_ROWS * KMAX), 0, st, a,
                           nr, 1);
        GT_LAUNCH("schur_top_fin");
    }
}

void run_top_bwd(const Schur& S, zd* x, int nr, hipStream_t st) {
    TopArgs a = top_args(S, nullptr, x);
    a.item = S.top_bwd.p;
    hipLaunchKernelGGL(schur_top_bwd, dim3((unsigned)S.h.top_bwd.size()), dim3(NTH), 0, st, a, nr);
    GT_LAUNCH("schur_top_bwd");
    if (!S.h.top_bfin.empty()) {
        a.item = S.top_bfin.p;
        hipLaunchKernelGGL(schur_top_fin, dim3((unsigned)S.h.top_bfin.size()), dim3(SCHUR_TOP_ROWS * KMAX), 0, st, a,
                           nr, 0);
        GT_LAUNCH("schur_top_fin");
    }
}

// the forward levels [l0, l1) of `ptr` / `items` (the tree below the dense top, or the top's own supernodes)
template <int KB>
void run_fwd_items(const Schur& S, SolveArgs a, int nr, hipStream_t st, const std::vector<int>& fp,
                   const std::vector<int>& cols, const std::vector<int>& finp, const SnItem* items, const SnItem* fins,
                   size_t l0 = 0, size_t l1 = SIZE_MAX) {
    constexpr size_t row_bytes = (KB + 1) * sizeof(zd);
    a.item = items;
    for (size_t l = l0; l + 1 < fp.size() && l < l1; ++l) {
        const int cnt = fp[l + 1] - fp[l];
        if (cnt == 0) continue;
        hipLaunchKernelGGL(schur_fwd<KB>, dim3(cnt), dim3(NTH), cols[l] * row_bytes, st, a, fp[l], nr);
        GT_LAUNCH("schur_fwd");
        const int nf = finp[l + 1] - finp[l];
        if (nf > 0) {
            SolveArgs af = a;
            af.item = fins;
            hipLaunchKernelGGL(schur_fin, dim3(nf), dim3(NTH), 0, st, af, finp[l], nr, 1);
            GT_LAUNCH("schur_fin");
        }
    }
}

template <int KB>
void run_fwd(const Schur& S, SolveArgs a, int nr, hipStream_t st) {
    run_fwd_items<KB>(S, a, nr, st, S.h.fwd_ptr, S.h.fwd_cols, S.h.ffin_ptr, S.fwd.p, S.ffin.p);
    if (S.h.top_T) run_top_fwd(S, a.b, nr, st);
}

template <int KB>
void run_bwd(const Schur& S, SolveArgs a, int nr, hipStream_t st) {
    constexpr size_t row_bytes = (KB + 1) * sizeof(zd);
    if (S.h.top_T) run_top_bwd(S, a.x, nr, st);
    const std::vector<int>& bp = S.h.bwd_ptr;
    a.item = S.bwd.p;
    for (size_t l = 0; l + 1 < bp.size(); ++l) {
        const int cnt = bp[l + 1] - bp[l];
        if (cnt == 0) continue;
        hipLaunchKernelGGL(schur_bwd<KB>, dim3(cnt), dim3(NTH), S.h.bwd_rows[l] * row_bytes, st, a, bp[l], nr);
        GT_LAUNCH("schur_bwd");
        const int nf = S.h.bfin_ptr[l + 1] - S.h.bfin_ptr[l];
        if (nf > 0) {
            SolveArgs af = a;
            af.item = S.bfin.p;
            hipLaunchKernelGGL(schur_fin, dim3(nf), dim3(NTH), 0, st, af, S.h.bfin_ptr[l], nr, 0);
            GT_LAUNCH("schur_fin");
        }
    }
}

}  // namespace

void Schur::upload(hipStream_t st) {
    auto nz = [](const std::vector<int>& v) { return v.empty() ? std::vector<int>{0} : v; };
    perm.upload(nz(h.perm), st);
    rows.upload(nz(h.rows), st);
    ctab.upload(h.ctab.empty() ? std::vector<int32_t>{-1, -1} : h.ctab, st);
    cidx.upload(h.cidx.empty() ? std::vector<int64_t>{0} : h.cidx, st);
    fwd.upload(h.fwd.empty() ? std::vector<SnItem>(1) : h.fwd, st);
    bwd.upload(h.bwd.empty() ? std::vector<SnItem>(1) : h.bwd, st);
    // the host fronts' panels (the device fronts' are written on the device): runs of consecutive host fronts
    panel.alloc(std::max<size_t>(1, h.panel.size()));
    if (h.panel.empty()) {
        panel.upload(std::vector<double>{0.0}, st);
    } else if (h.dev_order.empty()) {
        panel.upload(h.panel.data(), h.panel.size(), st);
    } else {
        const int ns = (int)h.nc.size();
        for (int s = 0; s < ns;) {
            if (h.on_dev[s]) {
                ++s;
                continue;
            }
            int e = s;
            while (e < ns && !h.on_dev[e]) ++e;
            const int64_t a = h.poff[s], b = h.poff[e];
            if (b > a)
                GT_HIP(hipMemcpyAsync(panel.p + a, h.panel.data() + a, (size_t)(b - a) * sizeof(double),
                                      hipMemcpyHostToDevice, st));
            s = e;
        }
        GT_HIP(hipStreamSynchronize(st));
    }
    part.alloc((size_t)std::max<int64_t>(1, h.slots) * SCHUR_TILE * KMAX);
    ffin.upload(h.ffin.empty() ? std::vector<SnItem>(1) : h.ffin, st);
    bfin.upload(h.bfin.empty() ? std::vector<SnItem>(1) : h.bfin, st);
    if (h.top_T) {
        top_col.upload(h.top_col, st);
        top_lo.upload(h.top_lo, st);
        top_cptr.upload(h.top_cptr, st);
        top_clen.upload(h.top_clen, st);
        top_chain.upload(h.top_chain, st);
        top_foff.upload(h.top_foff, st);
        top_boff.upload(h.top_boff, st);
        top_gptr.upload(h.top_gptr, st);
        top_gidx.upload(h.top_gidx.empty() ? std::vector<int64_t>{0} : h.top_gidx, st);
        top_fwd.upload(h.top_fwd, st);
        top_ffin.upload(h.top_ffin.empty() ? std::vector<TopItem>(1) : h.top_ffin, st);
        top_bwd.upload(h.top_bwd, st);
        top_bfin.upload(h.top_bfin.empty() ? std::vector<TopItem>(1) : h.top_bfin, st);
        top_part.alloc((size_t)std::max<int64_t>(1, h.top_slots) * SCHUR_TOP_ROWS * KMAX);
        tfwd.upload(h.tfwd.empty() ? std::vector<SnItem>(1) : h.tfwd, st);
        tffin.upload(h.tffin.empty() ? std::vector<SnItem>(1) : h.tffin, st);
    }
    cap = 0;
    ensure(1);
    GT_HIP(hipStreamSynchronize(st));
}

void Schur::ensure(int nr) {
    if (nr <= cap) return;
    if (!graphs.empty()) GT_HIP(hipDeviceSynchronize());    // no captured solve still running on the old buffers
    for (SchurGraph& g : graphs) (void)hipGraphExecDestroy(g.exec);
    graphs.clear();
    ubuf.release();
    zp.release();
    xp.release();
    gbuf.release();
    ubuf.alloc((size_t)std::max<int64_t>(1, h.upd_total) * nr);
    zp.alloc((size_t)std::max(1, h.n) * nr);
    xp.alloc((size_t)std::max(1, h.n) * nr);
    if (h.top_T) gbuf.alloc(h.top_col.size() * (size_t)nr);
    cap = nr;
}

void Schur::shrink(int nr) {
    if (nr >= cap) return;
    cap = 0;
    ensure(std::max(1, nr));
}

int64_t Schur::bytes_after(int nr) const {
    const int c = std::max(cap, nr);
    return (int64_t)sizeof(zd) * c *
           (std::max<int64_t>(1, h.upd_total) + 2 * (int64_t)std::max(1, h.n) + (h.top_T ? (int64_t)h.top_col.size() : 0));
}

int64_t Schur::device_bytes() const {
    return (int64_t)(perm.bytes() + rows.bytes() + ctab.bytes() + cidx.bytes() + fwd.bytes() + bwd.bytes() +
                     panel.bytes() + part.bytes() + ffin.bytes() + bfin.bytes() + ubuf.bytes() + zp.bytes() + xp.bytes() +
                     top_col.bytes() + top_lo.bytes() + top_cptr.bytes() + top_clen.bytes() + top_chain.bytes() +
                     top_foff.bytes() + top_boff.bytes() + top_gptr.bytes() + top_gidx.bytes() + top_Mf.bytes() +
                     top_Mb.bytes() + top_fwd.bytes() + top_ffin.bytes() + top_bwd.bytes() + top_bfin.bytes() +
                     top_part.bytes() + gbuf.bytes() + tfwd.bytes() + tffin.bytes());
}

namespace {

// the passes for nr right-hand sides: forward (pass 1), backward (pass 2) or both (3), with the kernels' batch width
// 1, 2, 4, 8, 12 or 16 (the least that holds nr: the padded slots are multiply-adds of zeros)
void launch_passes(Schur& S, const zd* b, zd* x, int nr, hipStream_t st, int passes) {
    SolveArgs a{nullptr, S.perm.p, S.rows.p, reinterpret_cast<const int2*>(S.ctab.p), S.cidx.p, S.panel.p, S.ubuf.p,
                S.zp.p, S.xp.p, b, x, S.part.p};
    auto go = [&](auto kb) {
        constexpr int KB = decltype(kb)::value;
        if (passes & 1) run_fwd<KB>(S, a, nr, st);
        if (passes & 2) run_bwd<KB>(S, a, nr, st);
    };
    if (nr == 1) go(std::integral_constant<int, 1>());
    else if (nr == 2) go(std::integral_constant<int, 2>());
    else if (nr <= 4) go(std::integral_constant<int, 4>());
    else if (nr <= 8) go(std::integral_constant<int, 8>());
    else if (nr <= 12) go(std::integral_constant<int, 12>());
    else go(std::integral_constant<int, 16>());
}

void launch_levels(Schur& S, const zd* b, zd* x, int nr, hipStream_t st) { launch_passes(S, b, x, nr, st, 3); }

}  // namespace

// With graphs, the level launches of a (nr, b, x) are captured once into a hipGraph and replayed: the same kernels
// with the same arguments, so the same bits, and one submission from the host instead of one per level. The
// sequence holds kernels only (no memset nodes). At most SCHUR_GRAPHS are kept, the oldest dropped first.
void schur_solve_device(Schur& S, const zd* b, zd* x, int nr, hipStream_t st, bool graph) {
    if (nr < 1 || nr > KMAX) fail(GT_ERR_ARG, "Schur solve: 1 to 16 right-hand sides");
    S.ensure(nr);
    if (!graph) {
        launch_levels(S, b, x, nr, st);
        return;
    }
    for (SchurGraph& g : S.graphs)
        if (g.nr == nr && g.b == b && g.x == x) {
            GT_HIP(hipGraphLaunch(g.exec, st));
            return;
        }
    hipGraph_t gr = nullptr;
    GT_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    try {
        launch_levels(S, b, x, nr, st);
    } catch (...) {
        (void)hipStreamEndCapture(st, &gr);
        if (gr) (void)hipGraphDestroy(gr);
        throw;
    }
    GT_HIP(hipStreamEndCapture(st, &gr));
    hipGraphExec_t exec = nullptr;
    const hipError_t e = hipGraphInstantiate(&exec, gr, nullptr, nullptr, 0);
    (void)hipGraphDestroy(gr);
    GT_HIP(e);
    if (S.graphs.size() >= SCHUR_GRAPHS) {
        (void)hipGraphExecDestroy(S.graphs.front().exec);
        S.graphs.erase(S.graphs.begin());
    }
    S.graphs.push_back({nr, b, x, exec});
    GT_HIP(hipGraphLaunch(exec, st));
}

void schur_forward_device(Schur& S, const zd* b, int nr, hipStream_t st) {
    if (nr < 1 || nr > KMAX) fail(GT_ERR_ARG, "Schur solve: 1 to 16 right-hand sides");
    S.ensure(nr);
    launch_passes(S, b, nullptr, nr, st, 1);
}

namespace {

__global__ __launch_bounds__(NTH) void upd_zero(zd* v, size_t n) {
    for (size_t e = blockIdx.x * (size_t)NTH + threadIdx.x; e < n; e += (size_t)gridDim.x * NTH) v[e] = zd();
}

}  // namespace

void schur_forward_last_device(Schur& S, const zd* b, int nr, hipStream_t st) {
    if (nr < 1 || nr > KMAX) fail(GT_ERR_ARG, "Schur solve: 1 to 16 right-hand sides");
    const int ns = (int)S.h.nc.size();
    if (ns == 0) return;
    S.ensure(nr);
    // every supernode but the last has a zero front: its update entries are cleared, not computed
    hipLaunchKernelGGL(upd_zero, dim3(1024), dim3(NTH), 0, st, S.ubuf.p, S.ubuf.n);
    GT_LAUNCH("upd_zero");
    if (S.h.top_T) {            // the last supernode is the top's root: the dense top alone
        run_top_fwd(S, b, nr, st);
        return;
    }
    size_t lvl = SIZE_MAX;      // the level holding the last supernode
    const std::vector<int>& fp = S.h.fwd_ptr;
    for (size_t l = 0; l + 1 < fp.size() && lvl == SIZE_MAX; ++l)
        for (int i = fp[l]; i < fp[l + 1]; ++i)
            if (S.h.fwd[i].s == ns - 1) {
                lvl = l;
                break;
            }
    if (lvl == SIZE_MAX) fail(GT_ERR_ARG, "the last supernode has no forward items");
    SolveArgs a{nullptr, S.perm.p, S.rows.p, reinterpret_cast<const int2*>(S.ctab.p), S.cidx.p, S.panel.p, S.ubuf.p,
                S.zp.p, S.xp.p, b, nullptr, S.part.p};
    auto go = [&](auto kb) {
        constexpr int KB = decltype(kb)::value;
        run_fwd_items<KB>(S, a, nr, st, fp, S.h.fwd_cols, S.h.ffin_ptr, S.fwd.p, S.ffin.p, lvl, lvl + 1);
    };
    if (nr == 1) go(std::integral_constant<int, 1>());
    else if (nr == 2) go(std::integral_constant<int, 2>());
    else if (nr <= 4) go(std::integral_constant<int, 4>());
    else if (nr <= 8) go(std::integral_constant<int, 8>());
    else if (nr <= 12) go(std::integral_constant<int, 12>());
    else go(std::integral_constant<int, 16>());
}

void schur_backward_device(Schur& S, zd* x, int nr, hipStream_t st) {
    if (nr < 1 || nr > KMAX) fail(GT_ERR_ARG, "Schur solve: 1 to 16 right-hand sides");
    if (nr > S.cap) fail(GT_ERR_ARG, "Schur backward pass without its forward pass");
    launch_passes(S, nullptr, x, nr, st, 2);
}

namespace {

// unit right-hand sides for M's columns k0 .. k0 + nb of the dense top (16 per batch, old numbering); the previous
// batch's units cleared first (kernels, not memsets)
__global__ __launch_bounds__(64) void top_units(zd* b, const int* perm, const int* col, int k_prev, int nb_prev, int k0,
                                                int nb) {
    const int p = threadIdx.x;
    if (p < nb_prev) b[(size_t)perm[col[k_prev + p]] * KMAX + p] = zd();
    __syncthreads();
    if (p < nb) b[(size_t)perm[col[k0 + p]] * KMAX + p] = zd(1.0);
}

__global__ __launch_bounds__(NTH) void top_zero(zd* v, size_t n) {
    for (size_t e = blockIdx.x * (size_t)NTH + threadIdx.x; e < n; e += (size_t)gridDim.x * NTH) v[e] = zd();
}

// M[i][k0 + p] = z at top column i for unit right-hand side p (real: the panels and units are)
__global__ __launch_bounds__(NTH) void top_extract(double* M, const zd* zp, const int* col, int nt, int k0, int nb) {
    const int e = blockIdx.x * NTH + threadIdx.x;
    const int i = e / KMAX, p = e - i * KMAX;
    if (i >= nt || p >= nb) return;
    M[(size_t)i * nt + k0 + p] = zp[(size_t)col[i] * KMAX + p].re;
}

}  // namespace

void Schur::top_device(hipStream_t st) {
    if (!h.top_T) return;
    const int nt = (int)h.top_col.size();
    ensure(KMAX);
    DevBuf<zd> bu((size_t)std::max(1, h.n) * KMAX);
    DevBuf<double> Mfull((size_t)nt * nt);
    const unsigned zb = 1024;
    hipLaunchKernelGGL(top_zero, dim3(zb), dim3(NTH), 0, st, bu.p, bu.n);
    hipLaunchKernelGGL(top_zero, dim3(zb), dim3(NTH), 0, st, ubuf.p, ubuf.n);     // the non-top children's entries
    GT_LAUNCH("top_zero");
    SolveArgs a{nullptr, perm.p, rows.p, reinterpret_cast<const int2*>(ctab.p), cidx.p, panel.p, ubuf.p, zp.p, xp.p,
                bu.p, nullptr, part.p};
    int k_prev = 0, nb_prev = 0;
    for (int k0 = 0; k0 < nt; k0 += KMAX) {
        const int nb = std::min(KMAX, nt - k0);
        hipLaunchKernelGGL(top_units, dim3(1), dim3(64), 0, st, bu.p, perm.p, top_col.p, k_prev, nb_prev, k0, nb);
        GT_LAUNCH("top_units");
        run_fwd_items<KMAX>(*this, a, KMAX, st, h.tfwd_ptr, h.tfwd_cols, h.tffin_ptr, tfwd.p, tffin.p);
        hipLaunchKernelGGL(top_extract, dim3(blocks_for((size_t)nt * KMAX, NTH)), dim3(NTH), 0, st, Mfull.p, zp.p,
                           top_col.p, nt, k0, nb);
        GT_LAUNCH("top_extract");
        k_prev = k0;
        nb_prev = nb;
    }
    std::vector<double> M((size_t)nt * nt);
    GT_HIP(hipMemcpyAsync(M.data(), Mfull.p, M.size() * sizeof(double), hipMemcpyDeviceToHost, st));
    GT_HIP(hipStreamSynchronize(st));
    std::vector<double> mf(h.top_fnnz), mb(h.top_bnnz);
    std::vector<int> idx(std::max(1, h.n), -1);               // new column -> top-column index
    for (int i = 0; i < nt; ++i) idx[h.top_col[i]] = i;
    for (int i = 0; i < nt; ++i) {
        for (int k = h.top_lo[i]; k <= i; ++k) mf[h.top_foff[i] + (k - h.top_lo[i])] = M[(size_t)i * nt + k];
        const int* ch = h.top_chain.data() + h.top_cptr[i];
        for (int e = 0; e < h.top_clen[i]; ++e) mb[h.top_boff[i] + e] = M[(size_t)idx[ch[e]] * nt + i];
    }
    top_Mf.upload(mf, st);
    top_Mb.upload(mb, st);
    shrink(1);
}

Schur::~Schur() {
    for (SchurGraph& g : graphs) (void)hipGraphExecDestroy(g.exec);
}

}  // namespace gt
