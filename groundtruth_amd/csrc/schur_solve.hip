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
__global__ __launch_bounds__(NTH) __attribute__((amdgpu_waves_per_eu(4))) void schur_fwd(SolveArgs a, int item0,
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
        constexpr int U = stage_depth<KB>();
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

template <int KB>
void run_fwd(const Schur& S, SolveArgs a, int nr, hipStream_t st) {
    constexpr size_t row_bytes = (KB + 1) * sizeof(zd);
    const std::vector<int>& fp = S.h.fwd_ptr;
    a.item = S.fwd.p;
    for (size_t l = 0; l + 1 < fp.size(); ++l) {
        const int cnt = fp[l + 1] - fp[l];
        if (cnt == 0) continue;
        hipLaunchKernelGGL(schur_fwd<KB>, dim3(cnt), dim3(NTH), S.h.fwd_cols[l] * row_bytes, st, a, fp[l], nr);
        GT_LAUNCH("schur_fwd");
        const int nf = S.h.ffin_ptr[l + 1] - S.h.ffin_ptr[l];
        if (nf > 0) {
            SolveArgs af = a;
            af.item = S.ffin.p;
            hipLaunchKernelGGL(schur_fin, dim3(nf), dim3(NTH), 0, st, af, S.h.ffin_ptr[l], nr, 1);
            GT_LAUNCH("schur_fin");
        }
    }
}

template <int KB>
void run_bwd(const Schur& S, SolveArgs a, int nr, hipStream_t st) {
    constexpr size_t row_bytes = (KB + 1) * sizeof(zd);
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
    panel.upload(h.panel.empty() ? std::vector<double>{0.0} : h.panel, st);
    part.alloc((size_t)std::max<int64_t>(1, h.slots) * SCHUR_TILE * KMAX);
    ffin.upload(h.ffin.empty() ? std::vector<SnItem>(1) : h.ffin, st);
    bfin.upload(h.bfin.empty() ? std::vector<SnItem>(1) : h.bfin, st);
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
    ubuf.alloc((size_t)std::max<int64_t>(1, h.upd_total) * nr);
    zp.alloc((size_t)std::max(1, h.n) * nr);
    xp.alloc((size_t)std::max(1, h.n) * nr);
    cap = nr;
}

void Schur::shrink(int nr) {
    if (nr >= cap) return;
    cap = 0;
    ensure(std::max(1, nr));
}

int64_t Schur::bytes_after(int nr) const {
    const int c = std::max(cap, nr);
    return (int64_t)sizeof(zd) * c * (std::max<int64_t>(1, h.upd_total) + 2 * (int64_t)std::max(1, h.n));
}

int64_t Schur::device_bytes() const {
    return (int64_t)(perm.bytes() + rows.bytes() + ctab.bytes() + cidx.bytes() + fwd.bytes() + bwd.bytes() +
                     panel.bytes() + part.bytes() + ffin.bytes() + bfin.bytes() + ubuf.bytes() + zp.bytes() + xp.bytes());
}

namespace {

// the passes for nr right-hand sides: forward (pass 1), backward (pass 2) or both (3), with the kernels' batch width
// the next power of two
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

void schur_backward_device(Schur& S, zd* x, int nr, hipStream_t st) {
    if (nr < 1 || nr > KMAX) fail(GT_ERR_ARG, "Schur solve: 1 to 16 right-hand sides");
    if (nr > S.cap) fail(GT_ERR_ARG, "Schur backward pass without its forward pass");
    launch_passes(S, nullptr, x, nr, st, 2);
}

Schur::~Schur() {
    for (SchurGraph& g : graphs) (void)hipGraphExecDestroy(g.exec);
}

}  // namespace gt
