// Level-scheduled supernodal triangular solves of B = L L^T on the device (the Schur block of the preconditioner,
// peec.py:3064-3083 `op.solve_S`).
//
// Forward (L z = b), leaves first, one launch per level of the assembly tree and one workgroup per supernode:
// the front vector [b at the supernode's columns ; 0 at its off-diagonal rows] gathers its children's update parts
// (extend-add, children in a fixed order, so the sums are the same bits every run), solves the dense diagonal block
// in 64-column blocks (one wavefront per triangle, lanes exchanging through shuffles) and updates the rest of the
// front with the panel; what is left at the off-diagonal rows is the parent's update.
// Backward (L^T x = z), roots first: each supernode takes its ancestors' x at its off-diagonal rows, then solves the
// transposed diagonal block from its last 64-column block up.
// The real factor multiplies complex vectors (the right-hand side's real and imaginary parts are two columns).
#include "schur.hpp"

namespace gt {

namespace {

struct SolveArgs {
    const int* perm;
    const int* c0;
    const int* nc;
    const int* nr;
    const int* child_ptr;
    const int* child;
    const int* rows;
    const int* rel;
    const int* levels;
    const int64_t* roff;
    const int64_t* poff;
    const int64_t* rel_off;
    const int64_t* foff;
    const double* panel;
    zd* fvec;
    zd* xp;
    const zd* b;
    zd* x;
};

constexpr int NTH = 256;
constexpr int BLK = 64;

__device__ __forceinline__ zd shfl(zd v, int lane) {
    return zd(__shfl(v.re, lane, 64), __shfl(v.im, lane, 64));
}

__global__ __launch_bounds__(NTH) void schur_fwd(SolveArgs a, int lvl0) {
    __shared__ double xs_raw[2 * BLK];
    zd* xs = reinterpret_cast<zd*>(xs_raw);
    const int s = a.levels[lvl0 + blockIdx.x];
    const int C = a.nc[s], R = a.nr[s], m = C + R, c0 = a.c0[s];
    zd* fv = a.fvec + a.foff[s];
    const double* L = a.panel + a.poff[s];
    const int tid = threadIdx.x;
    for (int q = tid; q < m; q += NTH) fv[q] = q < C ? a.b[a.perm[c0 + q]] : zd();
    __syncthreads();
    for (int e = a.child_ptr[s]; e < a.child_ptr[s + 1]; ++e) {
        const int k = a.child[e];
        const zd* u = a.fvec + a.foff[k] + a.nc[k];
        const int* rel = a.rel + a.rel_off[k];
        const int rk = a.nr[k];
        for (int i = tid; i < rk; i += NTH) {
            const int p = rel[i];
            fv[p] = fv[p] + u[i];
        }
        __syncthreads();
    }
    for (int blk = 0; blk < C; blk += BLK) {
        const int nb = min(BLK, C - blk);
        if (tid < BLK) {
            zd v = tid < nb ? fv[blk + tid] : zd();
            for (int q = 0; q < nb; ++q) {
                if (tid == q) {
                    const double d = L[(size_t)(blk + q) * m + blk + q];
                    v = zd(v.re / d, v.im / d);
                }
                const zd xq = shfl(v, q);
                if (tid > q && tid < nb) {
                    const double l = L[(size_t)(blk + q) * m + blk + tid];
                    v = zd(v.re - l * xq.re, v.im - l * xq.im);
                }
            }
            if (tid < nb) {
                xs[tid] = v;
                fv[blk + tid] = v;
            }
        }
        __syncthreads();
        for (int r = blk + nb + tid; r < m; r += NTH) {
            zd acc = fv[r];
            for (int q = 0; q < nb; ++q) {
                const double l = L[(size_t)(blk + q) * m + r];
                acc.re -= l * xs[q].re;
                acc.im -= l * xs[q].im;
            }
            fv[r] = acc;
        }
        __syncthreads();
    }
    for (int q = tid; q < C; q += NTH) a.xp[c0 + q] = fv[q];
}

__global__ __launch_bounds__(NTH) void schur_bwd(SolveArgs a, int lvl0) {
    __shared__ double ys_raw[2 * BLK];
    zd* ys = reinterpret_cast<zd*>(ys_raw);
    const int s = a.levels[lvl0 + blockIdx.x];
    const int C = a.nc[s], R = a.nr[s], m = C + R, c0 = a.c0[s];
    zd* w = a.fvec + a.foff[s];
    const double* L = a.panel + a.poff[s];
    const int* rows = a.rows + a.roff[s];
    const int tid = threadIdx.x;
    for (int q = tid; q < C; q += NTH) {
        zd acc = a.xp[c0 + q];
        const double* Lq = L + (size_t)q * m + C;
        for (int r = 0; r < R; ++r) {
            const zd xr = a.xp[rows[r]];
            acc.re -= Lq[r] * xr.re;
            acc.im -= Lq[r] * xr.im;
        }
        w[q] = acc;
    }
    __syncthreads();
    const int last = ((C - 1) / BLK) * BLK;
    for (int blk = last; blk >= 0; blk -= BLK) {
        const int nb = min(BLK, C - blk);
        if (tid < BLK) {
            zd v = tid < nb ? w[blk + tid] : zd();
            for (int q = nb - 1; q >= 0; --q) {
                if (tid == q) {
                    const double d = L[(size_t)(blk + q) * m + blk + q];
                    v = zd(v.re / d, v.im / d);
                }
                const zd yq = shfl(v, q);
                if (tid < q) {
                    const double l = L[(size_t)(blk + tid) * m + blk + q];
                    v = zd(v.re - l * yq.re, v.im - l * yq.im);
                }
            }
            if (tid < nb) {
                ys[tid] = v;
                w[blk + tid] = v;
            }
        }
        __syncthreads();
        for (int p = tid; p < blk; p += NTH) {
            zd acc = w[p];
            const double* Lp = L + (size_t)p * m + blk;
            for (int l = 0; l < nb; ++l) {
                acc.re -= Lp[l] * ys[l].re;
                acc.im -= Lp[l] * ys[l].im;
            }
            w[p] = acc;
        }
        __syncthreads();
    }
    for (int q = tid; q < C; q += NTH) {
        a.xp[c0 + q] = w[q];
        a.x[a.perm[c0 + q]] = w[q];
    }
}

}  // namespace

void Schur::upload(hipStream_t st) {
    perm.upload(h.perm, st);
    c0.upload(h.c0, st);
    nc.upload(h.nc, st);
    nr.upload(h.nr, st);
    child_ptr.upload(h.child_ptr, st);
    child.upload(h.child.empty() ? std::vector<int>{0} : h.child, st);
    rows.upload(h.rows.empty() ? std::vector<int>{0} : h.rows, st);
    rel.upload(h.rel.empty() ? std::vector<int>{0} : h.rel, st);
    fwd_levels.upload(h.fwd_levels, st);
    bwd_levels.upload(h.bwd_levels, st);
    roff.upload(h.roff, st);
    poff.upload(h.poff, st);
    rel_off.upload(h.rel_off, st);
    foff.upload(h.foff, st);
    panel.upload(h.panel, st);
    fvec.alloc(h.front_total > 0 ? h.front_total : 1);
    xp.alloc(h.n > 0 ? h.n : 1);
    GT_HIP(hipStreamSynchronize(st));
}

int64_t Schur::device_bytes() const {
    return (int64_t)(perm.bytes() + c0.bytes() + nc.bytes() + nr.bytes() + child_ptr.bytes() + child.bytes() +
                     rows.bytes() + rel.bytes() + fwd_levels.bytes() + bwd_levels.bytes() + roff.bytes() +
                     poff.bytes() + rel_off.bytes() + foff.bytes() + panel.bytes() + fvec.bytes() + xp.bytes());
}

void schur_solve_device(Schur& S, const zd* b, zd* x, hipStream_t st) {
    SolveArgs a{S.perm.p, S.c0.p, S.nc.p, S.nr.p, S.child_ptr.p, S.child.p, S.rows.p, S.rel.p, nullptr,
                S.roff.p, S.poff.p, S.rel_off.p, S.foff.p, S.panel.p, S.fvec.p, S.xp.p, b, x};
    const std::vector<int>& fp = S.h.fwd_levels_ptr;
    a.levels = S.fwd_levels.p;
    for (size_t l = 0; l + 1 < fp.size(); ++l) {
        const int cnt = fp[l + 1] - fp[l];
        if (cnt == 0) continue;
        hipLaunchKernelGGL(schur_fwd, dim3(cnt), dim3(NTH), 0, st, a, fp[l]);
        GT_LAUNCH("schur_fwd");
    }
    const std::vector<int>& bp = S.h.bwd_levels_ptr;
    a.levels = S.bwd_levels.p;
    for (size_t l = 0; l + 1 < bp.size(); ++l) {
        const int cnt = bp[l + 1] - bp[l];
        if (cnt == 0) continue;
        hipLaunchKernelGGL(schur_bwd, dim3(cnt), dim3(NTH), 0, st, a, bp[l]);
        GT_LAUNCH("schur_bwd");
    }
}

}  // namespace gt
