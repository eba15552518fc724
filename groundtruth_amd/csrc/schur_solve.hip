// Level-scheduled supernodal solves of B x = b on the device (the Schur block of the preconditioner,
// peec.py:3064-3083 `op.solve_S`), with each supernode's panel stored as Q = [L_CC^-1 ; -L_RC L_CC^-1]
// (schur_factor.cpp), so that neither direction has a serial chain:
//   forward (leaves first):  f = [b_C ; 0] + the children's update entries (pulled in a fixed order),
//                            [z_C ; u_R] = Q f_C + [0 ; f_R]          -- a dense product over the panel's rows
//   backward (roots first):  x_C = Q^T [z_C ; x_R]                     -- a dense product over its columns
// One launch per level; a large supernode is split over many workgroups (blocks of 64 rows forward, 16 columns
// backward), so the wide top separators run on many CUs. The real panel multiplies complex vectors.
#include "schur.hpp"

namespace gt {

namespace {

constexpr int NTH = 256;
constexpr int CH = 1024;          // vector entries staged in LDS at a time

struct SolveArgs {
    const int* perm;
    const int* c0;
    const int* nc;
    const int* nr;
    const int* rows;
    const int* item_s;
    const int* item_0;
    const int* item_n;
    const int64_t* roff;
    const int64_t* poff;
    const int64_t* uoff;
    const int64_t* fpos;
    const int64_t* cptr;
    const int64_t* cidx;
    const double* panel;
    zd* ubuf;
    zd* zp;
    zd* xp;
    const zd* b;
    zd* x;
};

// f at front position q of supernode s: its right-hand side (a column) plus the children's update entries
__device__ __forceinline__ zd front(const SolveArgs& a, int s, int q, int C, int c0) {
    zd acc = q < C ? a.b[a.perm[c0 + q]] : zd();
    const int64_t pos = a.fpos[s] + q;
    for (int64_t e = a.cptr[pos]; e < a.cptr[pos + 1]; ++e) acc = acc + a.ubuf[a.cidx[e]];
    return acc;
}

// One work item: rows [r0, r0 + 64) of supernode s. Lane = row, the four waves split the columns q (stride 4),
// their partial sums meet in LDS.
__global__ __launch_bounds__(NTH) void schur_fwd(SolveArgs a, int item0) {
    __shared__ double f_raw[2 * CH];
    __shared__ double part_raw[2 * NTH];
    zd* f = reinterpret_cast<zd*>(f_raw);
    zd* part = reinterpret_cast<zd*>(part_raw);
    const int s = a.item_s[item0 + blockIdx.x], r0 = a.item_0[item0 + blockIdx.x];
    const int C = a.nc[s], m = C + a.nr[s], c0 = a.c0[s];
    const double* Q = a.panel + a.poff[s];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r = r0 + lane;
    const int qlim = r < C ? r + 1 : C;
    zd acc;
    for (int cs = 0; cs < C; cs += CH) {
        const int ce = min(C, cs + CH);
        for (int q = cs + threadIdx.x; q < ce; q += NTH) f[q - cs] = front(a, s, q, C, c0);
        __syncthreads();
        if (r < m) {
            const int qe = min(ce, qlim);
            zd a1, a2, a3;
            int q = cs + wave;
            for (; q + 12 < qe; q += 16) {          // four independent loads in flight per lane
                const double w0 = Q[(size_t)q * m + r], w1 = Q[(size_t)(q + 4) * m + r];
                const double w2 = Q[(size_t)(q + 8) * m + r], w3 = Q[(size_t)(q + 12) * m + r];
                const zd v0 = f[q - cs], v1 = f[q + 4 - cs], v2 = f[q + 8 - cs], v3 = f[q + 12 - cs];
                acc.re += w0 * v0.re; acc.im += w0 * v0.im;
                a1.re += w1 * v1.re; a1.im += w1 * v1.im;
                a2.re += w2 * v2.re; a2.im += w2 * v2.im;
                a3.re += w3 * v3.re; a3.im += w3 * v3.im;
            }
            for (; q < qe; q += 4) {
                const double w = Q[(size_t)q * m + r];
                const zd v = f[q - cs];
                acc.re += w * v.re;
                acc.im += w * v.im;
            }
            acc = acc + a1 + a2 + a3;
        }
        __syncthreads();
    }
    part[threadIdx.x] = acc;
    __syncthreads();
    if (wave == 0 && r < m) {
        zd t = part[lane] + part[64 + lane] + part[128 + lane] + part[192 + lane];
        if (r < C) {
            a.zp[c0 + r] = t;
        } else {
            t = t + front(a, s, r, C, c0);
            a.ubuf[a.uoff[s] + r - C] = t;
        }
    }
}

// One work item: columns [q0, q0 + qn) of supernode s (qn <= 64); wave w takes columns q0 + w, q0 + w + 4, ...,
// its lanes stride the rows.
__global__ __launch_bounds__(NTH) void schur_bwd(SolveArgs a, int item0) {
    __shared__ double v_raw[2 * CH];
    zd* v = reinterpret_cast<zd*>(v_raw);
    const int s = a.item_s[item0 + blockIdx.x], q0 = a.item_0[item0 + blockIdx.x];
    const int C = a.nc[s], m = C + a.nr[s], c0 = a.c0[s];
    const int qend = min(C, q0 + a.item_n[item0 + blockIdx.x]);
    const double* Q = a.panel + a.poff[s];
    const int* rows = a.rows + a.roff[s];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    constexpr int K = 16;
    zd acc[K];
    for (int rs = q0; rs < m; rs += CH) {           // rows below the block's first column only
        const int re = min(m, rs + CH);
        for (int r = rs + threadIdx.x; r < re; r += NTH) v[r - rs] = r < C ? a.zp[c0 + r] : a.xp[rows[r - C]];
        __syncthreads();
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const int q = q0 + wave + 4 * k;
            if (q < qend) {
                const double* Qq = Q + (size_t)q * m;
                int r = max(q, rs) + lane;
                zd a1;
                for (; r + 64 < re; r += 128) {
                    const double w0 = Qq[r], w1 = Qq[r + 64];
                    const zd u0 = v[r - rs], u1 = v[r + 64 - rs];
                    acc[k].re += w0 * u0.re; acc[k].im += w0 * u0.im;
                    a1.re += w1 * u1.re; a1.im += w1 * u1.im;
                }
                if (r < re) {
                    const double w0 = Qq[r];
                    const zd u0 = v[r - rs];
                    acc[k].re += w0 * u0.re; acc[k].im += w0 * u0.im;
                }
                acc[k] = acc[k] + a1;
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int q = q0 + wave + 4 * k;
        if (q < qend) {
            double re = acc[k].re, im = acc[k].im;
            for (int off = 32; off > 0; off >>= 1) {
                re += __shfl_down(re, off, 64);
                im += __shfl_down(im, off, 64);
            }
            if (lane == 0) {
                a.xp[c0 + q] = zd(re, im);
                a.x[a.perm[c0 + q]] = zd(re, im);
            }
        }
    }
}

}  // namespace

void Schur::upload(hipStream_t st) {
    auto nz = [](const std::vector<int>& v) { return v.empty() ? std::vector<int>{0} : v; };
    auto nz64 = [](const std::vector<int64_t>& v) { return v.empty() ? std::vector<int64_t>{0} : v; };
    perm.upload(nz(h.perm), st);
    c0.upload(nz(h.c0), st);
    nc.upload(nz(h.nc), st);
    nr.upload(nz(h.nr), st);
    rows.upload(nz(h.rows), st);
    fwd_s.upload(nz(h.fwd_s), st);
    fwd_r0.upload(nz(h.fwd_r0), st);
    bwd_s.upload(nz(h.bwd_s), st);
    bwd_q0.upload(nz(h.bwd_q0), st);
    bwd_qn.upload(nz(h.bwd_qn), st);
    roff.upload(nz64(h.roff), st);
    poff.upload(nz64(h.poff), st);
    uoff.upload(nz64(h.uoff), st);
    fpos.upload(nz64(h.fpos), st);
    cptr.upload(nz64(h.cptr), st);
    cidx.upload(nz64(h.cidx), st);
    panel.upload(h.panel.empty() ? std::vector<double>{0.0} : h.panel, st);
    ubuf.alloc(h.upd_total > 0 ? h.upd_total : 1);
    zp.alloc(h.n > 0 ? h.n : 1);
    xp.alloc(h.n > 0 ? h.n : 1);
    GT_HIP(hipStreamSynchronize(st));
}

int64_t Schur::device_bytes() const {
    return (int64_t)(perm.bytes() + c0.bytes() + nc.bytes() + nr.bytes() + rows.bytes() + fwd_s.bytes() +
                     fwd_r0.bytes() + bwd_s.bytes() + bwd_q0.bytes() + bwd_qn.bytes() + roff.bytes() + poff.bytes() + uoff.bytes() +
                     fpos.bytes() + cptr.bytes() + cidx.bytes() + panel.bytes() + ubuf.bytes() + zp.bytes() +
                     xp.bytes());
}

void schur_solve_device(Schur& S, const zd* b, zd* x, hipStream_t st) {
    SolveArgs a{S.perm.p, S.c0.p, S.nc.p, S.nr.p, S.rows.p, nullptr, nullptr, nullptr, S.roff.p, S.poff.p, S.uoff.p,
                S.fpos.p, S.cptr.p, S.cidx.p, S.panel.p, S.ubuf.p, S.zp.p, S.xp.p, b, x};
    const std::vector<int>& fp = S.h.fwd_ptr;
    a.item_s = S.fwd_s.p;
    a.item_0 = S.fwd_r0.p;
    for (size_t l = 0; l + 1 < fp.size(); ++l) {
        const int cnt = fp[l + 1] - fp[l];
        if (cnt == 0) continue;
        hipLaunchKernelGGL(schur_fwd, dim3(cnt), dim3(NTH), 0, st, a, fp[l]);
        GT_LAUNCH("schur_fwd");
    }
    const std::vector<int>& bp = S.h.bwd_ptr;
    a.item_s = S.bwd_s.p;
    a.item_0 = S.bwd_q0.p;
    a.item_n = S.bwd_qn.p;
    for (size_t l = 0; l + 1 < bp.size(); ++l) {
        const int cnt = bp[l + 1] - bp[l];
        if (cnt == 0) continue;
        hipLaunchKernelGGL(schur_bwd, dim3(cnt), dim3(NTH), 0, st, a, bp[l]);
        GT_LAUNCH("schur_bwd");
    }
}

}  // namespace gt
