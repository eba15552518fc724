// Level-scheduled supernodal solves of B x = b on the device (the Schur block of the preconditioner,
// peec.py:3064-3083 `op.solve_S`), with each supernode's panel stored as Q = [L_CC^-1 ; -L_RC L_CC^-1]
// (schur_factor.cpp), so that neither direction has a serial chain:
//   forward (leaves first):  f = [b_C ; 0] + the children's update entries (pulled in a fixed order),
//                            [z_C ; u_R] = Q f_C + [0 ; f_R]          -- a dense product over the panel's rows
//   backward (roots first):  x_C = Q^T [z_C ; x_R]                     -- a dense product over its columns
// One launch per level; a large supernode is split over many workgroups (blocks of 64 rows forward, 16 columns
// backward), so the wide top separators run on many CUs. The real panel multiplies complex vectors.
#include "lp.hpp"
#include "schur.hpp"

namespace gt {

namespace {

constexpr int NTH = 256;

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

// Entries staged in LDS per pass for a batch of up to KB right-hand sides (16 KiB).
template <int KB>
constexpr int chunk_rows() { return KB == 1 ? 1024 : (KB <= 4 ? 256 : 64); }

// f at front position q of supernode s for right-hand side p of nr (interleaved vectors): its right-hand side
// (a column) plus the children's update entries, in the contributors' fixed order
__device__ __forceinline__ zd front(const SolveArgs& a, int s, int q, int C, int c0, int p, int nr) {
    zd acc = q < C ? a.b[(size_t)a.perm[c0 + q] * nr + p] : zd();
    const int64_t pos = a.fpos[s] + q;
    for (int64_t e = a.cptr[pos]; e < a.cptr[pos + 1]; ++e) acc = acc + a.ubuf[(size_t)a.cidx[e] * nr + p];
    return acc;
}

// One work item: rows [r0, r0 + 64) of supernode s, for nr <= KB right-hand sides. Lane = row; the four waves
// split the columns q (stride 4); their partial sums meet in LDS.
template <int KB>
__global__ __launch_bounds__(NTH) void schur_fwd(SolveArgs a, int item0, int nr) {
    constexpr int CB = chunk_rows<KB>();
    __shared__ double f_raw[2 * CB * KB];
    __shared__ double part_raw[2 * NTH];
    zd* f = reinterpret_cast<zd*>(f_raw);
    zd* part = reinterpret_cast<zd*>(part_raw);
    const int s = a.item_s[item0 + blockIdx.x], r0 = a.item_0[item0 + blockIdx.x];
    const int C = a.nc[s], m = C + a.nr[s], c0 = a.c0[s];
    const double* Q = a.panel + a.poff[s];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r = r0 + lane;
    const int qlim = r < C ? r + 1 : C;
    zd acc[KB];
    for (int cs = 0; cs < C; cs += CB) {
        const int ce = min(C, cs + CB);
        for (int e = threadIdx.x; e < (ce - cs) * nr; e += NTH) {
            const int q = cs + e / nr, p = e - (q - cs) * nr;
            f[(q - cs) * KB + p] = front(a, s, q, C, c0, p, nr);
        }
        __syncthreads();
        if (r < m) {
            const int qe = min(ce, qlim);
            int q = cs + wave;
            for (; q + 4 < qe; q += 8) {            // two independent panel loads in flight per lane
                const double w0 = Q[(size_t)q * m + r], w1 = Q[(size_t)(q + 4) * m + r];
                const zd* f0 = f + (q - cs) * KB;
                const zd* f1 = f + (q + 4 - cs) * KB;
#pragma unroll
                for (int p = 0; p < KB; ++p) {
                    if (p < nr) {
                        acc[p].re += w0 * f0[p].re + w1 * f1[p].re;
                        acc[p].im += w0 * f0[p].im + w1 * f1[p].im;
                    }
                }
            }
            for (; q < qe; q += 4) {
                const double w0 = Q[(size_t)q * m + r];
                const zd* f0 = f + (q - cs) * KB;
#pragma unroll
                for (int p = 0; p < KB; ++p) {
                    if (p < nr) {
                        acc[p].re += w0 * f0[p].re;
                        acc[p].im += w0 * f0[p].im;
                    }
                }
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int p = 0; p < KB; ++p) {
        if (p < nr) {
            part[threadIdx.x] = acc[p];
            __syncthreads();
            if (wave == 0 && r < m) {
                zd t = part[lane] + part[64 + lane] + part[128 + lane] + part[192 + lane];
                if (r < C) {
                    a.zp[(size_t)(c0 + r) * nr + p] = t;
                } else {
                    t = t + front(a, s, r, C, c0, p, nr);
                    a.ubuf[(size_t)(a.uoff[s] + r - C) * nr + p] = t;
                }
            }
            __syncthreads();
        }
    }
}

// One work item: columns [q0, q0 + qn) of supernode s, taken 4 K at a time: wave w holds columns qc + w + 4k,
// its lanes stride the rows; every panel load feeds all nr <= KB right-hand sides.
template <int KB>
__global__ __launch_bounds__(NTH) void schur_bwd(SolveArgs a, int item0, int nr) {
    constexpr int CB = chunk_rows<KB>();
    constexpr int K = KB == 1 ? 16 : (KB <= 4 ? 4 : 1);
    __shared__ double v_raw[2 * CB * KB];
    zd* v = reinterpret_cast<zd*>(v_raw);
    const int s = a.item_s[item0 + blockIdx.x], q0 = a.item_0[item0 + blockIdx.x];
    const int C = a.nc[s], m = C + a.nr[s], c0 = a.c0[s];
    const int qend = min(C, q0 + a.item_n[item0 + blockIdx.x]);
    const double* Q = a.panel + a.poff[s];
    const int* rows = a.rows + a.roff[s];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int qc = q0; qc < qend; qc += 4 * K) {
        zd acc[K][KB];
        for (int rs = qc; rs < m; rs += CB) {       // rows from the chunk's first column down
            const int re = min(m, rs + CB);
            for (int e = threadIdx.x; e < (re - rs) * nr; e += NTH) {
                const int r = rs + e / nr, p = e - (r - rs) * nr;
                v[(r - rs) * KB + p] = r < C ? a.zp[(size_t)(c0 + r) * nr + p] : a.xp[(size_t)rows[r - C] * nr + p];
            }
            __syncthreads();
#pragma unroll
            for (int k = 0; k < K; ++k) {
                const int q = qc + wave + 4 * k;
                if (q < qend) {
                    const double* Qq = Q + (size_t)q * m;
                    for (int r = max(q, rs) + lane; r < re; r += 64) {
                        const double w0 = Qq[r];
                        const zd* u = v + (r - rs) * KB;
#pragma unroll
                        for (int p = 0; p < KB; ++p) {
                            if (p < nr) {
                                acc[k][p].re += w0 * u[p].re;
                                acc[k][p].im += w0 * u[p].im;
                            }
                        }
                    }
                }
            }
            __syncthreads();
        }
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const int q = qc + wave + 4 * k;
#pragma unroll
            for (int p = 0; p < KB; ++p) {
                if (p < nr) {
                    double re = acc[k][p].re, im = acc[k][p].im;
                    for (int off = 32; off > 0; off >>= 1) {
                        re += __shfl_down(re, off, 64);
                        im += __shfl_down(im, off, 64);
                    }
                    if (lane == 0 && q < qend) {
                        a.xp[(size_t)(c0 + q) * nr + p] = zd(re, im);
                        a.x[(size_t)a.perm[c0 + q] * nr + p] = zd(re, im);
                    }
                }
            }
        }
    }
}

template <int KB>
void run_levels(const Schur& S, SolveArgs a, int nr, hipStream_t st) {
    const std::vector<int>& fp = S.h.fwd_ptr;
    a.item_s = S.fwd_s.p;
    a.item_0 = S.fwd_r0.p;
    for (size_t l = 0; l + 1 < fp.size(); ++l) {
        const int cnt = fp[l + 1] - fp[l];
        if (cnt == 0) continue;
        hipLaunchKernelGGL(schur_fwd<KB>, dim3(cnt), dim3(NTH), 0, st, a, fp[l], nr);
        GT_LAUNCH("schur_fwd");
    }
    const std::vector<int>& bp = S.h.bwd_ptr;
    a.item_s = S.bwd_s.p;
    a.item_0 = S.bwd_q0.p;
    a.item_n = S.bwd_qn.p;
    for (size_t l = 0; l + 1 < bp.size(); ++l) {
        const int cnt = bp[l + 1] - bp[l];
        if (cnt == 0) continue;
        hipLaunchKernelGGL(schur_bwd<KB>, dim3(cnt), dim3(NTH), 0, st, a, bp[l], nr);
        GT_LAUNCH("schur_bwd");
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
    ubuf.alloc((size_t)(h.upd_total > 0 ? h.upd_total : 1) * KMAX);
    zp.alloc((size_t)(h.n > 0 ? h.n : 1) * KMAX);
    xp.alloc((size_t)(h.n > 0 ? h.n : 1) * KMAX);
    GT_HIP(hipStreamSynchronize(st));
}

int64_t Schur::device_bytes() const {
    return (int64_t)(perm.bytes() + c0.bytes() + nc.bytes() + nr.bytes() + rows.bytes() + fwd_s.bytes() +
                     fwd_r0.bytes() + bwd_s.bytes() + bwd_q0.bytes() + bwd_qn.bytes() + roff.bytes() + poff.bytes() + uoff.bytes() +
                     fpos.bytes() + cptr.bytes() + cidx.bytes() + panel.bytes() + ubuf.bytes() + zp.bytes() +
                     xp.bytes());
}

void schur_solve_device(Schur& S, const zd* b, zd* x, int nr, hipStream_t st) {
    if (nr < 1 || nr > KMAX) fail(GT_ERR_ARG, "Schur solve: 1 to 16 right-hand sides");
    SolveArgs a{S.perm.p, S.c0.p, S.nc.p, S.nr.p, S.rows.p, nullptr, nullptr, nullptr, S.roff.p, S.poff.p, S.uoff.p,
                S.fpos.p, S.cptr.p, S.cidx.p, S.panel.p, S.ubuf.p, S.zp.p, S.xp.p, b, x};
    if (nr == 1) run_levels<1>(S, a, nr, st);
    else if (nr <= 4) run_levels<4>(S, a, nr, st);
    else run_levels<16>(S, a, nr, st);
}

}  // namespace gt
