// The L_p kernel spectra made on the device (peec.py:3276-3371 `_kernel_pair` / `_kernel_spectrum`,
// kernel.py:268-294 `strip_kernel`): the mutual partial inductance of two square strips at every grid offset, and
// the real, even spectrum of its circular embedding in the (P0, P1) FFT grid, as the quadrant the L_p product reads.
//
// The kernel is even in both offsets, so its DFT is real and separable into two cosine sums over the offsets' first
// quadrant (weights 1 at offset 0 and 2 elsewhere, for the +/- pair):
//   G(a, qj)  = sum_b w_b K(a, b) cos(2 pi qj b / P1)
//   F(qi, qj) = sum_a w_a G(a, qj) cos(2 pi qi a / P0) / (P0 P1)
// Each is a dense product with a cosine matrix on LDS tiles (a 64 x 64 tile per 256-thread workgroup, 4 x 4 outputs
// per thread), in FP64 with compensated (Neumaier) sums, so a spectrum entry carries the rounding of its terms and
// not that of a 1000-term running sum; the cosines come from the stack's twiddle tables (exact index q k mod P).
// Far from the diagonal (5 cells or more along or across) the kernel is the mutual of two filaments on the strips'
// centre lines (kernel.py:165-176): mu0/4pi times the second difference g(x + L) - 2 g(x) + g(x - L) of
// g(u) = u asinh(u / rho) - sqrt(u^2 + rho^2). In that closed form the three terms cancel to (L / x)^2 of their size,
// which leaves the reference's double-precision values up to 1.6e-10 off at the far end of a 256-cell board. Here
// the same quantity is taken as the integral it equals, int_{-L}^{L} (L - |t|) g''(x + t) dt with
// g''(u) = 1 / sqrt(u^2 + rho^2): positive terms, no cancellation, and analytic on each half interval with its
// nearest singularity at least four interval lengths away, so 8 Gauss-Legendre points per half give it to
// rounding. Within 4 cells along and across it is the near-field tape mutual, whose graded quadrature stays on the
// host (a 5 x 5 table per sheet pair, the reference's rule).
#include <cmath>

#include "stack.hpp"

namespace gt {

namespace {

constexpr int KG_PAIRS = LP_MAX_SHEETS * (LP_MAX_SHEETS + 1) / 2;
constexpr int KG_NEAR = 5;
constexpr int KG_GAUSS = 8;

struct FillArgs {
    double d, mu;               // pitch (m); mu0 / (4 pi)
    int n0, n1, S, npair;
    int axial_i;                // 1: the strips run along i (x strips), 0: along j (y strips)
    int ps[KG_PAIRS], pt[KG_PAIRS];
    double gx[KG_GAUSS], gw[KG_GAUSS];   // Gauss-Legendre nodes and weights on [0, 1]
    const double* dz;           // S x S face gaps
    const double* nearv;        // S x S x 5 x 5 near-field mutuals, [s][t][axial][lateral]
    double* K;                  // npair x n0 x n1
};


// K[pair][a][b]: the strips a cells apart along i and b along j
__global__ __launch_bounds__(256) void kg_fill(FillArgs f) {
    const int pair = blockIdx.y;
    const size_t e = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= (size_t)f.n0 * f.n1) return;
    const int a = (int)(e / f.n1), b = (int)(e % f.n1);
    const int s = f.ps[pair], t = f.pt[pair];
    const int ax = f.axial_i ? a : b, lat = f.axial_i ? b : a;
    double v;
    if (ax < KG_NEAR && lat < KG_NEAR) {
        v = f.nearv[((size_t)(s * f.S + t) * KG_NEAR + ax) * KG_NEAR + lat];
    } else {
        const double dz = f.dz[s * f.S + t];
        const double l = lat * f.d, x = ax * f.d, L = f.d;
        const double r2 = l * l + dz * dz;
        double acc = 0.0;
#pragma unroll
        for (int q = 0; q < KG_GAUSS; ++q) {
            const double tq = L * f.gx[q];
            const double up = x + tq, um = x - tq;
            acc += f.gw[q] * (L - tq) * (1.0 / sqrt(up * up + r2) + 1.0 / sqrt(um * um + r2));
        }
        v = f.mu * L * acc;
    }
    f.K[(size_t)pair * f.n0 * f.n1 + e] = v;
}

// C(m, n) = scale sum_k A(m, k) w_k cos(2 pi k n / P) for every pair; A(m, k) = A[pair * a_pair + m sa_m + k sa_k],
// C(m, n) at c_off[pair] + m sc_m + n sc_n (and at c_off2[pair] when it is >= 0: the transposed sheet pair)
struct CosArgs {
    const double* A;
    long a_pair, sa_m, sa_k;
    const zd* tw;
    int P;
    double* C;
    long sc_m, sc_n;
    long c_off[KG_PAIRS], c_off2[KG_PAIRS];
    double scale;
    int M, N, K;
};

constexpr int CT = 64, CK = 16;

__global__ __launch_bounds__(256) void kg_cos(CosArgs g) {
    __shared__ double As[CT][CK + 1];
    __shared__ double Bs[CK][CT + 1];
    const int pair = blockIdx.z;
    const int m0 = blockIdx.y * CT, n0 = blockIdx.x * CT;
    const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
    const double* A = g.A + (size_t)pair * g.a_pair;
    double acc[4][4], cmp[4][4];              // Neumaier-compensated sums
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[r][c] = cmp[r][c] = 0.0;
    for (int k0 = 0; k0 < g.K; k0 += CK) {
        for (int e = tid; e < CT * CK; e += 256) {
            const int mm = e / CK, kk = e % CK;
            const int m = m0 + mm, k = k0 + kk;
            As[mm][kk] = (m < g.M && k < g.K) ? A[(size_t)m * g.sa_m + (size_t)k * g.sa_k] : 0.0;
        }
        for (int e = tid; e < CK * CT; e += 256) {
            const int kk = e / CT, nn = e % CT;
            const int k = k0 + kk, n = n0 + nn;
            double v = 0.0;
            if (k < g.K && n < g.N) v = (k == 0 ? 1.0 : 2.0) * g.tw[(int)(((long)k * n) & (g.P - 1))].re;
            Bs[kk][nn] = v;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < CK; ++kk) {
            double av[4], bv[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) av[r] = As[ty * 4 + r][kk];
#pragma unroll
            for (int c = 0; c < 4; ++c) bv[c] = Bs[kk][tx * 4 + c];
#pragma unroll
            for (int r = 0; r < 4; ++r)
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    const double x = av[r] * bv[c], s0 = acc[r][c], t = s0 + x;
                    cmp[r][c] += fabs(s0) >= fabs(x) ? (s0 - t) + x : (x - t) + s0;
                    acc[r][c] = t;
                }
        }
        __syncthreads();
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int m = m0 + ty * 4 + r;
        if (m >= g.M) continue;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const int n = n0 + tx * 4 + c;
            if (n >= g.N) continue;
            const double v = (acc[r][c] + cmp[r][c]) * g.scale;
            const size_t o = (size_t)m * g.sc_m + (size_t)n * g.sc_n;
            g.C[g.c_off[pair] + o] = v;
            if (g.c_off2[pair] >= 0) g.C[g.c_off2[pair] + o] = v;
        }
    }
}

__global__ __launch_bounds__(256) void kg_to_float(const double* __restrict__ x, float* __restrict__ y, size_t n) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) y[i] = (float)x[i];
}

}  // namespace

void kernel_quadrant_device(gt_stack* st, int oi, const KernelGen& kg) {
    OrientHost& o = st->o[oi];
    hipStream_t s = st->ctx->stream;
    const int S = st->S, n0 = o.n0, n1 = o.n1, P0 = o.P0, P1 = o.P1;
    const int h0 = P0 / 2 + 1, h1 = P1 / 2 + 1;
    int npair = 0;
    FillArgs f{};
    f.d = kg.d;
    f.mu = kg.mu;
    f.n0 = n0;
    f.n1 = n1;
    f.S = S;
    f.axial_i = oi == 0 ? 1 : 0;
    // Gauss-Legendre rule of KG_GAUSS points: Newton's iteration on P_n from Chebyshev-like guesses, mapped to [0, 1]
    for (int q = 0; q < KG_GAUSS; ++q) {
        const int n = KG_GAUSS;
        double z = std::cos(3.14159265358979323846 * (q + 0.75) / (n + 0.5)), dp = 0;
        for (int it = 0; it < 100; ++it) {
            double p0 = 1.0, p1 = z;
            for (int k = 2; k <= n; ++k) {
                const double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
                p0 = p1;
                p1 = p2;
            }
            dp = n * (z * p1 - p0) / (z * z - 1.0);
            const double dz_ = p1 / dp;
            z -= dz_;
            if (std::fabs(dz_) < 1e-17) break;
        }
        f.gx[q] = 0.5 * (1.0 + z);
        f.gw[q] = 1.0 / ((1.0 - z * z) * dp * dp);   // 2 / ((1 - z^2) P_n'(z)^2), halved for [0, 1]
    }
    // one kernel per distinct face gap (sheet pairs at equal gaps have equal kernels), from a representative pair
    for (int g = 0; g < kg.G; ++g) {
        f.ps[npair] = kg.rep_s[g];
        f.pt[npair] = kg.rep_t[g];
        ++npair;
    }
    f.npair = npair;
    DevBuf<double> dz((size_t)S * S), nearv((size_t)S * S * KG_NEAR * KG_NEAR);
    dz.upload(kg.dz, (size_t)S * S, s);
    nearv.upload(kg.nearv, (size_t)S * S * KG_NEAR * KG_NEAR, s);
    f.dz = dz.p;
    f.nearv = nearv.p;
    DevBuf<double> K((size_t)npair * n0 * n1), G((size_t)npair * n0 * h1);
    f.K = K.p;
    hipLaunchKernelGGL(kg_fill, dim3(blocks_for((size_t)n0 * n1, 256), npair), dim3(256), 0, s, f);
    GT_LAUNCH("kg_fill");
    // G[pair][a][qj] = sum_b w_b K[pair][a][b] cos(2 pi qj b / P1)
    CosArgs c1{};
    c1.A = K.p;
    c1.a_pair = (long)n0 * n1;
    c1.sa_m = n1;
    c1.sa_k = 1;
    c1.tw = st->tw.at(P1).p;
    c1.P = P1;
    c1.C = G.p;
    c1.sc_m = h1;
    c1.sc_n = 1;
    for (int p = 0; p < npair; ++p) {
        c1.c_off[p] = (long)p * n0 * h1;
        c1.c_off2[p] = -1;
    }
    c1.scale = 1.0;
    c1.M = n0;
    c1.N = h1;
    c1.K = n1;
    hipLaunchKernelGGL(kg_cos, dim3((h1 + CT - 1) / CT, (n0 + CT - 1) / CT, npair), dim3(256), 0, s, c1);
    GT_LAUNCH("kg_cos");
    // F[qj][qi][g] = sum_a w_a G[g][a][qj] cos(2 pi qi a / P0) / (P0 P1)
    const size_t kq = (size_t)h1 * h0 * npair;
    o.K->alloc(kq);
    CosArgs c2{};
    c2.A = G.p;
    c2.a_pair = (long)n0 * h1;
    c2.sa_m = 1;
    c2.sa_k = h1;
    c2.tw = st->tw.at(P0).p;
    c2.P = P0;
    c2.C = o.K->p;
    c2.sc_m = (long)h0 * npair;
    c2.sc_n = npair;
    for (int p = 0; p < npair; ++p) {
        c2.c_off[p] = p;
        c2.c_off2[p] = -1;
    }
    c2.scale = 1.0 / ((double)P0 * (double)P1);
    c2.M = h1;
    c2.N = h0;
    c2.K = n0;
    hipLaunchKernelGGL(kg_cos, dim3((h0 + CT - 1) / CT, (h1 + CT - 1) / CT, npair), dim3(256), 0, s, c2);
    GT_LAUNCH("kg_cos");
    o.Kf->alloc(kq);
    hipLaunchKernelGGL(kg_to_float, dim3(blocks_for(kq, 256)), dim3(256), 0, s, o.K->p, o.Kf->p, kq);
    GT_LAUNCH("kg_to_float");
    GT_HIP(hipStreamSynchronize(s));
}

}  // namespace gt
