// The strips' coupling to the signal trace on the device (peec.py:2609-2656 `strip_coupling`): psi_k = d times
// the sum over the trace's segments of (current share) (segment direction . strip axis) times the segment's vector
// potential at the strip, per amp. The potential of a straight filament is the closed form
// mu0/4pi (asinh((L - s)/rho) + asinh(s / rho)) (kernel.py:297-312 `segment_vector_potential`), taken at the
// strip's centre on its sheet face toward the segment, or as the mean over 3 x 3 Gauss points of the strip's
// square when the centre lies within 3 pitches plus 2 widths of the segment. Segments along z or across the strip
// axis add nothing to that orientation.
//
// One workgroup per (sheet, row) of an orientation's strip layout (the CSR the L_p product uses), a thread per strip,
// every segment in turn: the strips' right-hand sides without a host pass over the board.
#include <cmath>

#include "stack.hpp"

namespace gt {

namespace {

constexpr int PSI_MAX_SHEETS = LP_MAX_SHEETS;

struct Seg {
    double a[3], u[3];          // start, unit direction
    double L, zmid, rho_min, reach;
    double coef[2];             // current share x (direction . axis) per orientation, 0 when it adds nothing
};

struct PsiArgs {
    const Seg* seg;
    int nseg;
    double d, mu;
    double z[PSI_MAX_SHEETS], half[PSI_MAX_SHEETS];
    double gq[3], gw[3];        // 3-point Gauss nodes on [-1/2, 1/2] (in pitches) and weights summing to 1
    const int* rowptr;
    const int* rowcol;
    int n0, lo, oi;
    double* psi;
};

__device__ __forceinline__ double seg_pot(const Seg& g, double px, double py, double pz, double mu) {
    const double qx = px - g.a[0], qy = py - g.a[1], qz = pz - g.a[2];
    const double s = qx * g.u[0] + qy * g.u[1] + qz * g.u[2];
    const double r2 = fmax(qx * qx + qy * qy + qz * qz - s * s, g.rho_min * g.rho_min);
    const double rho = sqrt(r2);
    return mu * (asinh((g.L - s) / rho) + asinh(s / rho));
}

__device__ __forceinline__ double seg_dist(const Seg& g, double px, double py, double pz) {
    const double qx = px - g.a[0], qy = py - g.a[1], qz = pz - g.a[2];
    const double t = fmin(fmax(qx * g.u[0] + qy * g.u[1] + qz * g.u[2], 0.0), g.L);
    const double ex = qx - t * g.u[0], ey = qy - t * g.u[1], ez = qz - t * g.u[2];
    return sqrt(ex * ex + ey * ey + ez * ez);
}

__global__ __launch_bounds__(256) void psi_strips(PsiArgs a) {
    const int row = blockIdx.x;                 // s * n0 + i
    const int s = row / a.n0, i = row - s * a.n0;
    const int kb = a.rowptr[row], ke = a.rowptr[row + 1];
    const double d = a.d;
    for (int k = kb + threadIdx.x; k < ke; k += 256) {
        const int j = a.rowcol[k] & 0xffff;
        const double cx = i * d + (a.oi == 0 ? 0.5 * d : 0.0);
        const double cy = j * d + (a.oi == 1 ? 0.5 * d : 0.0);
        double acc = 0.0;
        for (int q = 0; q < a.nseg; ++q) {
            const Seg g = a.seg[q];
            const double coef = g.coef[a.oi];
            if (coef == 0.0) continue;
            const double cz = a.z[s] + (g.zmid > a.z[s] ? a.half[s] : -a.half[s]);
            double val;
            if (seg_dist(g, cx, cy, cz) < g.reach) {
                val = 0.0;
                for (int ua = 0; ua < 3; ++ua)
                    for (int ul = 0; ul < 3; ++ul) {
                        // offsets along the strip axis (ua) and across it (ul)
                        const double dx = (a.oi == 0 ? a.gq[ua] : a.gq[ul]) * d;
                        const double dy = (a.oi == 0 ? a.gq[ul] : a.gq[ua]) * d;
                        val += a.gw[ua] * a.gw[ul] * seg_pot(g, cx + dx, cy + dy, cz, a.mu);
                    }
            } else {
                val = seg_pot(g, cx, cy, cz, a.mu);
            }
            acc += coef * val;
        }
        a.psi[a.lo + k] = d * acc;
    }
}

// The vias' coupling to the trace's vertical segments (peec.py:2658-2666 `branch_coupling`): per via, the sum over
// segments of (current share) times the mutual of the segment and the via's axis, the segment's potential integrated
// along the axis by a 16-point Gauss rule times the cosine between them (the via's radius the least distance).
constexpr int VIA_GAUSS = 16;

struct ViaArgs {
    const Seg* seg;
    const double* via;          // per via: a (3), b (3), radius
    int nseg, nv;
    double mu;
    double gx[VIA_GAUSS], gw[VIA_GAUSS];     // nodes on [0, 1] and weights summing to 1
    double* out;
};

__global__ __launch_bounds__(64) void psi_vias(ViaArgs a) {
    const int v = blockIdx.x * 64 + threadIdx.x;
    if (v >= a.nv) return;
    const double* p = a.via + 7 * (size_t)v;
    const double dx = p[3] - p[0], dy = p[4] - p[1], dz = p[5] - p[2];
    const double L2 = sqrt(dx * dx + dy * dy + dz * dz);
    double acc = 0.0;
    if (L2 > 0) {
        for (int q = 0; q < a.nseg; ++q) {
            Seg g = a.seg[q];
            g.rho_min = p[6];
            double sum = 0.0;
            for (int k = 0; k < VIA_GAUSS; ++k)
                sum += a.gw[k] * seg_pot(g, p[0] + a.gx[k] * dx, p[1] + a.gx[k] * dy, p[2] + a.gx[k] * dz, a.mu);
            const double cosang = (g.u[0] * dx + g.u[1] * dy + g.u[2] * dz) / L2;
            acc += g.coef[0] * cosang * L2 * sum;
        }
    }
    a.out[v] = acc;
}

// Gauss-Legendre nodes and weights of order n on [-1, 1]: Newton's iteration on P_n from the asymptotic roots.
void gauss_legendre(int n, double* x, double* w) {
    for (int i = 0; i < n; ++i) {
        double t = std::cos(M_PI * (i + 0.75) / (n + 0.5)), dp = 0;
        for (int it = 0; it < 100; ++it) {
            double p0 = 1, p1 = t;
            for (int k = 2; k <= n; ++k) {
                const double pk = ((2 * k - 1) * t * p1 - (k - 1) * p0) / k;
                p0 = p1;
                p1 = pk;
            }
            dp = n * (t * p1 - p0) / (t * t - 1);
            const double dt = p1 / dp;
            t -= dt;
            if (std::fabs(dt) < 1e-16) break;
        }
        double p0 = 1, p1 = t;
        for (int k = 2; k <= n; ++k) {
            const double pk = ((2 * k - 1) * t * p1 - (k - 1) * p0) / k;
            p0 = p1;
            p1 = pk;
        }
        dp = n * (t * p1 - p0) / (t * t - 1);
        x[i] = t;
        w[i] = 2.0 / ((1 - t * t) * dp * dp);
    }
}

}  // namespace

void via_coupling_device(gt_ctx* ctx, int nv, const double* vias, int nseg, const gt_segment* segs, double mu,
                         double* out) {
    hipStream_t s = ctx->stream;
    std::vector<Seg> hs;
    for (int q = 0; q < nseg; ++q) {
        const gt_segment& g = segs[q];
        double ab[3];
        for (int c = 0; c < 3; ++c) ab[c] = g.b[c] - g.a[c];
        // vertical segments only: along z with no run in the plane
        if (!(std::fabs(ab[2]) > 0) || !(std::hypot(ab[0], ab[1]) < 1e-12)) continue;
        const double L = std::sqrt(ab[0] * ab[0] + ab[1] * ab[1] + ab[2] * ab[2]);
        Seg v{};
        for (int c = 0; c < 3; ++c) {
            v.a[c] = g.a[c];
            v.u[c] = ab[c] / L;
        }
        v.L = L;
        v.coef[0] = g.current;
        hs.push_back(v);
    }
    if (nv == 0) return;
    if (hs.empty()) {
        std::fill(out, out + nv, 0.0);
        return;
    }
    for (int v = 0; v < nv; ++v)
        if (!(vias[7 * v + 6] > 0)) fail(GT_ERR_ARG, "a via's radius must be positive");
    DevBuf<unsigned char> buf(hs.size() * sizeof(Seg) + (size_t)nv * 8 * sizeof(double));
    ViaArgs a{};
    a.seg = (const Seg*)buf.p;
    a.via = (const double*)(buf.p + hs.size() * sizeof(Seg));
    a.out = (double*)(buf.p + hs.size() * sizeof(Seg) + (size_t)nv * 7 * sizeof(double));
    a.nseg = (int)hs.size();
    a.nv = nv;
    a.mu = mu;
    double x[VIA_GAUSS], w[VIA_GAUSS];
    gauss_legendre(VIA_GAUSS, x, w);
    for (int k = 0; k < VIA_GAUSS; ++k) {
        a.gx[k] = 0.5 * (x[k] + 1.0);
        a.gw[k] = 0.5 * w[k];
    }
    GT_HIP(hipMemcpyAsync((void*)a.seg, hs.data(), hs.size() * sizeof(Seg), hipMemcpyHostToDevice, s));
    GT_HIP(hipMemcpyAsync((void*)a.via, vias, (size_t)nv * 7 * sizeof(double), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(psi_vias, dim3((nv + 63) / 64), dim3(64), 0, s, a);
    GT_LAUNCH("psi_vias");
    GT_HIP(hipMemcpyAsync(out, a.out, (size_t)nv * sizeof(double), hipMemcpyDeviceToHost, s));
    GT_HIP(hipStreamSynchronize(s));
}

void strip_coupling_device(gt_stack* st, double d, double mu, int nseg, const gt_segment* segs, const double* sheet_z,
                           const double* sheet_t, double* psi) {
    hipStream_t s = st->ctx->stream;
    if (!(d > 0)) fail(GT_ERR_ARG, "the pitch must be positive");
    std::vector<Seg> hs;
    for (int q = 0; q < nseg; ++q) {
        const gt_segment& g = segs[q];
        Seg v{};
        double ab[3];
        for (int c = 0; c < 3; ++c) {
            v.a[c] = g.a[c];
            ab[c] = g.b[c] - g.a[c];
        }
        if (std::hypot(ab[0], ab[1]) == 0) continue;        // along z: only vias couple to it
        const double L = std::sqrt(ab[0] * ab[0] + ab[1] * ab[1] + ab[2] * ab[2]);
        for (int c = 0; c < 3; ++c) v.u[c] = ab[c] / L;
        v.L = L;
        v.zmid = 0.5 * (g.a[2] + g.b[2]);
        v.rho_min = std::max(0.5 * g.width, 1e-7);
        v.reach = 3 * d + 2 * g.width;
        for (int oi = 0; oi < 2; ++oi) v.coef[oi] = std::fabs(ab[oi]) < 1e-15 * L ? 0.0 : g.current * ab[oi] / L;
        hs.push_back(v);
    }
    DevBuf<double> segbuf(std::max<size_t>(1, hs.size() * sizeof(Seg) / sizeof(double)));
    if (!hs.empty())
        GT_HIP(hipMemcpyAsync(segbuf.p, hs.data(), hs.size() * sizeof(Seg), hipMemcpyHostToDevice, s));
    PsiArgs a{};
    a.seg = reinterpret_cast<const Seg*>(segbuf.p);
    a.nseg = (int)hs.size();
    a.d = d;
    a.mu = mu;
    for (int k = 0; k < st->S; ++k) {
        a.z[k] = sheet_z[k];
        a.half[k] = 0.5 * sheet_t[k];
    }
    const double g3 = std::sqrt(0.6);
    a.gq[0] = -0.5 * g3;
    a.gq[1] = 0.0;
    a.gq[2] = 0.5 * g3;
    a.gw[0] = 5.0 / 18.0;
    a.gw[1] = 8.0 / 18.0;
    a.gw[2] = 5.0 / 18.0;
    a.psi = psi;
    for (int oi = 0; oi < 2; ++oi) {
        const OrientHost& o = st->o[oi];
        if (!o.present) continue;
        a.rowptr = o.rowptr.p;
        a.rowcol = o.rowcol.p;
        a.n0 = o.n0;
        a.lo = o.lo;
        a.oi = oi;
        hipLaunchKernelGGL(psi_strips, dim3(st->S * o.n0), dim3(256), 0, s, a);
        GT_LAUNCH("psi_strips");
    }
    GT_HIP(hipStreamSynchronize(s));
}

}  // namespace gt
