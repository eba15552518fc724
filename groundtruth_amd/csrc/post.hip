// Post-processing on the device (the reference's Solution.node_field, peec.py:1151-1199): the complex sheet-current
// density (Kx, Ky) at every grid node of every layer, from the strip currents of a solution vector still on the
// device. At a node, the x strips (i - 1, j) and (i, j) and the y strips (i, j - 1) and (i, j) that exist and carry
// current are averaged and divided by the pitch; the sums run in the order the reference's numpy adds them (the
// strip starting at the node first), so the values equal the host's to rounding (numpy divides a complex by a real
// through the reciprocal). A strip is found in its orientation's row of the stack's strip layout (sorted by column)
// by binary search.
#include <algorithm>

#include "stack.hpp"

namespace gt {

namespace {

struct FieldArgs {
    const zd* x;
    zd* kx;
    zd* ky;
    double d;
    int S, nx, ny;
    const int* rowptr[2];
    const int* rowcol[2];
    int n0[2], lo[2];
    bool present[2];
};

// the current of strip (s, i, j) of orientation oi, 0 when there is none
__device__ __forceinline__ zd strip_at(const FieldArgs& a, int oi, int s, int i, int j) {
    if (!a.present[oi] || i < 0 || i >= a.n0[oi] || j < 0) return zd();
    const int* rc = a.rowcol[oi];
    int lo = a.rowptr[oi][s * a.n0[oi] + i], hi = a.rowptr[oi][s * a.n0[oi] + i + 1];
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        const int c = rc[mid] & 0xffff;
        if (c == j) return a.x[a.lo[oi] + mid];
        if (c < j) lo = mid + 1;
        else hi = mid;
    }
    return zd();
}

__global__ __launch_bounds__(256) void node_field(FieldArgs a) {
    const int idx = blockIdx.x * 256 + threadIdx.x;
    const int plane = a.nx * a.ny;
    if (idx >= a.S * plane) return;
    const int s = idx / plane, r = idx - s * plane, i = r / a.ny, j = r - i * a.ny;
    const zd x0 = strip_at(a, 0, s, i, j), x1 = strip_at(a, 0, s, i - 1, j);
    const zd y0 = strip_at(a, 1, s, i, j), y1 = strip_at(a, 1, s, i, j - 1);
    zd kx = zd() + x0;
    kx = kx + x1;
    zd ky = zd() + y0;
    ky = ky + y1;
    const double cx = (double)((x0.re != 0 || x0.im != 0) + (x1.re != 0 || x1.im != 0));
    const double cy = (double)((y0.re != 0 || y0.im != 0) + (y1.re != 0 || y1.im != 0));
    const double ax = cx > 1 ? cx : 1.0, ay = cy > 1 ? cy : 1.0;
    a.kx[idx] = zd(kx.re / ax / a.d, kx.im / ax / a.d);
    a.ky[idx] = zd(ky.re / ay / a.d, ky.im / ay / a.d);
}

// The loop's sums over a solution vector (peec.py:2787-2811 and :2668-2700): the coupling psi . Ib, the strips' and
// branches' share of the dipole moment m = 1/2 sum r x I dl, and the injection nodes' potentials sum_n a_n V_n. Strip
// (s, i, j) of the x orientation runs from (i d, j d) to ((i + 1) d, j d) at height z_s, so r x dl = d (0, z_s, -j d);
// of the y orientation, d (-z_s, 0, i d). A row (s, i) of the strip layout therefore needs three sums: psi . I,
// sum I and sum j I. One workgroup per row writes them; one workgroup then weighs and adds the rows, the branches
// and the injections, always in the same order (the result does not depend on the launch).
struct LoopArgs {
    const zd* x;
    const double* psi;
    const double* bm;          // per branch: 1/2 r_a x r_b (3 doubles), or null
    const int* inj_idx;        // injection terms: index into x and amps
    const double* inj_amps;
    int n_inj;
    zd* part;                  // 3 sums per row, the x orientation's rows first
    zd* out;                   // psi . Ib, m (3), sum a V
    double d;
    double z[LP_MAX_SHEETS];
    int S, ns, nb;
    const int* rowptr[2];
    const int* rowcol[2];
    int n0[2], lo[2], rows[2];
};

template <int NT>
__device__ __forceinline__ zd block_sum(zd v, double* sh) {
    const int t = threadIdx.x;
    for (int off = 32; off > 0; off >>= 1) {
        v.re += __shfl_down(v.re, off, 64);
        v.im += __shfl_down(v.im, off, 64);
    }
    __syncthreads();
    if ((t & 63) == 0) {
        sh[2 * (t >> 6)] = v.re;
        sh[2 * (t >> 6) + 1] = v.im;
    }
    __syncthreads();
    zd r;
    if (t == 0)
        for (int w = 0; w < NT / 64; ++w) r = r + zd(sh[2 * w], sh[2 * w + 1]);
    return r;                  // thread 0 holds the sum
}

__global__ __launch_bounds__(256) void loop_rows(LoopArgs a) {
    __shared__ double sh[8];
    const int oi = blockIdx.x < (unsigned)a.rows[0] ? 0 : 1;
    const int row = blockIdx.x - (oi ? a.rows[0] : 0);
    const int kb = a.rowptr[oi][row], ke = a.rowptr[oi][row + 1];
    zd A, B, Cj;
    for (int k = kb + (int)threadIdx.x; k < ke; k += 256) {
        const zd v = a.x[a.lo[oi] + k];
        const double p = a.psi[a.lo[oi] + k];
        const double j = (double)(a.rowcol[oi][k] & 0xffff);
        A = A + p * v;
        B = B + v;
        Cj = Cj + j * v;
    }
    const zd sa = block_sum<256>(A, sh), sb = block_sum<256>(B, sh), sc = block_sum<256>(Cj, sh);
    if (threadIdx.x == 0) {
        zd* o = a.part + 3 * (size_t)blockIdx.x;
        o[0] = sa;
        o[1] = sb;
        o[2] = sc;
    }
}

__global__ __launch_bounds__(256) void loop_final(LoopArgs a) {
    __shared__ double sh[8];
    zd P, m0, m1, m2, V;
    const double d = a.d, h = 0.5;
    for (int r = threadIdx.x; r < a.rows[0]; r += 256) {        // x strips: r x dl = d (0, z_s, -j d)
        const zd* o = a.part + 3 * (size_t)r;
        const int s = r / a.n0[0];
        P = P + o[0];
        m1 = m1 + (h * d * a.z[s]) * o[1];
        m2 = m2 + (-h * d * d) * o[2];
    }
    for (int r = threadIdx.x; r < a.rows[1]; r += 256) {        // y strips: r x dl = d (-z_s, 0, i d)
        const zd* o = a.part + 3 * (size_t)(a.rows[0] + r);
        const int s = r / a.n0[1], i = r - s * a.n0[1];
        P = P + o[0];
        m0 = m0 + (-h * d * a.z[s]) * o[1];
        m2 = m2 + (h * d * d * i) * o[1];
    }
    for (int k = a.ns + threadIdx.x; k < a.nb; k += 256) {      // branches
        const zd v = a.x[k];
        P = P + a.psi[k] * v;
        if (a.bm) {
            const double* b = a.bm + 3 * (size_t)(k - a.ns);
            m0 = m0 + b[0] * v;
            m1 = m1 + b[1] * v;
            m2 = m2 + b[2] * v;
        }
    }
    for (int q = threadIdx.x; q < a.n_inj; q += 256) V = V + a.inj_amps[q] * a.x[a.inj_idx[q]];
    const zd sP = block_sum<256>(P, sh), s0 = block_sum<256>(m0, sh), s1 = block_sum<256>(m1, sh),
             s2 = block_sum<256>(m2, sh), sV = block_sum<256>(V, sh);
    if (threadIdx.x == 0) {
        a.out[0] = sP;
        a.out[1] = s0;
        a.out[2] = s1;
        a.out[3] = s2;
        a.out[4] = sV;
    }
}

}  // namespace

void loop_terms_device(gt_stack* st, const zd* x, const double* psi, const double* bm, int n_inj, const int* inj_idx,
                       const double* inj_amps, double d, const double* sheet_z, double* out) {
    if (!(d > 0)) fail(GT_ERR_ARG, "the pitch must be positive");
    const int n = st->nb + st->nk;
    for (int q = 0; q < n_inj; ++q)
        if (inj_idx[q] < st->nb || inj_idx[q] >= n)
            fail(GT_ERR_ARG, "an injection index is not a kept node's potential in the solution vector");
    hipStream_t s = st->ctx->stream;
    LoopArgs a{};
    a.x = x;
    a.psi = psi;
    a.bm = bm;
    a.d = d;
    a.S = st->S;
    a.ns = st->ns;
    a.nb = st->nb;
    for (int k = 0; k < st->S; ++k) a.z[k] = sheet_z[k];
    for (int oi = 0; oi < 2; ++oi) {
        const OrientHost& o = st->o[oi];
        const bool on = o.present && o.cnt > 0;
        a.rowptr[oi] = o.rowptr.p;
        a.rowcol[oi] = o.rowcol.p;
        a.n0[oi] = o.n0;
        a.lo[oi] = o.lo;
        a.rows[oi] = on ? st->S * o.n0 : 0;
    }
    const int rows = a.rows[0] + a.rows[1];
    // one buffer: the row sums, the 5 results, the injection terms
    const size_t part_b = (size_t)std::max(rows, 1) * 3 * sizeof(zd), out_b = 5 * sizeof(zd);
    const size_t inj_b = (size_t)n_inj * (sizeof(int) + sizeof(double));
    DevBuf<unsigned char> buf(part_b + out_b + inj_b + 16);
    a.part = (zd*)buf.p;
    a.out = (zd*)(buf.p + part_b);
    a.inj_amps = (const double*)(buf.p + part_b + out_b);
    a.inj_idx = (const int*)(buf.p + part_b + out_b + (size_t)n_inj * sizeof(double));
    a.n_inj = n_inj;
    if (n_inj) {
        GT_HIP(hipMemcpyAsync((void*)a.inj_amps, inj_amps, n_inj * sizeof(double), hipMemcpyHostToDevice, s));
        GT_HIP(hipMemcpyAsync((void*)a.inj_idx, inj_idx, n_inj * sizeof(int), hipMemcpyHostToDevice, s));
    }
    if (rows) {
        hipLaunchKernelGGL(loop_rows, dim3(rows), dim3(256), 0, s, a);
        GT_LAUNCH("loop_rows");
    }
    hipLaunchKernelGGL(loop_final, dim3(1), dim3(256), 0, s, a);
    GT_LAUNCH("loop_final");
    GT_HIP(hipMemcpyAsync(out, a.out, out_b, hipMemcpyDeviceToHost, s));
    GT_HIP(hipStreamSynchronize(s));
}

void node_field_device(gt_stack* st, const zd* x, double d, zd* kx, zd* ky) {
    if (!(d > 0)) fail(GT_ERR_ARG, "the pitch must be positive");
    FieldArgs a{};
    a.x = x;
    a.kx = kx;
    a.ky = ky;
    a.d = d;
    a.S = st->S;
    a.nx = st->nx;
    a.ny = st->ny;
    for (int oi = 0; oi < 2; ++oi) {
        const OrientHost& o = st->o[oi];
        a.present[oi] = o.present && o.cnt > 0;
        a.rowptr[oi] = o.rowptr.p;
        a.rowcol[oi] = o.rowcol.p;
        a.n0[oi] = o.n0;
        a.lo[oi] = o.lo;
    }
    const long total = (long)a.S * a.nx * a.ny;
    if (total == 0) return;
    hipLaunchKernelGGL(node_field, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st->ctx->stream, a);
    GT_LAUNCH("node_field");
}

}  // namespace gt
