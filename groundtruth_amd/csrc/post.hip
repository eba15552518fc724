// Post-processing on the device (the reference's Solution.node_field, peec.py:1151-1199): the complex sheet-current
// density (Kx, Ky) at every grid node of every layer, from the strip currents of a solution vector still on the
// device. At a node, the x strips (i - 1, j) and (i, j) and the y strips (i, j - 1) and (i, j) that exist and carry
// current are averaged and divided by the pitch; the sums run in the order the reference's numpy adds them (the
// strip starting at the node first), so the values equal the host's to rounding (numpy divides a complex by a real
// through the reciprocal). A strip is found in its orientation's row of the stack's strip layout (sorted by column)
// by binary search.
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

}  // namespace

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
