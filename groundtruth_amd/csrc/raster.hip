// Copper rasterisation on the device (peec.py:329-983 `build_grid` and helpers): the per-sample and per-cell work
// of turning copper polygons into a grid, where the board's area is.
//
// - Lattice labels (`contains_xy` over a lattice, peec.py:379-385, :791-797): one workgroup per lattice column
//   x = X. The column's crossing edges (those with min x <= X < max x, listed on the host) are sorted by their
//   crossing height in LDS; a sample's crossings below it are counted by binary search, and the few within a rounding
//   window of it are decided by the exact orientation sign. A sample is inside when the count is odd and it lies on
//   no edge (contains_xy is the interior): on-edge tests also take the column's vertical edges and the vertices on X.
//   Pieces claim samples in order, the first keeping one (peec.py:385).
// - The fine grid's cells (`_coarse_from_fine`, peec.py:774-849): one thread per coarse cell flood-fills the cell's
//   s x s samples into pieces, keeps the largest (ties: the piece met first in row-major order) when it spans s
//   samples, and one thread per cell border counts the kept pieces' fine strips across it, adding 1/s and the
//   offsets in the order numpy's add.at does.
// - Line covers (`_line_cover`, peec.py:533-583): one workgroup per line; its crossings sorted in LDS bound the
//   even-odd spans, and each thread sums the spans' pieces over its cells.
// - Point location and nearest points (`covers`, `intersects_xy`, `shortest_line`): a thread per point over the
//   edges.
//
// The file is compiled without floating-point contraction: crossings, sample decisions and nearest points round as
// the reference's numpy expressions do.
#pragma clang fp contract(off)

#include <cmath>

#include "stack.hpp"

namespace gt {

namespace {

constexpr int RS_NTH = 256;
constexpr int RS_CAP = 4096;                 // crossings of one lattice column or cover line (LDS)
constexpr int RS_SMAX = 8;                   // largest sub-grid factor
constexpr double RS_ERRB = 3.3306690738754716e-16;   // (3 + 16 eps) eps: float orientation error bound

// ---- exact orientation: sign of (b - a) x (p - a) (Shewchuk's float filter, then an exact expansion)
__host__ __device__ inline void two_sum(double a, double b, double& x, double& y) {
    x = a + b;
    const double bv = x - a, av = x - bv;
    y = (a - av) + (b - bv);
}
__host__ __device__ inline void two_diff(double a, double b, double& x, double& y) {
    x = a - b;
    const double bv = a - x, av = x + bv;
    y = (a - av) + (bv - b);
}
__host__ __device__ inline void two_prod(double a, double b, double& x, double& y) {
    x = a * b;
    y = fma(a, b, -x);
}
// h = e + b exactly (e nonoverlapping, increasing magnitude, zeros dropped); returns the new length
__host__ __device__ inline int grow(const double* e, int n, double b, double* h) {
    double q = b;
    int m = 0;
    for (int i = 0; i < n; ++i) {
        double s, t;
        two_sum(q, e[i], s, t);
        q = s;
        if (t != 0.0) h[m++] = t;
    }
    if (q != 0.0 || m == 0) h[m++] = q;
    return m;
}

__host__ __device__ inline int orient(double ax, double ay, double bx, double by, double px, double py) {
    const double l = (bx - ax) * (py - ay), r = (by - ay) * (px - ax), det = l - r;
    const double bound = RS_ERRB * (fabs(l) + fabs(r));
    if (det > bound) return 1;
    if (-det > bound) return -1;
    double A[2], B[2], C[2], D[2];
    two_diff(bx, ax, A[1], A[0]);
    two_diff(py, ay, B[1], B[0]);
    two_diff(by, ay, C[1], C[0]);
    two_diff(px, ax, D[1], D[0]);
    double e[20], h[20];
    int n = 0;
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) {
            double p, q;
            two_prod(A[i], B[j], p, q);
            n = grow(e, n, q, h);
            for (int k = 0; k < n; ++k) e[k] = h[k];
            n = grow(e, n, p, h);
            for (int k = 0; k < n; ++k) e[k] = h[k];
            two_prod(C[i], D[j], p, q);
            n = grow(e, n, -q, h);
            for (int k = 0; k < n; ++k) e[k] = h[k];
            n = grow(e, n, -p, h);
            for (int k = 0; k < n; ++k) e[k] = h[k];
        }
    const double top = e[n - 1];
    return (top > 0.0) - (top < 0.0);
}

// (b - a) box holds p and p lies on the line through a, b: p on the closed segment
__device__ inline bool on_edge(const double* E, double px, double py) {
    const double ax = E[0], ay = E[1], bx = E[2], by = E[3];
    if (px < fmin(ax, bx) || px > fmax(ax, bx) || py < fmin(ay, by) || py > fmax(ay, by)) return false;
    return orient(ax, ay, bx, by, px, py) == 0;
}

// Bitonic sort of n <= RS_CAP (key, id) pairs in LDS by key then id (entries n .. pow2 padded with +inf).
__device__ void lds_sort(double* key, int* id, int n) {
    int m = 1;
    while (m < n) m <<= 1;
    for (int i = n + threadIdx.x; i < m; i += RS_NTH) {
        key[i] = INFINITY;
        id[i] = 0x7fffffff;
    }
    __syncthreads();
    for (int k = 2; k <= m; k <<= 1)
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = threadIdx.x; i < m; i += RS_NTH) {
                const int l = i ^ j;
                if (l > i) {
                    const bool up = (i & k) == 0;
                    const bool gt = key[i] > key[l] || (key[i] == key[l] && id[i] > id[l]);
                    if (gt == up) {
                        const double tk = key[i];
                        key[i] = key[l];
                        key[l] = tk;
                        const int ti = id[i];
                        id[i] = id[l];
                        id[l] = ti;
                    }
                }
            }
            __syncthreads();
        }
}

// ---- lattice labels
struct LabelArgs {
    const double* edges;        // the piece's edges, ax ay bx by
    const int* cptr;            // per window column: crossing edges (half-open in x) ...
    const int* cid;
    const int* sptr;            // ... and edges with an end on the column's x (vertices and vertical edges there)
    const int* sid;
    const double* xs;           // lattice column x (all columns)
    const double* ys;           // lattice sample y (all samples)
    int p0, q0, q1, nv;
    double eps;                 // crossing heights within eps of a sample are decided exactly
    int k;                      // the piece's label
    int* lab;                   // nu x nv
    signed char* state;         // nu x nv, or NULL: inside (1) and on an edge (2) written instead of labels
};

__global__ __launch_bounds__(RS_NTH) void rs_label(LabelArgs a) {
    __shared__ double yk[RS_CAP];
    __shared__ int yi[RS_CAP];
    const int c = blockIdx.x, p = a.p0 + c;
    const double X = a.xs[p];
    const int c0 = a.cptr[c], n = a.cptr[c + 1] - c0;
    for (int i = threadIdx.x; i < n; i += RS_NTH) {
        const int e = a.cid[c0 + i];
        const double* E = a.edges + 4 * (size_t)e;
        yk[i] = E[1] + (X - E[0]) * (E[3] - E[1]) / (E[2] - E[0]);
        yi[i] = e;
    }
    __syncthreads();
    lds_sort(yk, yi, n);
    const int s0 = a.sptr[c], s1 = a.sptr[c + 1];
    for (int q = a.q0 + threadIdx.x; q < a.q1; q += RS_NTH) {
        int* L = a.lab + (size_t)p * a.nv + q;
        if (!a.state && *L >= 0) continue;          // claimed by an earlier piece
        const double py = a.ys[q];
        int lo = 0, hi = n;                          // first crossing at or above py - eps
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (yk[mid] < py - a.eps) lo = mid + 1;
            else hi = mid;
        }
        int below = lo;
        bool edge = false;
        for (int i = lo; i < n && yk[i] <= py + a.eps; ++i) {
            const double* E = a.edges + 4 * (size_t)yi[i];
            const int o = orient(E[0], E[1], E[2], E[3], X, py);
            if (o == 0) edge = true;
            else if ((E[0] < E[2]) == (o > 0)) ++below;     // the edge passes below the sample
        }
        for (int i = s0; i < s1 && !edge; ++i) edge = on_edge(a.edges + 4 * (size_t)a.sid[i], X, py);
        if (a.state) a.state[(size_t)p * a.nv + q] = (signed char)((below & 1) | (edge ? 2 : 0));
        else if ((below & 1) && !edge) *L = a.k;
    }
}

// ---- the fine grid's cells
struct CellArgs {
    const int* lab;             // mx x my fine labels
    int nx, ny, s, my;
    int* part;                  // nx x ny: the kept piece's label, -1 none
    double* fill;
    unsigned char* on;          // mx x my: the sample belongs to its cell's kept piece
};

__global__ __launch_bounds__(RS_NTH) void rs_cells(CellArgs a) {
    const int cell = blockIdx.x * RS_NTH + threadIdx.x;
    if (cell >= a.nx * a.ny) return;
    const int I = cell / a.ny, J = cell - I * a.ny, s = a.s;
    int l[RS_SMAX * RS_SMAX], comp[RS_SMAX * RS_SMAX], stack[RS_SMAX * RS_SMAX];
    for (int u = 0; u < s; ++u)
        for (int v = 0; v < s; ++v) {
            l[u * s + v] = a.lab[(size_t)(I * s + u) * a.my + J * s + v];
            comp[u * s + v] = -1;
        }
    int best = -1, best_size = 0, nc = 0;
    for (int st = 0; st < s * s; ++st) {             // components in row-major order of their first sample
        if (l[st] < 0 || comp[st] >= 0) continue;
        int top = 0, size = 0;
        stack[top++] = st;
        comp[st] = nc;
        while (top > 0) {
            const int w = stack[--top];
            ++size;
            const int u = w / s, v = w - u * s;
            const int nb[4] = {u > 0 ? w - s : -1, u + 1 < s ? w + s : -1, v > 0 ? w - 1 : -1, v + 1 < s ? w + 1 : -1};
            for (int t = 0; t < 4; ++t)
                if (nb[t] >= 0 && comp[nb[t]] < 0 && l[nb[t]] == l[st]) {
                    comp[nb[t]] = nc;
                    stack[top++] = nb[t];
                }
        }
        if (size > best_size) {
            best_size = size;
            best = nc;
        }
        ++nc;
    }
    const bool keep = best >= 0 && best_size >= s;
    int label = -1;
    for (int w = 0; w < s * s; ++w) {
        const bool mine = keep && comp[w] == best;
        if (mine) label = l[w];
        a.on[(size_t)(I * s + w / s) * a.my + J * s + w % s] = mine;
    }
    a.part[cell] = label;
    a.fill[cell] = keep ? (double)best_size / (double)(s * s) : 0.0;
}

struct CrossArgs {
    const int* lab;
    const unsigned char* on;
    int nx, ny, s, my;
    double* cross_x;            // (nx - 1) x ny
    double* off_x;
    double* cross_y;            // nx x (ny - 1)
    double* off_y;
};

__global__ __launch_bounds__(RS_NTH) void rs_cross(CrossArgs a) {
    const int t = blockIdx.x * RS_NTH + threadIdx.x;
    const int nbx = (a.nx - 1) * a.ny, nby = a.nx * (a.ny - 1), s = a.s;
    if (t < nbx) {                                   // border between cells (I, J) and (I + 1, J)
        const int I = t / a.ny, J = t - I * a.ny, p = I * s + s - 1;
        double cr = 0.0, of = 0.0;
        for (int v = 0; v < s; ++v) {
            const size_t q = (size_t)J * s + v, i0 = (size_t)p * a.my + q, i1 = i0 + a.my;
            if (a.lab[i0] >= 0 && a.lab[i0] == a.lab[i1] && a.on[i0] && a.on[i1]) {
                cr += 1.0 / s;
                of += ((v + 0.5) / s - 0.5) / s;
            }
        }
        a.cross_x[t] = cr;
        a.off_x[t] = cr > 0 ? of / cr : 0.0;
    } else if (t < nbx + nby) {                      // border between cells (I, J) and (I, J + 1)
        const int u = t - nbx, I = u / (a.ny - 1), J = u - I * (a.ny - 1), q = J * s + s - 1;
        double cr = 0.0, of = 0.0;
        for (int w = 0; w < s; ++w) {
            const size_t i0 = (size_t)(I * s + w) * a.my + q;
            if (a.lab[i0] >= 0 && a.lab[i0] == a.lab[i0 + 1] && a.on[i0] && a.on[i0 + 1]) {
                cr += 1.0 / s;
                of += ((w + 0.5) / s - 0.5) / s;
            }
        }
        a.cross_y[u] = cr;
        a.off_y[u] = cr > 0 ? of / cr : 0.0;
    }
}

// ---- line covers
struct CoverArgs {
    const double* edges;        // ua va ub vb
    const int* lptr;            // per line: crossing edges
    const int* lid;
    double u0, du, v0, d;
    int n_cells;
    double* out;                // n_lines x n_cells
};

__global__ __launch_bounds__(RS_NTH) void rs_cover(CoverArgs a) {
    __shared__ double vk[RS_CAP];
    __shared__ int vi[RS_CAP];
    const int m = blockIdx.x;
    const int c0 = a.lptr[m], n = a.lptr[m + 1] - c0;
    const double u = a.u0 + (double)m * a.du;
    for (int i = threadIdx.x; i < n; i += RS_NTH) {
        const int e = a.lid[c0 + i];
        const double* E = a.edges + 4 * (size_t)e;
        vk[i] = E[1] + (u - E[0]) * (E[3] - E[1]) / (E[2] - E[0]);
        vi[i] = e;
    }
    __syncthreads();
    lds_sort(vk, vi, n);
    double* row = a.out + (size_t)m * a.n_cells;
    const int nc = a.n_cells;
    if (n & 1) {                                     // a degenerate line: no cover
        for (int c = threadIdx.x; c < nc; c += RS_NTH) row[c] = 0.0;
        return;
    }
    const double vlo = a.v0 - a.d / 2, top = nc * a.d, d = a.d;
    // spans clipped to [0, top] in place: vk[2i], vk[2i + 1]
    for (int i = threadIdx.x; i < n; i += RS_NTH) vk[i] = fmin(fmax(vk[i] - vlo, 0.0), top);
    __syncthreads();
    const int ns = n / 2;
    for (int c = threadIdx.x; c < nc; c += RS_NTH) {
        // the first span whose last cell is c or later
        int lo = 0, hi = ns;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            const double b = vk[2 * mid + 1];
            if (min((int)(b / d), nc - 1) < c) lo = mid + 1;
            else hi = mid;
        }
        double acc = 0.0, full = 0.0;
        for (int i = lo; i < ns; ++i) {
            const double sa = vk[2 * i], sb = vk[2 * i + 1];
            if (!(sb > sa)) continue;
            const int ca = min((int)(sa / d), nc - 1), cb = min((int)(sb / d), nc - 1);
            if (ca > c) break;
            if (ca == cb) acc += sb - sa;
            else if (c == ca) acc += (ca + 1) * d - sa;
            else if (c == cb) acc += sb - cb * d;
            else full += d;
        }
        row[c] = acc + full;
    }
}

// ---- point location and nearest points
struct PointArgs {
    const double* edges;
    long long ne;
    const double* pts;          // x y
    long long np;
    signed char* inside;
    signed char* onb;
    double* nearest;            // x y
};

__global__ __launch_bounds__(RS_NTH) void rs_locate(PointArgs a) {
    const long long t = (long long)blockIdx.x * RS_NTH + threadIdx.x;
    if (t >= a.np) return;
    const double px = a.pts[2 * t], py = a.pts[2 * t + 1];
    bool par = false, edge = false;
    for (long long e = 0; e < a.ne; ++e) {
        const double* E = a.edges + 4 * e;
        const double ax = E[0], ay = E[1], bx = E[2], by = E[3];
        const bool up = ay <= py && py < by, down = by <= py && py < ay;
        const bool inbox = fmin(ax, bx) <= px && px <= fmax(ax, bx) && fmin(ay, by) <= py && py <= fmax(ay, by);
        if (!(up || down || inbox)) continue;
        const int o = orient(ax, ay, bx, by, px, py);
        if ((up && o > 0) || (down && o < 0)) par = !par;
        if (inbox && o == 0) edge = true;
    }
    a.inside[t] = par;
    a.onb[t] = edge;
}

__global__ __launch_bounds__(RS_NTH) void rs_nearest(PointArgs a) {
    const long long t = (long long)blockIdx.x * RS_NTH + threadIdx.x;
    if (t >= a.np) return;
    const double px = a.pts[2 * t], py = a.pts[2 * t + 1];
    double best = INFINITY, bx_ = 0.0, by_ = 0.0;
    for (long long e = 0; e < a.ne; ++e) {
        const double* E = a.edges + 4 * e;
        const double ux = E[2] - E[0], uy = E[3] - E[1];
        const double L = ux * ux + uy * uy;
        double tt = L > 0 ? ((px - E[0]) * ux + (py - E[1]) * uy) / L : 0.0;
        tt = fmin(fmax(tt, 0.0), 1.0);
        const double qx = E[0] + tt * ux, qy = E[1] + tt * uy;
        const double d2 = (qx - px) * (qx - px) + (qy - py) * (qy - py);
        if (d2 < best) {
            best = d2;
            bx_ = qx;
            by_ = qy;
        }
    }
    a.nearest[2 * t] = bx_;
    a.nearest[2 * t + 1] = by_;
}

// Nearest points by bins: the edges listed in every bin their bounding box meets; a point searches rings of bins
// outward from its own (unclamped) bin and stops once its best distance is below the ring's reach, since an edge not
// yet seen lies outside the square searched. Ties go to the lower edge index, as a scan in order gives them.
struct BinArgs {
    const double* edges;
    const int* bptr;            // per bin (row-major, bx major): edges
    const int* bid;
    int nbx, nby;
    double bx0, by0, bsz;
    const double* pts;
    long long np;
    double* nearest;
};

__global__ __launch_bounds__(RS_NTH) void rs_nearest_binned(BinArgs a) {
    const long long t = (long long)blockIdx.x * RS_NTH + threadIdx.x;
    if (t >= a.np) return;
    const double px = a.pts[2 * t], py = a.pts[2 * t + 1];
    const int ci = (int)floor((px - a.bx0) / a.bsz), cj = (int)floor((py - a.by0) / a.bsz);
    double best = INFINITY, bx_ = 0.0, by_ = 0.0;
    long long be = -1;
    const int rmax = max(max(abs(ci), abs(ci - a.nbx + 1)), max(abs(cj), abs(cj - a.nby + 1)));
    for (int r = 0; r <= rmax; ++r) {
        for (int i = ci - r; i <= ci + r; ++i) {
            if (i < 0 || i >= a.nbx) continue;
            const bool side = i == ci - r || i == ci + r;
            for (int j = cj - r; j <= cj + r; j += (side ? 1 : 2 * r)) {
                if (j >= 0 && j < a.nby) {
                    const int b = i * a.nby + j;
                    for (int k = a.bptr[b]; k < a.bptr[b + 1]; ++k) {
                        const int e = a.bid[k];
                        const double* E = a.edges + 4 * (size_t)e;
                        const double ux = E[2] - E[0], uy = E[3] - E[1];
                        const double L = ux * ux + uy * uy;
                        double tt = L > 0 ? ((px - E[0]) * ux + (py - E[1]) * uy) / L : 0.0;
                        tt = fmin(fmax(tt, 0.0), 1.0);
                        const double qx = E[0] + tt * ux, qy = E[1] + tt * uy;
                        const double d2 = (qx - px) * (qx - px) + (qy - py) * (qy - py);
                        if (d2 < best || (d2 == best && e < be)) {
                            best = d2;
                            be = e;
                            bx_ = qx;
                            by_ = qy;
                        }
                    }
                }
                if (r == 0) break;
            }
        }
        const double reach = r * a.bsz;
        if (best < reach * reach * (1.0 - 1e-12)) break;
    }
    a.nearest[2 * t] = bx_;
    a.nearest[2 * t + 1] = by_;
}

// first index in sorted xs with xs[i] >= v (lower) or > v (upper)
int lower(const double* xs, int n, double v) { return (int)(std::lower_bound(xs, xs + n, v) - xs); }
int upper(const double* xs, int n, double v) { return (int)(std::upper_bound(xs, xs + n, v) - xs); }

// One label pass per piece over its window of the lattice, the pieces in order (the first to claim keeps a sample).
void label_lattice(gt_ctx* ctx, int ng, const int64_t* eptr, const double* edges, const int32_t* win, int nu,
                   const double* xs_h, int nv, const double* ys_h, int* lab_d, const DevBuf<double>& xs,
                   const DevBuf<double>& ys, signed char* state_d = nullptr) {
    hipStream_t st = ctx->stream;
    for (int i = 1; i < nu; ++i)
        if (!(xs_h[i] > xs_h[i - 1])) fail(GT_ERR_ARG, "lattice columns not increasing");
    // the lists of every piece in one upload
    std::vector<int> cptr, cid, sptr, sid;
    std::vector<size_t> c_at(ng), s_at(ng), cid_at(ng), sid_at(ng);
    std::vector<double> eps(ng);
    for (int k = 0; k < ng; ++k) {
        const int p0 = std::max(0, win[4 * k]), p1 = std::min(nu, win[4 * k + 1]);
        const int64_t e0 = eptr[k], e1 = eptr[k + 1];
        c_at[k] = cptr.size();
        s_at[k] = sptr.size();
        cid_at[k] = cid.size();
        sid_at[k] = sid.size();
        const int nc = std::max(0, p1 - p0);
        std::vector<int> cc(nc + 1, 0), sc(nc + 1, 0);
        double big = 0.0;
        auto each = [&](auto&& f) {
            for (int64_t e = e0; e < e1; ++e) {
                const double* E = edges + 4 * e;
                big = std::max({big, std::fabs(E[0]), std::fabs(E[1]), std::fabs(E[2]), std::fabs(E[3])});
                const double lo = std::min(E[0], E[2]), hi = std::max(E[0], E[2]);
                const int a0 = std::max(lower(xs_h, nu, lo), p0), a1 = std::min(lower(xs_h, nu, hi), p1);
                f(0, e - e0, a0, a1);
                for (const double x : {E[0], E[2]}) {
                    const int b0 = std::max(lower(xs_h, nu, x), p0), b1 = std::min(upper(xs_h, nu, x), p1);
                    f(1, e - e0, b0, b1);
                    if (E[0] == E[2]) break;                 // a vertical edge: one column
                }
            }
        };
        each([&](int kind, int64_t, int a0, int a1) {
            for (int p = a0; p < a1; ++p) (kind ? sc : cc)[p - p0 + 1]++;
        });
        for (int c = 0; c < nc; ++c) {
            cc[c + 1] += cc[c];
            sc[c + 1] += sc[c];
            if (cc[c + 1] - cc[c] > RS_CAP) fail(GT_ERR_ARG, "a lattice column crosses more than 4096 edges of one piece");
        }
        std::vector<int> cl(cc.back()), sl(sc.back()), cf(cc.begin(), cc.end() - 1), sf(sc.begin(), sc.end() - 1);
        each([&](int kind, int64_t e, int a0, int a1) {
            for (int p = a0; p < a1; ++p) {
                if (kind) sl[sf[p - p0]++] = (int)e;
                else cl[cf[p - p0]++] = (int)e;
            }
        });
        cptr.insert(cptr.end(), cc.begin(), cc.end());
        sptr.insert(sptr.end(), sc.begin(), sc.end());
        cid.insert(cid.end(), cl.begin(), cl.end());
        sid.insert(sid.end(), sl.begin(), sl.end());
        eps[k] = 1e-9 * (1.0 + big);
    }
    DevBuf<double> E(4 * (size_t)eptr[ng]);
    E.upload(edges, E.n, st);
    DevBuf<int> dcp(cptr.size()), dci(std::max<size_t>(1, cid.size())), dsp(sptr.size()),
        dsi(std::max<size_t>(1, sid.size()));
    dcp.upload(cptr.data(), cptr.size(), st);
    dci.upload(cid.data(), cid.size(), st);
    dsp.upload(sptr.data(), sptr.size(), st);
    dsi.upload(sid.data(), sid.size(), st);
    for (int k = 0; k < ng; ++k) {
        const int p0 = std::max(0, win[4 * k]), p1 = std::min(nu, win[4 * k + 1]);
        const int q0 = std::max(0, win[4 * k + 2]), q1 = std::min(nv, win[4 * k + 3]);
        if (p0 >= p1 || q0 >= q1 || eptr[k + 1] == eptr[k]) continue;
        LabelArgs a{E.p + 4 * eptr[k], dcp.p + c_at[k], dci.p + cid_at[k], dsp.p + s_at[k], dsi.p + sid_at[k],
                    xs.p, ys.p, p0, q0, q1, nv, eps[k], k, lab_d, state_d};
        hipLaunchKernelGGL(rs_label, dim3(p1 - p0), dim3(RS_NTH), 0, st, a);
        GT_LAUNCH("rs_label");
    }
}

void need_arg(bool ok, const char* msg) {
    if (!ok) fail(GT_ERR_ARG, msg);
}

void check_groups(int ng, const int64_t* eptr, const double* edges, const int32_t* win) {
    need_arg(ng >= 0 && (ng == 0 || (eptr && win)), "NULL argument");
    for (int k = 0; k < ng; ++k)
        if (eptr[k + 1] < eptr[k]) fail(GT_ERR_ARG, "edge offsets decrease");
    if (ng && eptr[ng] > 0 && !edges) fail(GT_ERR_ARG, "NULL edges");
    if (ng && eptr[ng] >= (int64_t(1) << 31)) fail(GT_ERR_ARG, "more than 2^31 edges");
}

}  // namespace

void raster_label(gt_ctx* ctx, int32_t ng, const int64_t* eptr, const double* edges, const int32_t* win, int32_t nu,
                  const double* xs, int32_t nv, const double* ys, int32_t* lab) {
    check_groups(ng, eptr, edges, win);
    need_arg(nu > 0 && nv > 0 && xs && ys && lab, "bad lattice");
    hipStream_t st = ctx->stream;
    DevBuf<double> dx(nu), dy(nv);
    dx.upload(xs, nu, st);
    dy.upload(ys, nv, st);
    DevBuf<int> L((size_t)nu * nv);
    L.upload(lab, L.n, st);
    label_lattice(ctx, ng, eptr, edges, win, nu, xs, nv, ys, L.p, dx, dy);
    GT_HIP(hipMemcpyAsync(lab, L.p, L.bytes(), hipMemcpyDeviceToHost, st));
    GT_HIP(hipStreamSynchronize(st));
}

void raster_state(gt_ctx* ctx, int64_t ne, const double* edges, const int32_t* win, int32_t nu, const double* xs,
                  int32_t nv, const double* ys, int8_t* state) {
    need_arg(ne >= 0 && (ne == 0 || edges) && win && xs && ys && state && nu > 0 && nv > 0, "bad argument");
    if (ne >= (int64_t(1) << 31)) fail(GT_ERR_ARG, "more than 2^31 edges");
    hipStream_t st = ctx->stream;
    DevBuf<double> dx(nu), dy(nv);
    dx.upload(xs, nu, st);
    dy.upload(ys, nv, st);
    DevBuf<signed char> S((size_t)nu * nv);
    GT_HIP(hipMemsetAsync(S.p, 0, S.bytes(), st));
    const int64_t eptr[2] = {0, ne};
    DevBuf<int> none(1);
    label_lattice(ctx, 1, eptr, edges, win, nu, xs, nv, ys, none.p, dx, dy, S.p);
    GT_HIP(hipMemcpyAsync(state, S.p, S.bytes(), hipMemcpyDeviceToHost, st));
    GT_HIP(hipStreamSynchronize(st));
}

void raster_fine(gt_ctx* ctx, int32_t ng, const int64_t* eptr, const double* edges, const int32_t* win, int32_t nx,
                 int32_t ny, int32_t s, const double* xf, const double* yf, int32_t* part, double* fill,
                 double* cross_x, double* off_x, double* cross_y, double* off_y, int32_t* fine) {
    check_groups(ng, eptr, edges, win);
    need_arg(nx >= 2 && ny >= 2 && s >= 2 && s <= RS_SMAX, "sub-grid factor must be 2 .. 8 and the grid 2 x 2 or more");
    need_arg(xf && yf && part && fill && cross_x && off_x && cross_y && off_y, "NULL argument");
    hipStream_t st = ctx->stream;
    const int mx = nx * s, my = ny * s;
    DevBuf<double> dx(mx), dy(my);
    dx.upload(xf, mx, st);
    dy.upload(yf, my, st);
    DevBuf<int> L((size_t)mx * my);
    GT_HIP(hipMemsetAsync(L.p, 0xff, L.bytes(), st));          // every sample -1: no copper
    label_lattice(ctx, ng, eptr, edges, win, mx, xf, my, yf, L.p, dx, dy);
    const size_t ncell = (size_t)nx * ny, nbx = (size_t)(nx - 1) * ny, nby = (size_t)nx * (ny - 1);
    DevBuf<unsigned char> on((size_t)mx * my);
    DevBuf<int> dpart(ncell);
    DevBuf<double> out(ncell + 2 * nbx + 2 * nby);
    double* dfill = out.p;
    CellArgs ca{L.p, nx, ny, s, my, dpart.p, dfill, on.p};
    hipLaunchKernelGGL(rs_cells, dim3((unsigned)((ncell + RS_NTH - 1) / RS_NTH)), dim3(RS_NTH), 0, st, ca);
    GT_LAUNCH("rs_cells");
    CrossArgs cr{L.p, on.p, nx, ny, s, my, out.p + ncell, out.p + ncell + nbx, out.p + ncell + 2 * nbx,
                 out.p + ncell + 2 * nbx + nby};
    hipLaunchKernelGGL(rs_cross, dim3((unsigned)((nbx + nby + RS_NTH - 1) / RS_NTH)), dim3(RS_NTH), 0, st, cr);
    GT_LAUNCH("rs_cross");
    std::vector<double> h(out.n);
    GT_HIP(hipMemcpyAsync(part, dpart.p, dpart.bytes(), hipMemcpyDeviceToHost, st));
    GT_HIP(hipMemcpyAsync(h.data(), out.p, out.bytes(), hipMemcpyDeviceToHost, st));
    if (fine) GT_HIP(hipMemcpyAsync(fine, L.p, L.bytes(), hipMemcpyDeviceToHost, st));
    GT_HIP(hipStreamSynchronize(st));
    std::copy(h.begin(), h.begin() + ncell, fill);
    std::copy(h.begin() + ncell, h.begin() + ncell + nbx, cross_x);
    std::copy(h.begin() + ncell + nbx, h.begin() + ncell + 2 * nbx, off_x);
    std::copy(h.begin() + ncell + 2 * nbx, h.begin() + ncell + 2 * nbx + nby, cross_y);
    std::copy(h.begin() + ncell + 2 * nbx + nby, h.end(), off_y);
}

void raster_cover(gt_ctx* ctx, int64_t ne, const double* edges, double u0, double du, int32_t n_lines, double v0,
                  double d, int32_t n_cells, double* out) {
    need_arg(ne >= 0 && (ne == 0 || edges) && out && n_lines >= 0 && n_cells >= 0 && du > 0 && d > 0, "bad argument");
    if (n_lines == 0 || n_cells == 0) return;
    if (ne >= (int64_t(1) << 31)) fail(GT_ERR_ARG, "more than 2^31 edges");
    hipStream_t st = ctx->stream;
    // the lines each edge crosses, half open (peec.py:557-559)
    std::vector<int> cnt(n_lines + 1, 0);
    std::vector<int> m0(ne), m1(ne);
    for (int64_t e = 0; e < ne; ++e) {
        const double* E = edges + 4 * e;
        const double lo = std::min(E[0], E[2]), hi = std::max(E[0], E[2]);
        const auto clip = [&](double v) { return (int)std::min<double>(std::max(std::ceil(v), 0.0), n_lines); };
        m0[e] = clip((lo - u0) / du);
        m1[e] = clip((hi - u0) / du);
        for (int m = m0[e]; m < m1[e]; ++m) cnt[m + 1]++;
    }
    for (int m = 0; m < n_lines; ++m) {
        if (cnt[m + 1] > RS_CAP) fail(GT_ERR_ARG, "a cover line crosses more than 4096 edges");
        cnt[m + 1] += cnt[m];
    }
    std::vector<int> ids(std::max(1, cnt.back())), at(cnt.begin(), cnt.end() - 1);
    for (int64_t e = 0; e < ne; ++e)
        for (int m = m0[e]; m < m1[e]; ++m) ids[at[m]++] = (int)e;
    DevBuf<double> E(std::max<int64_t>(1, 4 * ne));
    E.upload(edges, 4 * ne, st);
    DevBuf<int> lp(cnt.size()), li(ids.size());
    lp.upload(cnt.data(), cnt.size(), st);
    li.upload(ids.data(), ids.size(), st);
    DevBuf<double> o((size_t)n_lines * n_cells);
    CoverArgs a{E.p, lp.p, li.p, u0, du, v0, d, n_cells, o.p};
    hipLaunchKernelGGL(rs_cover, dim3(n_lines), dim3(RS_NTH), 0, st, a);
    GT_LAUNCH("rs_cover");
    GT_HIP(hipMemcpyAsync(out, o.p, o.bytes(), hipMemcpyDeviceToHost, st));
    GT_HIP(hipStreamSynchronize(st));
}

void raster_points(gt_ctx* ctx, int64_t ne, const double* edges, int64_t np, const double* pts, int8_t* inside,
                   int8_t* onb, double* nearest) {
    need_arg(ne >= 0 && np >= 0 && (ne == 0 || edges) && (np == 0 || pts), "NULL argument");
    if (np == 0) return;
    hipStream_t st = ctx->stream;
    DevBuf<double> E(std::max<int64_t>(1, 4 * ne)), P(2 * np);
    E.upload(edges, 4 * ne, st);
    P.upload(pts, 2 * np, st);
    const dim3 grid((unsigned)((np + RS_NTH - 1) / RS_NTH));
    if (inside || onb) {
        DevBuf<signed char> io(2 * np);
        PointArgs a{E.p, ne, P.p, np, io.p, io.p + np, nullptr};
        hipLaunchKernelGGL(rs_locate, grid, dim3(RS_NTH), 0, st, a);
        GT_LAUNCH("rs_locate");
        std::vector<signed char> h(2 * np);
        GT_HIP(hipMemcpyAsync(h.data(), io.p, io.bytes(), hipMemcpyDeviceToHost, st));
        GT_HIP(hipStreamSynchronize(st));
        if (inside) std::copy(h.begin(), h.begin() + np, inside);
        if (onb) std::copy(h.begin() + np, h.end(), onb);
    }
    if (nearest) {
        if (ne == 0) fail(GT_ERR_ARG, "nearest points of no edges");
        DevBuf<double> q(2 * np);
        if (ne <= 256) {
            PointArgs a{E.p, ne, P.p, np, nullptr, nullptr, q.p};
            hipLaunchKernelGGL(rs_nearest, grid, dim3(RS_NTH), 0, st, a);
            GT_LAUNCH("rs_nearest");
        } else {
            if (ne >= (int64_t(1) << 31)) fail(GT_ERR_ARG, "more than 2^31 edges");
            double x0 = INFINITY, y0 = INFINITY, x1 = -INFINITY, y1 = -INFINITY;
            for (int64_t e = 0; e < ne; ++e) {
                const double* Q = edges + 4 * e;
                x0 = std::min({x0, Q[0], Q[2]});
                x1 = std::max({x1, Q[0], Q[2]});
                y0 = std::min({y0, Q[1], Q[3]});
                y1 = std::max({y1, Q[1], Q[3]});
            }
            // about 8 edges a bin, at most 2048 bins a side
            const double span = std::max({x1 - x0, y1 - y0, 1e-9});
            const double area = std::max((x1 - x0) * (y1 - y0), span * span * 1e-6);
            double bsz = std::max(std::sqrt(area * 8.0 / (double)ne), span / 2048.0);
            const int nbx = std::max(1, (int)std::ceil((x1 - x0) / bsz) + 1);
            const int nby = std::max(1, (int)std::ceil((y1 - y0) / bsz) + 1);
            std::vector<int> cnt((size_t)nbx * nby + 1, 0);
            auto range = [&](const double* Q, int& i0, int& i1, int& j0, int& j1) {
                i0 = std::clamp((int)std::floor((std::min(Q[0], Q[2]) - x0) / bsz), 0, nbx - 1);
                i1 = std::clamp((int)std::floor((std::max(Q[0], Q[2]) - x0) / bsz), 0, nbx - 1);
                j0 = std::clamp((int)std::floor((std::min(Q[1], Q[3]) - y0) / bsz), 0, nby - 1);
                j1 = std::clamp((int)std::floor((std::max(Q[1], Q[3]) - y0) / bsz), 0, nby - 1);
            };
            for (int64_t e = 0; e < ne; ++e) {
                int i0, i1, j0, j1;
                range(edges + 4 * e, i0, i1, j0, j1);
                for (int i = i0; i <= i1; ++i)
                    for (int j = j0; j <= j1; ++j) cnt[(size_t)i * nby + j + 1]++;
            }
            for (size_t b = 0; b + 1 < cnt.size(); ++b) cnt[b + 1] += cnt[b];
            if (cnt.back() >= (int64_t(1) << 31) - 1) fail(GT_ERR_ARG, "too many edge bins");
            std::vector<int> ids(std::max(1, cnt.back())), at(cnt.begin(), cnt.end() - 1);
            for (int64_t e = 0; e < ne; ++e) {
                int i0, i1, j0, j1;
                range(edges + 4 * e, i0, i1, j0, j1);
                for (int i = i0; i <= i1; ++i)
                    for (int j = j0; j <= j1; ++j) ids[at[(size_t)i * nby + j]++] = (int)e;
            }
            DevBuf<int> bp(cnt.size()), bi(ids.size());
            bp.upload(cnt.data(), cnt.size(), st);
            bi.upload(ids.data(), ids.size(), st);
            BinArgs a{E.p, bp.p, bi.p, nbx, nby, x0, y0, bsz, P.p, np, q.p};
            hipLaunchKernelGGL(rs_nearest_binned, grid, dim3(RS_NTH), 0, st, a);
            GT_LAUNCH("rs_nearest_binned");
        }
        GT_HIP(hipMemcpyAsync(nearest, q.p, q.bytes(), hipMemcpyDeviceToHost, st));
        GT_HIP(hipStreamSynchronize(st));
    }
}

}  // namespace gt
