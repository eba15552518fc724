// Launch arguments of the L_p product kernels (lp.hip), for a batch of right-hand sides.
#pragma once

#include "gt_common.hpp"

namespace gt {

constexpr int LP_MAX_SHEETS = 8;
constexpr int KMAX = 16;              // right-hand sides one batched launch takes

// Points one workgroup's LDS holds in the row passes (a tile of rows) and in the column pass (every sheet's
// column): 64 KiB of complex values.
template <class T>
struct LpPts;
// points of the LDS passes' workgroup buffers: rows / cols for the usual sizes, big for the long transforms (a row
// of 4096 points, or S * P0 up to 8192 in the column pass: 128 KB of FP64 LDS, launched only where needed)
template <>
struct LpPts<double> { static constexpr int rows = 4096, cols = 4096, big = 8192; };
template <>
struct LpPts<float> { static constexpr int rows = 8192, cols = 8192, big = 8192; };

template <class T>
struct OrientArgs {
    int n0, n1;               // dense strip grid: rows (axis i) and columns (axis j)
    int P0, P1, log0, log1;   // FFT sizes
    int TR;                   // rows per workgroup in the row passes
    int lo, cnt;              // the orientation's strips: x[lo, lo + cnt)
    const int* rowptr;        // S * n0 + 1: first strip of each (sheet, row), relative to lo
    const int* rowcol;        // cnt: (i << 16) | j of each strip
    const T* K;               // kernel quadrant [qj][qi][g], scaled by 1 / (P0 P1): one value per distinct face gap g
    cpx<T>* T1;               // [s][j][i] after the row FFT (S * P1 * n0), right-hand side 0
    cpx<T>* T2;               // [t][j][i] after the column pass
    const cpx<T>* tw0;        // twiddles of P0
    const cpx<T>* tw1;        // twiddles of P1
    const cpx<T>* twq;        // twiddles of P0 / 4 (the residue-class column pass)
};

// The right-hand sides of one batched product: input and output vectors, and each one's operator data.
template <class T>
struct RhsSet {
    int nr;
    const cpx<T>* x[KMAX];
    cpx<T>* y[KMAX];
    T w[KMAX];                        // angular frequency
    const cpx<T>* r_sheet[KMAX];      // S strip impedances
    const cpx<T>* z_tail[KMAX];       // branch impedances
    int nb[KMAX];                     // offset of the node potentials in x and y (the stack's branch count)
};

template <class T>
struct LpArgs {
    OrientArgs<T> o[2];
    int nblk_rows[2];         // workgroups of each orientation in the row passes
    int nblk_cols[2];         // workgroups of each orientation in the column pass (= its P1)
    int S;
    int mode;                 // 0: y = L_p x (strip vectors); 1: the system product's strip rows (node potentials
                              // at rhs.nb[p] in x)
    size_t t_stride;          // elements of T1 / T2 per right-hand side
    int G;                    // distinct kernels (face gaps) per pixel; sheets t, s read K[.][gmap[t S + s]]
    const int* gmap;          // S x S, device memory
    int gpad[LP_MAX_SHEETS * LP_MAX_SHEETS];   // the same map at row stride LP_MAX_SHEETS, in the arguments
    int toeplitz;             // the map is g = |t - s| with G = S (equally spaced sheets)
    const int* a_keep;
    const int* b_keep;
    RhsSet<T> rhs;
};

// The branch and node rows of the product. The right-hand sides may belong to different stacks of one core (a
// board and its what-ifs): the strips are common, each has its own branches, incidence and node adjacency.
template <class T>
struct TailArgs {
    int ns, nk;
    int nbr[KMAX];
    const int* a_keep[KMAX];
    const int* b_keep[KMAX];
    const int* adj_ptr[KMAX]; // nk + 1
    const int* adj[KMAX];     // (branch << 1) | (1 if the node is the branch's b end)
    RhsSet<T> rhs;
};

struct LpShape {
    size_t lds_rows = 0, lds_cols = 0;
    // launch shapes of the register passes per batch-size class (1; 2-4; 5-16 right-hand sides): rows per row-pass
    // workgroup and waves per column-pass workgroup. They change no arithmetic; an autotuner may set them.
    int rw[3] = {4, 4, 4}, cw[3] = {1, 1, 1};
    // the column pass by residue classes of bins mod 4 (lp_cols_res4), for stacks whose S sheets of a column
    // exceed one workgroup's LDS
    bool split = false;
};

template <class T>
void launch_lp(const LpArgs<T>& a, const LpShape& sh, hipStream_t st);
template <class T>
void launch_tail_bottom(const TailArgs<T>& a, hipStream_t st);

}  // namespace gt
