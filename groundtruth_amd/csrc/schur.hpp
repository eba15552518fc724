// The Schur complement factor: B = Ar^T Ar over the kept nodes (the graph Laplacian of strips and branches, one
// reference node per connected piece removed, so B is symmetric positive definite), ordered by geometric nested
// dissection and factorised B = L L^T by a multifrontal supernodal Cholesky on the host; the supernodes are the
// nested-dissection tree's nodes (leaf boxes and separators), so the assembly tree has the dissection's depth.
// The device keeps the panels and solves L L^T x = b level by level (schur_solve.hip).
#pragma once

#include <vector>

#include "gt_common.hpp"

namespace gt {

struct SchurHost {
    int n = 0;                          // kept nodes
    std::vector<int> perm;              // new -> old
    // supernodes in postorder (children first); columns of supernode s are new indices [c0[s], c0[s] + nc[s])
    std::vector<int> c0, nc, nr;        // first column, columns, off-diagonal rows
    std::vector<int64_t> roff;          // offset of s's off-diagonal rows in `rows`
    std::vector<int> rows;              // off-diagonal row indices (new numbering), ascending per supernode
    std::vector<int64_t> poff;          // offset of s's panel in `panel`: (nc + nr) x nc, column-major
    // the solve panel Q = [L_CC^-1 ; -L_RC L_CC^-1] (column-major, m x nc): the forward solve of a supernode is
    // [x_C ; u_R] = Q f_C + [0 ; f_R] and the backward x_C = Q^T [z_C ; x_R]: dense products, no serial chain
    std::vector<double> panel;
    std::vector<int> parent;            // assembly-tree parent (-1 at a root)
    std::vector<int> child_ptr, child;  // children per supernode (CSR, ascending)
    std::vector<int64_t> uoff;          // offset of s's update vector (nr entries) in the update buffer
    int64_t upd_total = 0;
    std::vector<int64_t> fpos;          // offset of s's front positions (nc + nr) in the contributor CSR
    std::vector<int64_t> cptr;          // contributors of each front position: update-buffer indices, children
    std::vector<int64_t> cidx;          //   in ascending order (a fixed summation order)
    int levels = 0;
    std::vector<int> fwd_ptr, bwd_ptr;  // work items per level (forward: leaves first; backward: roots first)
    std::vector<int> fwd_s, fwd_r0;     // forward items: supernode, first row of its block
    std::vector<int> bwd_s, bwd_q0, bwd_qn;   // backward items: supernode, first column, columns
    int64_t nnz_L = 0;
    int max_cols = 0;

    // factorise B given by its strictly-off-diagonal edges (a, b pairs, weight 1 each, repeated edges add) and the
    // diagonal (degree incl. edges to reference nodes); xyz: (sheet, i, j) per node
    void build(int n, const std::vector<int>& ea, const std::vector<int>& eb, const std::vector<double>& diag,
               const int32_t* xyz);
    // x = B^-1 b on the host (complex values interleaved), with the same panels as the device
    void solve_host(const double* b, double* x) const;
};

constexpr int SCHUR_ROWS = 64;         // panel rows per forward work item
constexpr int SCHUR_COLS = 16;         // panel columns per backward work item of a wide supernode
constexpr int SCHUR_COLS_MAX = 64;     // a supernode of at most this many columns is one backward item

struct Schur {
    SchurHost h;
    DevBuf<int> perm, c0, nc, nr, rows, fwd_s, fwd_r0, bwd_s, bwd_q0, bwd_qn;
    DevBuf<int64_t> roff, poff, uoff, fpos, cptr, cidx;
    DevBuf<double> panel;
    DevBuf<zd> ubuf;                    // update vectors of one forward solve (up to 16 right-hand sides)
    DevBuf<zd> zp, xp;                  // forward and backward solutions in the new numbering

    void upload(hipStream_t s);
    int64_t device_bytes() const;
};

// x = B^-1 b on the device for nr <= 16 right-hand sides, interleaved: entry (node, p) at node * nr + p, in the old
// numbering (b and x distinct)
void schur_solve_device(Schur& S, const zd* b, zd* x, int nr, hipStream_t s);

}  // namespace gt
