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
    std::vector<double> panel;
    std::vector<int> parent;            // assembly-tree parent (-1 at a root)
    std::vector<int> child_ptr, child;  // children per supernode (CSR, ascending)
    std::vector<int64_t> rel_off;       // offset of s's row map in `rel` (nr[s] entries)
    std::vector<int> rel;               // position of each off-diagonal row of s in its parent's front
    std::vector<int64_t> foff;          // offset of s's front vector (nc + nr) in the device scratch
    int64_t front_total = 0;
    std::vector<int> fwd_levels_ptr, fwd_levels;   // supernodes by height (leaves first)
    std::vector<int> bwd_levels_ptr, bwd_levels;   // supernodes by depth (roots first)
    int64_t nnz_L = 0;
    int max_cols = 0;

    // factorise B given by its strictly-off-diagonal edges (a, b pairs, weight 1 each, repeated edges add) and the
    // diagonal (degree incl. edges to reference nodes); xyz: (sheet, i, j) per node
    void build(int n, const std::vector<int>& ea, const std::vector<int>& eb, const std::vector<double>& diag,
               const int32_t* xyz);
    // x = B^-1 b on the host (complex values interleaved)
    void solve_host(const double* b, double* x) const;
};

struct Schur {
    SchurHost h;
    DevBuf<int> perm, c0, nc, nr, child_ptr, child, rows, rel, fwd_levels, bwd_levels;
    DevBuf<int64_t> roff, poff, rel_off, foff;
    DevBuf<double> panel;
    DevBuf<zd> fvec;                    // front vectors of one solve
    DevBuf<zd> xp;                      // the solution in the new numbering

    void upload(hipStream_t s);
    int64_t device_bytes() const;
};

// x = B^-1 b on the device; b and x hold n complex values in the old numbering (may alias? no)
void schur_solve_device(Schur& S, const zd* b, zd* x, hipStream_t s);

}  // namespace gt
