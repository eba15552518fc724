// The Schur complement factor: B = Ar^T Ar over the kept nodes (the graph Laplacian of strips and branches, one
// reference node per connected piece removed, so B is symmetric positive definite), ordered by geometric nested
// dissection (each sheet's grid on its own under a top separator of the nodes that branches join across sheets,
// or all sheets cut together) and factorised B = L L^T by a multifrontal supernodal Cholesky on the host; the supernodes are the
// nested-dissection tree's nodes (leaf boxes and separators), so the assembly tree has the dissection's depth.
// The device keeps the panels and solves L L^T x = b level by level (schur_solve.hip).
#pragma once

#include <cstdint>
#include <memory>
#include <new>
#include <utility>

#include <vector>

#include "gt_common.hpp"

namespace gt {

// One work item of a device solve with everything it needs about its supernode (one load). An item covers one chunk
// [k0, k1) of its tile's reduction (columns forward, rows backward); a tile whose reduction spans several chunks is a
// group of items that write partial sums to slots slot0 + kidx, added in order by a finalize item (one per group,
// launched after the level).
struct alignas(16) SnItem {
    int32_t s;          // supernode
    int32_t t0;         // first row of the tile (forward) or first column (backward)
    int32_t kw;         // forward: column phases per wave x 1 (64-row tile) or x 4 (16-row tile); backward: row phases
    int32_t C;          // columns
    int32_t m;          // rows (columns + off-diagonal rows)
    int32_t c0;         // first column (new numbering)
    int32_t k0, k1;     // the chunk of the reduction
    int32_t slot0;      // first partial-sum slot of the group, -1 for a tile of one chunk
    int32_t kidx;       // chunk index in the group
    int32_t ks;         // chunks in the group
    int32_t pad0;
    int64_t poff;       // panel offset
    int64_t aux;        // forward: update-vector offset; backward: offset of the off-diagonal row list
    int64_t fpos;       // first front position
    int64_t pad;
};

// A work item of the top of the tree's dense products (schur_solve.hip): forward, rows [i0, i0 + cnt) of one top
// supernode (top-column indices) over the chunk [k0, k1) of their reduction; backward, columns [i0, i0 + cnt) over the
// chunk [k0, k1) of their chain. A tile split over several chunks writes partial sums (slot0 + kidx), added in chunk
// order by a finalize item.
struct TopItem {
    int32_t i0, cnt, k0, k1, slot0, kidx, ks, pad;
};

// An allocator whose default construction leaves its values unset, for buffers that their writers fill entirely
// (the solve panels: gigabytes that a pass of zeros would only page in twice).
template <class T>
struct UninitAlloc : std::allocator<T> {
    template <class U>
    struct rebind {
        using other = UninitAlloc<U>;
    };
    UninitAlloc() = default;
    template <class U>
    UninitAlloc(const UninitAlloc<U>&) noexcept {}
    template <class U>
    void construct(U* p) noexcept { ::new ((void*)p) U; }
    template <class U, class... A>
    void construct(U* p, A&&... a) { ::new ((void*)p) U(std::forward<A>(a)...); }
};

struct SchurHost {
    int n = 0;                          // kept nodes
    std::vector<int> perm;              // new -> old
    // supernodes in postorder (children first); columns of supernode s are new indices [c0[s], c0[s] + nc[s])
    std::vector<int> c0, nc, nr;        // first column, columns, off-diagonal rows
    std::vector<int64_t> roff;          // offset of s's off-diagonal rows in `rows`
    std::vector<int> rows;              // off-diagonal row indices (new numbering), ascending per supernode
    std::vector<int64_t> poff;          // offset of s's panel in `panel`: (nc + nr) x nc, column-major
    // the solve panel Q = [L_CC^-1 ; -L_RC L_CC^-1] (column-major, m x nc): the forward solve of a supernode is
    // [x_C ; u_R] = Q f_C + [0 ; f_R] and the backward x_C = Q^T [z_C ; x_R]: dense products, no serial chain.
    // Every entry is written: a host front's by its factorisation, a device front's by the device (its host copy
    // holds them only after fetch_panels).
    std::vector<double, UninitAlloc<double>> panel;
    std::vector<int> parent;            // assembly-tree parent (-1 at a root)
    std::vector<int> child_ptr, child;  // children per supernode (CSR, ascending)
    std::vector<int64_t> uoff;          // offset of s's update vector (nr entries) in the update buffer
    int64_t upd_total = 0;
    std::vector<int64_t> fpos;          // offset of s's front positions (nc + nr) in the contributor CSR
    std::vector<int64_t> cptr;          // contributors of each front position: update-buffer indices, children
    std::vector<int64_t> cidx;          //   in ascending order (a fixed summation order)
    // the same lists for the device, two per front position: {a, b} with -1 for none, or, for a position with more
    // than two contributors, {-2 - cptr[pos], count} (the list in cidx)
    std::vector<int32_t> ctab;
    int levels = 0;
    std::vector<int> fwd_ptr, bwd_ptr;  // work items per level (forward: leaves first; backward: roots first)
    std::vector<SnItem> fwd, bwd;       // the work items
    std::vector<int> fwd_cols, bwd_rows;      // per level: right-hand-side rows staged (LDS sizing)
    std::vector<int> ffin_ptr, bfin_ptr;      // per level: finalize items of its split tiles
    std::vector<SnItem> ffin, bfin;
    int64_t slots = 0;                  // partial-sum slots (SCHUR_TILE x 16 right-hand sides each)
    // The top of the assembly tree (the supernodes of depth < top_T) is solved as dense products with the explicit
    // inverse M of its block of L: forward z_T = M g, where g holds the right-hand side at the top columns plus the
    // update entries their non-top children send up (those carry the whole rest of the tree's forward solve), and
    // backward x_T = M^T z_T. Two or three launches each way replace 2 top_T level launches and their finalizes.
    // top_T is the deepest that keeps the top columns and M's entries within SCHUR_TOP_COLS and SCHUR_TOP_NNZ, and
    // 0 (no dense top) below SCHUR_TOP_MIN_T. M is made on the device after the factor (Schur::top_device) by
    // forward solves of unit vectors through the top supernodes' own level items (tfwd).
    int top_T = 0;
    std::vector<int> top_col;           // top columns (new numbering), ascending: top-column index i
    std::vector<int> top_lo;            // per i: first index of its supernode's top subtree (M's row i spans [lo, i])
    std::vector<int64_t> top_foff;      // per i: offset of row i's entries M[i][lo .. i] in the forward store
    std::vector<int> top_cptr, top_clen;        // per i: its chain (its supernode's columns, then its top
    std::vector<int> top_chain;                 //   ancestors', ascending; new numbering) at top_chain[top_cptr[i] ..]
    std::vector<int64_t> top_boff;      // per i: offset of column i's entries M[chain[e]][i] in the backward store
    int64_t top_fnnz = 0, top_bnnz = 0;
    std::vector<int64_t> top_gptr;      // per i: the update-buffer entries its non-top children add (ascending)
    std::vector<int64_t> top_gidx;
    std::vector<TopItem> top_fwd, top_ffin, top_bwd, top_bfin;
    int64_t top_slots = 0;              // partial-sum slots of the dense tiles (SCHUR_TOP_ROWS x 16 each)
    std::vector<int> tfwd_ptr, tfwd_cols, tffin_ptr;        // the top supernodes' forward level items (M's set-up)
    std::vector<SnItem> tfwd, tffin;
    int64_t nnz_L = 0;
    int max_cols = 0;
    // Fronts factorised on the device (schur_dense.hip): those of at least dev_min rows and every ancestor of one.
    // Their panels are made on the device; `build` leaves them out of the host's numeric and panel work and hands
    // over what their assembly needs: the original entries (front row, front column, value; duplicates summed) and
    // the update matrices of their host children (nr x nr, column-major).
    int dev_min = INT32_MAX;
    // nodes of a nested-dissection leaf box at most (one dense supernode at the bottom of the tree)
    int leaf = 32;
    // Sheets apart: the sheets are coupled only through branches, so the nodes those join across sheets (at most
    // SHEETS_APART_MAX of them) form the top separator and each sheet's grid is dissected on its own below it; the
    // separators span one sheet, not all. Else (or past the limit) the dissection cuts every sheet at once.
    bool sheets_apart = true;
    static constexpr int SHEETS_APART_MAX = 4096;
    std::vector<char> on_dev;
    std::vector<int> dev_order;                 // device fronts in postorder
    std::vector<int64_t> trip_off;              // per device front (index into dev_order): its original entries
    std::vector<int32_t> trip_r, trip_c;
    std::vector<double> trip_v;
    std::vector<std::vector<double>> upd_host;  // update matrices of host fronts whose parent is on the device
    bool host_panels = true;                    // false while device fronts' panels exist only on the device

    // factorise B given by its strictly-off-diagonal edges (a, b pairs, weight 1 each, repeated edges add) and the
    // diagonal (degree incl. edges to reference nodes); xyz: (sheet, i, j) per node
    void build(int n, const std::vector<int>& ea, const std::vector<int>& eb, const std::vector<double>& diag,
               const int32_t* xyz);
    // x = B^-1 b on the host (complex values interleaved), with the same panels as the device
    void solve_host(const double* b, double* x) const;
    // the dense top's index arrays and items (top_T chosen), from the supernodes' depths
    void build_top(const std::vector<int>& dep);
};

// Work items of the device solves. Forward: a tile of 64 panel rows whose columns are split 4 ways inside each wave
// (at most 32 columns), or, for a wider supernode, a tile of 16 rows with the columns split 16 ways over the
// workgroup. Backward: a tile of 16 columns, each column's rows split 16 ways inside a wave, or, for a supernode of
// 17 to 32 columns and at most 64 rows, one tile of 32 columns with the rows split 8 ways. Every lane takes at most
// SCHUR_PV panel values per item (all loaded at once), which sets the chunk sizes.
constexpr int SCHUR_PV = 8;
constexpr int SCHUR_TOP_COLS = 8192;    // top columns of the dense top at most
constexpr int64_t SCHUR_TOP_NNZ = int64_t(4) << 20;    // entries of its forward store at most
constexpr int SCHUR_TOP_MIN_T = 3;      // a dense top of fewer levels saves no launches
constexpr int SCHUR_TOP_CHUNK = 128;    // reduction length of one dense item
constexpr int SCHUR_TOP_ROWS = 64;      // rows (forward) or columns (backward) of one dense item
constexpr int SCHUR_NARROW = 32;       // forward: the widest supernode of 64-row tiles
constexpr int SCHUR_TILE = 16;         // rows (forward, wide) or columns (backward) of a split tile

// A captured Schur solve (the level launches of one batch size, input and output)
struct SchurGraph {
    int nr;
    const zd* b;
    zd* x;
    hipGraphExec_t exec;
};
constexpr size_t SCHUR_GRAPHS = 32;

struct Schur {
    SchurHost h;
    std::vector<SchurGraph> graphs;     // captured solves, replayed when the same (nr, b, x) comes again
    Schur() = default;
    Schur(const Schur&) = delete;
    Schur& operator=(const Schur&) = delete;
    ~Schur();
    DevBuf<int> perm, rows, ctab;
    DevBuf<int64_t> cidx;
    DevBuf<SnItem> fwd, bwd, ffin, bfin;
    DevBuf<zd> part;                    // partial sums of split tiles
    DevBuf<double> panel;
    DevBuf<zd> ubuf;                    // update vectors of one forward solve (cap right-hand sides)
    DevBuf<zd> zp, xp;                  // forward and backward solutions in the new numbering
    // the dense top: its index arrays, M's forward (row) and backward (column) stores, its items, partial sums, and
    // g for the batch
    DevBuf<int> top_col, top_lo, top_cptr, top_clen, top_chain;
    DevBuf<int64_t> top_foff, top_boff, top_gptr, top_gidx;
    DevBuf<double> top_Mf, top_Mb;
    DevBuf<TopItem> top_fwd, top_ffin, top_bwd, top_bfin;
    DevBuf<zd> top_part, gbuf;
    DevBuf<SnItem> tfwd, tffin;
    int cap = 0;                        // right-hand sides ubuf / zp / xp hold (indexed with the batch's stride)

    void upload(hipStream_t s);
    // ubuf / zp / xp for nr right-hand sides: grown (dropping the captured graphs, whose pointers move) when nr > cap
    void ensure(int nr);
    // back to nr right-hand sides when they hold more (set-up work that took a larger batch)
    void shrink(int nr);
    // the bytes ubuf / zp / xp would hold after ensure(nr)
    int64_t bytes_after(int nr) const;
    // the panels of the device fronts (schur_dense.hip), after upload: assembly, blocked partial Cholesky and solve
    // panels, straight into `panel`
    void factor_device(hipStream_t s);
    // the device fronts' panels copied into h.panel (for solve_host)
    void fetch_panels(hipStream_t s);
    // M of the dense top (after the panels are final): forward solves of unit vectors through tfwd, 16 at a time
    void top_device(hipStream_t s);
    int64_t device_bytes() const;
};

// x = B^-1 b on the device for nr <= 16 right-hand sides, interleaved: entry (node, p) at node * nr + p, in the old
// numbering (b and x distinct)
void schur_solve_device(Schur& S, const zd* b, zd* x, int nr, hipStream_t s, bool graph);
// The two passes of that solve apart (the same kernels, so the same bits): the forward pass leaves M^-1 P b in zp
// (new numbering, interleaved; M the factor, B = P^T M M^T P), the backward pass writes x = P^T M^-T zp. Between them
// a caller may change zp (the Woodbury correction on the last supernode's rows).
void schur_forward_device(Schur& S, const zd* b, int nr, hipStream_t s);
void schur_backward_device(Schur& S, zd* x, int nr, hipStream_t s);
// The forward pass for right-hand sides that vanish outside the last supernode's columns (the root Woodbury form's
// unit columns): every other supernode's front is zero, so its update entries are cleared instead of computed and
// only the last supernode's level runs (or the dense top, when it holds it). zp then holds, on the last supernode's
// rows, the same bits as after schur_forward_device.
void schur_forward_last_device(Schur& S, const zd* b, int nr, hipStream_t s);

}  // namespace gt
