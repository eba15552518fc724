// Host-side objects behind the C ABI: the context (device + stream), the resident stack and the per-frequency
// operator, and the device work each owns.
#pragma once

#include <atomic>

#include <algorithm>
#include <map>
#include <memory>
#include <mutex>
#include <vector>

#include "gt_common.hpp"
#include "lp.hpp"
#include "schur.hpp"

namespace gt {
struct GmresWork;
}

struct gt_ctx_s {
    int device = 0;
    hipStream_t stream = nullptr;
    size_t lds_max = 65536;
    // every call that works on this context's stream and scratch holds it (calls from several host threads run one
    // at a time)
    std::mutex mu;
    int schur_dev_min = 256;                  // fronts of at least this many rows are factorised on the device
    bool graphs = false;                      // Schur solves captured in hipGraphs and replayed (gt_ctx_graphs)
    bool gs_delayed = true;                   // GMRES orthogonalisation: delayed CGS2, else CGS + DGKS re-pass
    int schur_leaf = 32;                      // nested-dissection leaf size of the factorisations (gt_ctx_schur_leaf)
    bool schur_sheets_apart = true;           // each sheet dissected on its own (gt_ctx_schur_ordering)
    // Woodbury form of the stacks factored from here on (gt_ctx_woodbury): GT_WOODBURY_AUTO keeps W = B^-1 U resident
    // while it has at most 256 columns and wb_stored_max bytes, else applies B^-1 U t by a second Schur solve
    int woodbury = 0;
    int64_t wb_stored_max = int64_t(16) << 30;
    std::unique_ptr<gt::GmresWork> gmres;     // GMRES device memory kept between calls (gmres.cpp)
    // L_p launch timing (gt_ctx_lp_timing): an event pair on the stream around every L_p launch while enabled
    bool lp_timing = false;
    std::vector<hipEvent_t> lp_events;
    size_t lp_used = 0;                       // pairs recorded
    int64_t lp_rhs = 0;                       // right-hand sides those launches carried
    gt_ctx_s();
    ~gt_ctx_s();
};

namespace gt {

struct OrientHost {
    bool present = false;
    int n0 = 0, n1 = 0, P0 = 0, P1 = 0, log0 = 0, log1 = 0, TR = 1, lo = 0, cnt = 0;
    size_t t_off = 0;          // offset of this orientation's L_p intermediates in T1 / T2
    DevBuf<int> rowptr, rowcol;
    // the kernel spectra, shared by the stacks made like this one (gt_stack_create_like: same grid and sheet gaps)
    std::shared_ptr<DevBuf<double>> K = std::make_shared<DevBuf<double>>();
    std::shared_ptr<DevBuf<float>> Kf = std::make_shared<DevBuf<float>>();
    bool k_shared = false;     // K and Kf came from another stack (not counted in this one's device bytes)
};

// Scratch of one batched product / preconditioner application (grown to the largest batch used).
struct Work {
    int cap = 0;              // right-hand sides the buffers hold
    DevBuf<zd> T1, T2;        // L_p intermediates, t_elems per right-hand side
    DevBuf<zd> s, z;          // Schur right-hand sides and solutions, interleaved (n_keep x batch)
    DevBuf<zd> t;             // Woodbury: Q U^T z per right-hand side (t_cols each)
    DevBuf<zd> z2;            // implicit Woodbury form: B^-1 U t, interleaved like z
    DevBuf<zd> u;             // root form: Y_r^T w per right-hand side (t_cols each)
    int t_cols = 0;
    void ensure(const gt_stack_s* st, int nr, int kt = 0);
    // back to nr right-hand sides when the buffers hold more (set-up work that took a larger batch)
    void shrink(const gt_stack_s* st, int nr);
    // back to what a new stack holds (one right-hand side, its own branches' Woodbury width)
    void reset(const gt_stack_s* st);
    int64_t bytes() const {
        return (int64_t)(T1.bytes() + T2.bytes() + s.bytes() + z.bytes() + t.bytes() + z2.bytes() +
                         u.bytes());
    }
    // the bytes it would hold after ensure(st, nr, kt)
    int64_t bytes_after(const gt_stack_s* st, int nr, int kt) const;
};

}  // namespace gt

namespace gt {

// What a stack shares with the what-if stacks derived from it (peec.py:1871-1909 `with_branches`: "shares grids and
// kernels"): the strip layout, the kernel spectra and twiddles, and the Schur factor of the Laplacian of its own
// branches.
struct StackCore {
    OrientHost o[2];
    std::map<int, DevBuf<zd>> tw;
    std::map<int, DevBuf<zf>> twf;
    DevBuf<int> node_sheet;
    DevBuf<unsigned char> strip_sheet;        // ns: sheet of each strip
    std::vector<unsigned char> h_strip_sheet;
    LpShape shape;
    size_t t_elems = 0;                       // complex elements of the L_p intermediates (both orientations)
    int kG = 0;                               // distinct kernels per pixel (face gaps) and the sheets' map to them
    std::vector<int> kgmap;
    DevBuf<int> kgmap_dev;
    std::unique_ptr<Schur> schur;
    // strips narrower than the pitch (gt_stack_set_narrow, peec.py:1732-1843): per strip its resistance factor less
    // one (1 / r_width - 1) and its narrow self term (H), and the side-by-side neighbours' mutual corrections (H) as a
    // CSR over the strips; absent while every strip is a pitch wide
    bool narrow = false;
    DevBuf<double> nw_fac, nw_dl, nw_val;
    DevBuf<int> nw_ptr, nw_col;
};

}  // namespace gt

struct gt_stack_s {
    std::shared_ptr<gt::StackCore> core;
    // the core's members under their own names
    gt::OrientHost (&o)[2];
    std::map<int, gt::DevBuf<gt::zd>>& tw;
    std::map<int, gt::DevBuf<gt::zf>>& twf;
    gt::DevBuf<int>& node_sheet;
    gt::DevBuf<unsigned char>& strip_sheet;
    std::vector<unsigned char>& h_strip_sheet;
    gt::LpShape& shape;
    size_t& t_elems;
    int& kG;
    std::vector<int>& kgmap;
    gt::DevBuf<int>& kgmap_dev;
    std::unique_ptr<gt::Schur>& schur;

    gt_ctx* ctx = nullptr;
    int S = 0, nx = 0, ny = 0, nxb = 0, nyb = 0, nbr = 0, ns = 0, nb = 0, nk = 0;
    gt::DevBuf<int> a_keep, b_keep, adj_ptr, adj;
    std::vector<int> h_a_keep, h_b_keep;     // host copies (Woodbury columns, the factor's pattern)
    gt::Work work;
    int sessions = 0;                         // sessions open on this stack (its scratch is theirs while any is)
    gt::DevBuf<double> W;                     // Woodbury: B^-1 U, k columns of n_keep (column-major)
    bool wb_implicit = false;                 // W not stored: B^-1 (U t) by a second Schur solve per product
    gt::DevBuf<double> G;                     // implicit form: U^T B^-1 U (k x k, row-major)
    // Root form of the implicit one, when every branch end lies in the factor's last supernode: with Y = M^-1 P U
    // (nonzero on that supernode's yr_n rows only) the correction is w_r -= Y_r Q Y_r^T w_r between the solve's
    // passes, and G = Y_r^T Y_r
    bool wb_root = false;
    // Node columns (gt_schur_factor_nodes): the k Woodbury columns are unit columns at kept nodes, not branch
    // incidence columns (sheets of different strip admittance); always in the root form
    bool wb_nodes = false;
    gt::DevBuf<double> Yr;                    // k columns of yr_n (column-major)
    int yr_c0 = 0, yr_n = 0;
    int k = 0;
    int k_root = 0;                           // the columns of the first stack of its lineage (copied first into W)
    bool derived = false;                     // a what-if stack: its core belongs to the stack it came from
    uint64_t id = 0;                          // unique in the process
    std::vector<uint64_t> lineage;            // the ids of the stacks it was derived from (parent first)

    explicit gt_stack_s(std::shared_ptr<gt::StackCore> c = std::make_shared<gt::StackCore>())
        : core(std::move(c)), o(core->o), tw(core->tw), twf(core->twf), node_sheet(core->node_sheet),
          strip_sheet(core->strip_sheet), h_strip_sheet(core->h_strip_sheet), shape(core->shape),
          t_elems(core->t_elems), kG(core->kG), kgmap(core->kgmap), kgmap_dev(core->kgmap_dev), schur(core->schur), id(next_id()) {}
    gt_stack_s(const gt_stack_s&) = delete;
    gt_stack_s& operator=(const gt_stack_s&) = delete;

    int64_t device_bytes() const;
    gt::LpArgs<double> lp_args(int mode) const;
    static uint64_t next_id() {
        static std::atomic<uint64_t> counter{0};
        return ++counter;
    }
};

struct gt_op_s {
    gt_stack* st = nullptr;
    double w = 0;
    gt::DevBuf<gt::zd> r_sheet, z_tail, dinv_sheet, dinv_tail, q;
    gt::DevBuf<gt::zf> r_sheet_f, z_tail_f;   // the product's impedances for single-precision products
    gt::zd s_scale;
    gt::DevBuf<gt::zd> s_in, s_out;           // per-sheet scales of a node-column stack (empty: s_scale)
    bool per_sheet = false;
    int k = 0;
};

namespace gt {
// device work of a batch of operators on one host stack (ops.hip): y_p = K_p x_p, z_p = M_p^-1 r_p for p < nr <= 16
// (T: the storage of the vectors, double or float; the preconditioner's arithmetic and Schur solve are double).
// Each operator belongs to st or to a what-if stack derived from it (check_batch).
void check_batch(const gt_stack* st, int nr, gt_op* const* ops);
template <class T>
void ops_matvec(gt_stack* st, int nr, gt_op* const* ops, const cpx<T>* const* x, cpx<T>* const* y);
// y_p += the narrow strips' corrections of x_p (ops.hip): with ops, the operator's (z_sheet (1 / r_width - 1) x +
// j w (dL x + M_near x)); without, the L_p product's (dL x + M_near x). Nothing when the stack has none.
template <class T>
void narrow_apply(const gt_stack* st, int nr, gt_op* const* ops, const cpx<T>* const* x, cpx<T>* const* y);
template <class T>
void ops_precond(gt_stack* st, int nr, gt_op* const* ops, const cpx<T>* const* r, cpx<T>* const* z);
void schur_apply(gt_stack* st, const zd* b, zd* x, int nr);   // B^-1 b, nr interleaved right-hand sides
// The no-allocation scratch (noalloc.hip): looked up once per device at context creation; its address on a device
void noalloc_scratch_setup(int device);
void* noalloc_scratch(int device);
// (Kx, Ky) at every node of every sheet, S x nx x ny complex each, from the strip currents of x (post.hip)
void node_field_device(gt_stack* st, const zd* x, double d, zd* kx, zd* ky);
void loop_terms_device(gt_stack* st, const zd* x, const double* psi, const double* bm, int n_inj, const int* inj_idx,
                       const double* inj_amps, double d, const double* sheet_z, double* out);
void via_coupling_device(gt_ctx* ctx, int nv, const double* vias, int nseg, const gt_segment* segs, double mu,
                         double* out);
// frees the context's kept GMRES workspace (gmres.cpp)
void ctx_trim_gmres(gt_ctx* ctx);
// copper rasterisation (raster.hip)
void raster_label(gt_ctx* ctx, int32_t ng, const int64_t* eptr, const double* edges, const int32_t* win, int32_t nu,
                  const double* xs, int32_t nv, const double* ys, int32_t* lab);
void raster_fine(gt_ctx* ctx, int32_t ng, const int64_t* eptr, const double* edges, const int32_t* win, int32_t nx,
                 int32_t ny, int32_t s, const double* xf, const double* yf, int32_t* part, double* fill,
                 double* cross_x, double* off_x, double* cross_y, double* off_y, int32_t* fine);
void raster_state(gt_ctx* ctx, int64_t ne, const double* edges, const int32_t* win, int32_t nu, const double* xs,
                  int32_t nv, const double* ys, int8_t* state);
void raster_cover(gt_ctx* ctx, int64_t ne, const double* edges, double u0, double du, int32_t n_lines, double v0,
                  double d, int32_t n_cells, double* out);
void raster_points(gt_ctx* ctx, int64_t ne, const double* edges, int64_t np, const double* pts, int8_t* inside,
                   int8_t* onb, double* nearest);
// psi of every strip (ns values at psi, device memory) from the trace's segments (psi.hip)
void strip_coupling_device(gt_stack* st, double d, double mu, int nseg, const gt_segment* segs, const double* sheet_z,
                           const double* sheet_t, double* psi);
// What the device kernel generation needs beyond the stack's own sizes (kernel_gen.hip)
struct KernelGen {
    double d = 0, mu = 0;       // pitch (m), mu0 / (4 pi)
    const double* dz = nullptr;     // S x S face gaps
    const double* nearv = nullptr;  // S x S x 5 x 5 near-field mutuals [s][t][axial][lateral]
    int G = 0;                      // distinct kernels: one per face gap, made from the sheet pair (rep_s, rep_t)
    int rep_s[LP_MAX_SHEETS * LP_MAX_SHEETS] = {}, rep_t[LP_MAX_SHEETS * LP_MAX_SHEETS] = {};
};
// The kernel quadrant of orientation oi (its FFT sizes and twiddles set) made on the device, FP64 and FP32 copies
void kernel_quadrant_device(gt_stack* st, int oi, const KernelGen& kg);
// Woodbury columns on the device: the base's W copied (none without a base), then B^-1 U' for the added branches.
// In the implicit form (st->wb_implicit) only G = U^T B^-1 U is kept: the base's block copied, the added branches'
// columns from the same solves, their rows by symmetry.
void woodbury_extend(gt_stack* st, const gt_stack* base, int n_extra, const int* a_extra, const int* b_extra);
// G = U^T W (k x k row-major) of the stack's branches, from its device W (or its device G), into host memory
void woodbury_gram_device(const gt_stack* st, double* G_host);
constexpr int WB_STORED_MAX_K = 256;     // branches of one stack with W resident (one workgroup's LDS for U^T z)
constexpr int WB_IMPLICIT_MAX_K = 8192;  // branches of one stack in the implicit form
constexpr int WB_IMPLICIT_SMALL_K = 2048;   // ... up to which its product stages U^T z in 32 KB of LDS (else 128 KB)
constexpr int WB_NODES_MAX_K = 2048;     // branch end nodes of a stack of sheets of different admittance (root form)
}  // namespace gt
