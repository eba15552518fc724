/*
 * gtpeec.h - C ABI of libgtpeec.so, the MI355X (gfx950) plane-current PEEC solve.
 *
 * Plain pointers and sizes only; no exceptions cross this boundary. Every call returns GT_OK or an error code, and
 * gt_last_error() gives the calling thread's last message. Complex numbers are interleaved (re, im) doubles.
 *
 * Object model (the lifetimes of the reference's device seam, src/groundtruth/solver/gpu.py:617-666):
 *   gt_ctx    one HIP device and the stream every call on it is ordered on (gpu.py:247 `use`, one process per GPU).
 *   gt_stack  a board resident on the device: strip layout, incidence, kernel spectra, Schur factor
 *             (peec.py:1555 PlaneStack; uploaded once, as gpu.py:617-645 `_stack_data`).
 *   gt_op     one frequency's data on a stack (peec.py:1505 _Operator; gpu.py:504-520 `_op`).
 *
 * Unknowns (peec.py:1596-1656): x = [Ib (nb) ; V_keep (n_keep)], nb = nxb + nyb + nbr, strips sorted by
 * (sheet, i, j) per orientation (x strips first), lumped branches after the strips.
 */
#ifndef GTPEEC_H
#define GTPEEC_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GT_ABI_VERSION 1

#define GT_OK 0
#define GT_ERR_ARG 1          /* a bad argument (the caller's bug): ValueError                                   */
#define GT_ERR_OOM 2          /* device or host allocation failed: MemoryError (peec.py:2428-2429)               */
#define GT_ERR_DEVICE 3       /* a HIP call failed (no device, launch failure)                                    */
#define GT_ERR_NOCONV 4       /* GMRES did not converge: SolverError (peec.py:2599-2600)                          */
#define GT_ERR_NUMERIC 5      /* the Schur factorisation met a zero or non-finite pivot: SolverError             */
#define GT_ERR_UNSUPPORTED 6  /* a shape this library has no kernel for (FFT size, LDS budget)                      */

typedef struct gt_ctx_s gt_ctx;
typedef struct gt_stack_s gt_stack;
typedef struct gt_op_s gt_op;

/* The calling thread's last error message ("" after success). */
const char* gt_last_error(void);
int gt_abi_version(void);

/* ---------------------------------------------------------------- context
 * Replaces the device selection of gpu.py:247 `use` / :120 `probe`: binds HIP device `device` and creates the
 * stream all work of this context is ordered on. */
int gt_ctx_create(int device, gt_ctx** out);
void gt_ctx_destroy(gt_ctx* ctx);
int gt_ctx_sync(gt_ctx* ctx);
/* Frees the device memory the context keeps between solves (its GMRES workspace, grown to the largest batch it has
 * served; the next solve makes it again). No counterpart in the reference; PlaneStack calls it, with torch's cache
 * emptied, before it refuses a model whose footprint exceeds 0.8 of the free memory (gpu.py:483-488). */
int gt_ctx_trim(gt_ctx* ctx);
/* The context's hipStream_t (for callers that order their own device work with it). */
void* gt_ctx_stream(gt_ctx* ctx);
/* Measurement (no counterpart in the reference: per-kernel HBM figures against the device's peak). While enabled,
 * every L_p launch of the system products on this context is bracketed by a HIP event pair on its stream;
 * enabling (or disabling) clears the record. gt_ctx_lp_time waits for the stream and gives the launches recorded,
 * the right-hand sides they carried and their summed duration in ms. */
int gt_ctx_lp_timing(gt_ctx* ctx, int enable);
/* The factorisations made on this context (gt_schur_factor) take every front of at least m_min rows, and its
 * ancestors, on the device (assembly, blocked partial Cholesky, solve panels); smaller fronts stay on the host.
 * Default 256; INT32_MAX keeps the whole factorisation on the host. (No counterpart in the reference, which
 * factorises with SuperLU on the host, peec.py:2379-2424.) */
int gt_ctx_schur_device_fronts(gt_ctx* ctx, int32_t m_min);
/* The nested-dissection leaf size of the factorisations made on this context: boxes of at most `nodes` nodes
 * (4-512; default 32) are one dense supernode each. Larger leaves give a shallower tree (fewer solve levels) and
 * denser leaf panels. (No counterpart in the reference, whose SuperLU orders by COLAMD.) */
int gt_ctx_schur_leaf(gt_ctx* ctx, int32_t nodes);
/* The nested-dissection ordering of the factorisations made on this context. GT_ND_SHEETS_APART (the default): the
 * kept nodes that branches join across sheets (at most 4096) form the top separator and each sheet's grid is
 * dissected on its own below it, so every other separator spans one sheet; past 4096 such nodes, or with
 * GT_ND_JOINT, every separator cuts all sheets at once. (No counterpart in the reference, whose SuperLU orders by
 * COLAMD.) */
#define GT_ND_JOINT 0
#define GT_ND_SHEETS_APART 1
int gt_ctx_schur_ordering(gt_ctx* ctx, int32_t ordering);
/* The Woodbury form of the stacks factorised on this context from now on: how the preconditioner's branch
 * correction is applied (the reference factorises a board's branches into S and keeps a what-if's S^-1 U dense on
 * the host, peec.py:2364-2377). GT_WOODBURY_STORED keeps W = B^-1 U resident (n_keep x k doubles, at most 256 branches); GT_WOODBURY_IMPLICIT keeps
 * only G = U^T B^-1 U (k x k, at most 8192 branches) and applies B^-1 U t by a second Schur solve in each
 * preconditioner product. GT_WOODBURY_AUTO (the default) stores W while k <= 256 and W takes at most
 * stored_max_bytes (default 16 GiB). A derived stack takes its base's form. (No counterpart in the reference.) */
#define GT_WOODBURY_AUTO 0
#define GT_WOODBURY_STORED 1
#define GT_WOODBURY_IMPLICIT 2
int gt_ctx_woodbury(gt_ctx* ctx, int32_t form, int64_t stored_max_bytes);
/* A stack in the implicit form whose branch ends all lie in the factor's last supernode (the sheet-apart ordering
 * puts every node a branch joins across sheets there) takes its root form, reported by gt_stack_woodbury as
 * GT_WOODBURY_ROOT: with Y = M^-1 P U (M the factor, B = P^T M M^T P), nonzero on that supernode's rows only, the
 * correction is w -= Y Q Y^T w between the Schur solve's forward and backward passes (one solve per product, no
 * second one) and G = Y^T Y. A what-if derived from it takes the second solve for its own right-hand sides. */
#define GT_WOODBURY_ROOT 3
/* Schur solves on this context captured in hipGraphs: the level launches of each batch size are recorded once and
 * replayed with one submission; the same kernels and arguments, so the same bits. Off by default: on the MI355X a
 * replay takes as long as the launches it replaces (the levels' time is their own latency on the device, not the
 * host's submissions). (No counterpart in the reference.) */
int gt_ctx_graphs(gt_ctx* ctx, int enable);
/* The GMRES orthogonalisation of the solves on this context. GT_GS_CGS_DGKS: classical Gram-Schmidt with the DGKS
 * re-pass where |w| falls below 0.7 of itself, as peec.py:3093-3224 `_gmres` (three reads of the basis per
 * Arnoldi step when the re-pass is taken). GT_GS_DELAYED_CGS2 (the default): the same second pass, always taken and
 * one step late, fused with the next step's projection (two reads per step; the low-synchronisation GMRES of
 * Swirydowicz et al., Numer. Linear Algebra Appl. 28 (2021) e2343). Both give the same Krylov space; iteration
 * counts and solutions agree to rounding. (No counterpart in the reference beyond its CGS + DGKS.) */
#define GT_GS_CGS_DGKS 0
#define GT_GS_DELAYED_CGS2 1
int gt_ctx_orthogonalization(gt_ctx* ctx, int mode);
int gt_ctx_lp_time(gt_ctx* ctx, int64_t* launches, int64_t* rhs, double* ms);

/* ---------------------------------------------------------------- stack
 * Replaces gpu.py:617-666 `_stack_data` / `_orients` (kernels, strip indices and incidence uploaded once per stack
 * and device) for PlaneStack (peec.py:1555-1988). The kernel spectra are the real, even spectra of
 * peec.py:3276-3371 `_kernel_pair` / `_kernel_spectrum` for the FFT sizes given here (any size >= 2n-1 embeds the
 * linear convolution), stored as their [0, P1/2] x [0, P0/2] quadrant, pixel-major over sheet pairs: element
 * ((qj * (P0/2 + 1) + qi) * S + t) * S + s, already scaled by 1 / (P0 * P1). */
typedef struct {
    int32_t n_sheets;         /* S: current sheets (one per layer)                                               */
    int32_t nx, ny;           /* nodes per layer                                                                  */
    int32_t nxb, nyb;         /* x and y strips                                                                   */
    int32_t nbr;              /* lumped branches (vias) after the strips                                          */
    int32_t n_keep;           /* kept node potentials (one reference node per connected piece dropped)           */
    const int32_t* x_sij;     /* nxb x 3 (sheet, i, j) of the x strips, sorted                                    */
    const int32_t* y_sij;     /* nyb x 3                                                                          */
    const int32_t* a_keep;    /* nb: kept index of each branch's node a (+1 end), -1 for a reference node        */
    const int32_t* b_keep;    /* nb: kept index of node b (-1 end)                                                */
    const int32_t* node_sheet;/* n_keep: sheet of each kept node                                                  */
    int32_t px[2];            /* FFT size of the x strips (P0 along i, P1 along j), powers of two                */
    int32_t py[2];            /* FFT size of the y strips                                                         */
    const double* kx;         /* x kernel quadrant, (px[1]/2+1)(px[0]/2+1) S S doubles                            */
    const double* ky;         /* y kernel quadrant                                                                */
} gt_stack_desc;

int gt_stack_create(gt_ctx* ctx, const gt_stack_desc* desc, gt_stack** out);

/* The same with the kernel spectra made on the device (peec.py:3276-3371 `_kernel_pair`, kernel.py:268-294
 * `strip_kernel`): desc->kx / desc->ky are ignored. The far field is the filament closed form at every offset; the
 * near field (offsets below 5 cells along and across) comes from the caller, the reference's graded quadrature of
 * the tape mutual (kernel.py:222-265); the spectrum is two cosine sums over the even kernel, made into the same
 * quadrant layout. */
typedef struct {
    double d;                 /* grid pitch (m)                                                                   */
    double mu;                /* mu0 / (4 pi) as the caller computes it                                          */
    const double* dz;         /* S x S: face gap between sheets s and t (m), symmetric                          */
    const double* near_field; /* S x S x 5 x 5: [s][t][axial][lateral] tape mutuals at offsets of 0..4 cells (H) */
} gt_kernel_desc;
int gt_stack_create_gen(gt_ctx* ctx, const gt_stack_desc* desc, const gt_kernel_desc* kernels, gt_stack** out);
/* The resident kernel quadrant of orientation 0 (x strips) or 1 (y strips), in the layout of gt_stack_desc. */
int gt_stack_kernel(const gt_stack* stack, int32_t orient, double* out);
void gt_stack_destroy(gt_stack* stack);
/* Device bytes the stack holds (the footprint gpu.py:593-605 `device_bytes` estimates); for a what-if stack
 * (gt_stack_derive), only what it adds to its base's. */
int64_t gt_stack_device_bytes(const gt_stack* stack);

/* L_p I (peec.py:2115-2135 Lp_matvec) on device vectors of nxb + nyb complex values: scatter, row FFT, column
 * FFT + kernel multiply-add + inverse column FFT, inverse row FFT, gather. */
int gt_lp_matvec(gt_stack* stack, const void* I_dev, void* out_dev);
/* The same for nr <= 16 vectors in one launch (peec.py:2115-2135 applied to each): the kernel spectra are read once
 * for the batch. */
int gt_lp_matvec_batch(gt_stack* stack, int32_t nr, const void* const* I_dev, void* const* out_dev);
/* gt_lp_matvec_batch that allocates nothing. No counterpart in the reference: it is for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc). The
 * same kernels and launch shapes, so each output equals gt_lp_matvec_batch's bit for bit; the intermediates live in
 * a fixed device array of GT_LP_NOALLOC_SCRATCH_BYTES per GPU (held until the process ends), not in the stack's
 * workspace. A batch needs 2 * 16 B * nr * (the stack's per-product intermediate elements) of it: a batch above that
 * returns GT_ERR_ARG with the sizes.
 * Calls that overlap on one GPU (from other contexts or threads, before an earlier call's kernels finish) share the
 * array and race: order them on one stream. */
#define GT_LP_NOALLOC_SCRATCH_BYTES (64u << 20)
int gt_lp_matvec_batch_noalloc(gt_stack* stack, int32_t nr, const void* const* I_dev, void* const* out_dev);

/* The complex sheet-current density (Kx, Ky), in A/m per A of signal, at every grid node of every sheet from the
 * strip currents of a solution vector on the device (x_dev: nb + n_keep complex values; kx_dev, ky_dev: S x nx x ny
 * complex each, row-major, written): at a node, the mean of the touching strips along each axis that carry current,
 * over the pitch d (m). Replaces peec.py:1151-1199 `Solution.node_field` (copper=False) for a served map, without
 * copying the solution to the host; the values equal the host's to rounding. */
int gt_node_field(gt_stack* stack, const void* x_dev, double d, void* kx_dev, void* ky_dev);

/* psi of every strip (peec.py:2609-2656 `strip_coupling`): the integral of the trace's vector potential along each
 * strip per amp of loop current, the strips' right-hand side (rhs_top = -j w psi). One trace segment from a to b (m),
 * `width` wide (m), carrying `current` amps per amp of loop current. psi_dev: ns = nxb + nyb doubles (device).
 * sheet_z, sheet_t: each sheet's height and copper thickness (m); d: the pitch (m); mu: mu0 / (4 pi). */
typedef struct {
    double a[3], b[3];
    double width;
    double current;
} gt_segment;
int gt_strip_coupling(gt_stack* stack, double d, double mu, int32_t nseg, const gt_segment* segs,
                      const double* sheet_z, const double* sheet_t, void* psi_dev);

/* The loop's sums over a solution on the device, for Z_loop and the moment without copying the solution to the host
 * (peec.py:2787-2811 `RouteExcitation.solve`'s Z = R_trace + sum a_k V[n_k] + j w (L_trace + psi . Ib), and
 * peec.py:2668-2700 `PlaneStack.moment`'s m = 1/2 sum r x I dl over the strips and branches). x_dev: nb + n_keep
 * complex values; psi_dev: nb doubles (the strips' then the branches' coupling); bm_dev: per branch 1/2 r_a x r_b
 * (3 doubles, m), or NULL to leave the branches out of m; inj_idx / inj_amps (host): the injection terms, each an index
 * into x_dev (a kept node's potential, nb <= idx < nb + n_keep) and its amps; d: the pitch (m); sheet_z (host): each
 * sheet's height (m). out (host, 10 doubles, written): psi . Ib, the strips' and branches' m (3 complex values, the
 * trace's own part not included) and sum a V, each as (re, im). */
int gt_loop_terms(gt_stack* stack, const void* x_dev, const void* psi_dev, const void* bm_dev, int32_t n_inj,
                  const int32_t* inj_idx, const double* inj_amps, double d, const double* sheet_z, double* out);

/* psi of via branches (peec.py:2658-2666 `branch_coupling`): each via's mutual with the trace's vertical segments
 * (those with no run in the plane), times the segments' current shares. vias (host): per via its ends a, b (m) and
 * its radius (m), 7 doubles; mu: mu0 / (4 pi); out (host, nv doubles, written). */
int gt_via_coupling(gt_ctx* ctx, int32_t nv, const double* vias, int32_t nseg, const gt_segment* segs, double mu,
                    double* out);

/* ---------------------------------------------------------------- copper rasterisation
 * Replaces the shapely point and line queries of peec.py:329-983 `build_grid` and its helpers over a board's area.
 * Copper is given as pieces, each a run of ring edges (ax, ay, bx, by per edge, mm; edge_ptr[k] .. edge_ptr[k + 1]
 * for piece k) whose even-odd region is the piece's copper. windows: per piece the lattice columns [p0, p1) and
 * samples [q0, q1) that can hold it (4 int32). Point tests are exact: a point on an edge is on the boundary whatever
 * the edge's slope, as shapely's predicates decide it. All arrays are host memory. */

/* Lattice labels (peec.py:379-385 `contains_xy` of each net's copper at the nodes and strip midpoints): labels[p * nv
 * + q] (in/out) takes k for the first piece k whose interior holds the point (xs[p], ys[q]), where it is still < 0.
 * xs increasing. A lattice column crossing more than 4096 edges of one piece is refused. */
int gt_raster_label(gt_ctx* ctx, int32_t n_pieces, const int64_t* edge_ptr, const double* edges,
                    const int32_t* windows, int32_t nu, const double* xs, int32_t nv, const double* ys,
                    int32_t* labels);

/* The sub-grid (peec.py:774-849 `_coarse_from_fine`): the fine lattice xf (nx sub) x yf (ny sub) labelled by piece
 * as above, each coarse cell's largest same-piece group of fine samples (4-neighbours inside the cell; ties to the
 * group met first in row-major order) kept when it has at least `sub` samples: part (nx x ny, its piece or -1), fill
 * (its share of the cell); and per cell border the share of its `sub` fine strips that join kept groups of one piece
 * (cross_x (nx-1) x ny, cross_y nx x (ny-1)) and their mean offset from the border's middle (off_x, off_y), each a
 * fraction of the pitch. fine_labels (nx sub x ny sub) may be NULL. sub: 2 .. 8. */
int gt_raster_fine(gt_ctx* ctx, int32_t n_pieces, const int64_t* edge_ptr, const double* edges,
                   const int32_t* windows, int32_t nx, int32_t ny, int32_t sub, const double* xf, const double* yf,
                   int32_t* part, double* fill, double* cross_x, double* off_x, double* cross_y, double* off_y,
                   int32_t* fine_labels);

/* Lattice points against one region (peec.py:472 `intersects_xy` of a net's nodes): state[p * nv + q] = 1 inside,
 * 2 on an edge, 0 outside, for the points of the window (4 int32: columns, then samples); 0 elsewhere. */
int gt_raster_state(gt_ctx* ctx, int64_t n_edges, const double* edges, const int32_t* window, int32_t nu,
                    const double* xs, int32_t nv, const double* ys, int8_t* state);

/* Line covers (peec.py:533-583 `_line_cover`): on each line u = u0 + m du (m < n_lines) the length of each interval
 * [v0 + (c - 1/2) d, v0 + (c + 1/2) d] (c < n_cells) inside the even-odd region of the edges (ua, va, ub, vb per
 * edge); out: n_lines x n_cells (written). A line with an odd number of crossings covers nothing. */
int gt_raster_cover(gt_ctx* ctx, int64_t n_edges, const double* edges, double u0, double du, int32_t n_lines,
                    double v0, double d, int32_t n_cells, double* out);

/* Points against one region's edges (peec.py:400 `covers`, :472-474 `intersects_xy` and `shortest_line`): inside
 * (even-odd, the boundary aside), on_edge, and the nearest point of the edges (x, y per point; the first of equal
 * distances; edges binned when there are many). Any output may be NULL. */
int gt_raster_points(gt_ctx* ctx, int64_t n_edges, const double* edges, int64_t n_points, const double* points,
                     int8_t* inside, int8_t* on_edge, double* nearest);

/* Strips narrower than the pitch (peec.py:1732-1843 `PlaneStack._widths`, applied at :2129-2132 and :2159): per strip
 * r_factor = 1 / r_width (its resistance and internal impedance are its sheet's / r_width; NULL: all 1) and dL, the
 * partial self inductance its narrow copper adds to the full strip's (H; NULL: none), and M_near, the side-by-side
 * neighbours' mutual corrections (H), a CSR over the strips (rowptr ns + 1, col, val; nnz 0: none). The operator
 * then adds z_sheet (1 / r_width - 1) I + j w (dL I + M_near I) to the strip rows and the L_p product dL I + M_near I;
 * the preconditioner keeps each sheet's full-strip admittance. Shared with the stack's what-if stacks. */
int gt_stack_set_narrow(gt_stack* stack, const double* r_factor, const double* dL, int64_t nnz, const int32_t* rowptr,
                        const int32_t* col, const double* val);

/* The L_p product's launch shapes, tuned (no counterpart in the reference): for each batch-size class (1, 2-4, 5-16
 * right-hand sides) the rows per row-pass workgroup (1, 2, 4) and waves per column-pass workgroup (1, 2, 4) that
 * ran `reps` products of 1, 4 and 8 right-hand sides fastest on this stack. The shapes change no arithmetic, so a
 * product's bits do not depend on them. rw_out, cw_out: 3 values each, the plan kept. gt_stack_set_lp_plan applies
 * a plan found before (the same for every stack of one shape). Both hold for the what-if stacks of the stack. */
int gt_stack_lp_autotune(gt_stack* stack, int32_t reps, int32_t* rw_out, int32_t* cw_out);
int gt_stack_set_lp_plan(gt_stack* stack, const int32_t* rw, const int32_t* cw);

/* ---------------------------------------------------------------- Schur complement preconditioner
 * Replaces peec.py:2345-2445 `_factorise` (SuperLU of S = Ar^T D^-1 Ar in a host thread) and the C shim
 * solver/accelerate_shim.c:21-59 (gt_factor / gt_solve / gt_free). With every strip of a stack alike, D^-1 is one
 * admittance d_s per strip and S = d_s (B + U C U^T): B = Ar^T Ar is the real graph Laplacian of the kept nodes
 * (strips and branches, weight 1), U = the branches' incidence columns, C = diag(d_br / d_s - 1). B is factorised
 * once per stack (nested dissection, multifrontal supernodal Cholesky: the small fronts on the host, the large ones
 * on the device, gt_ctx_schur_device_fronts), its triangular factors are solved on the device level by level, and
 * each frequency only changes the k x k Woodbury matrix, whose columns W = B^-1 U are device solves.
 *
 * node_xyz: n_keep x 3 (sheet, i, j) of the kept nodes (the nested-dissection geometry).
 * G_out: k x k (k = nbr) real, row-major: U^T B^-1 U, from which the caller forms each frequency's
 *        Q = C (I + G C)^-1 (peec.py:2357-2377's Woodbury identity). May be NULL when nbr == 0. */
int gt_schur_factor(gt_stack* stack, const int32_t* node_xyz, double* G_out);
/* The Schur factor of a stack whose sheets have different strip admittances (layers of different copper thickness;
 * peec.py:2379-2424 factorises S = Ar^T D^-1 Ar for any D). With lambda_s each sheet's strip admittance and Lambda
 * the diagonal of lambda over the kept nodes, S = Lambda^1/2 (B + E) Lambda^1/2 exactly: B = Ar^T Ar (the same real
 * Laplacian as gt_schur_factor's, branches at weight 1) and E = sum_m (d_m u~_m u~_m^T - u_m u_m^T) with
 * u~_m = Lambda^-1/2 u_m, nonzero only at the branches' kept end nodes. The Woodbury columns are therefore the
 * unit columns P of those nodes, ascending: cols[0 .. *n_cols) (at most 2048, the caller's array holds nbr * 2), and
 * G_out (n_cols x n_cols, row-major) = P^T B^-1 P. Each frequency's caller forms E_v = P^T E P and
 * Q = E_v (I + G E_v)^-1, and passes s_in = s_out = lambda^-1/2 per sheet in gt_op_desc. Every end node must lie in
 * the factor's last supernode (the sheet-apart ordering puts the nodes that branches join across sheets there), so
 * the correction runs on that supernode's rows between the solve's passes (the root form); otherwise
 * GT_ERR_UNSUPPORTED. Such a stack derives no what-if stacks (gt_stack_derive refuses): a what-if of it is a stack
 * of its own. */
int gt_schur_factor_nodes(gt_stack* stack, const int32_t* node_xyz, int32_t* n_cols, int32_t* cols, double* G_out);
/* The Woodbury form the factorised stack took (GT_WOODBURY_STORED, GT_WOODBURY_IMPLICIT or GT_WOODBURY_ROOT) and its
 * rank k (the branches). (No counterpart in the reference.) */
int gt_stack_woodbury(const gt_stack* stack, int32_t* form, int32_t* k);
/* A what-if stack (peec.py:1871-1909 `with_branches`): `base` plus n_extra branches (candidate vias) whose kept end
 * nodes are a_keep_extra / b_keep_extra (-1: a reference node), appended after the base's branches. It shares the
 * base's spectra, strip layout and Schur factor (held while any stack uses them; both order their work on the
 * base's context); its own are the incidence and the Woodbury columns W = B^-1 [U U'] of the base's factor, the added ones
 * made by device solves. The added branches are not in B: a caller's C for them is d_br / d_s, not d_br / d_s - 1.
 * G_out: k x k (k = base branches + n_extra), row-major U^T B^-1 U, or NULL. A branch that joins two pieces of
 * copper the base keeps apart changes the reference nodes; such a what-if needs a stack of its own. */
int gt_stack_derive(gt_stack* base, int32_t n_extra, const int32_t* a_keep_extra, const int32_t* b_keep_extra,
                    double* G_out, gt_stack** out);
/* The device bytes gt_stack_derive with these arguments would hold (*held: the stack's incidence and Woodbury
 * columns; it keeps no scratch of its own, its solves running in its base's batches) and need at its peak (*peak:
 * with the added columns' set-up solves), without making it: what a caller admits before the stack exists
 * (simulate.py:3930-3961 admits a task by its planned memory before it starts). (No counterpart in the reference.) */
int gt_stack_derive_bytes(const gt_stack* base, int32_t n_extra, const int32_t* a_keep_extra,
                          const int32_t* b_keep_extra, int64_t* held, int64_t* peak);
/* A stack of another copper topology on the same grid as `donor` (the same sheets, pitch and face gaps, so the same
 * inductance kernels): peec.py:3288-3290 reuses a board's kernels by their (nx, ny, d, dz) key for any model on that
 * grid; here the new stack shares the donor's resident kernel spectra (held while either stack lives) and makes its
 * own strip layout, incidence and (gt_schur_factor) factor. A topology-changing what-if (a moved slot). desc as for
 * gt_stack_create, without kernels (kx, ky ignored); GT_ERR_ARG when the grid or FFT sizes differ. */
int gt_stack_create_like(gt_ctx* ctx, const gt_stack_desc* desc, const gt_stack* donor, gt_stack** out);
/* Statistics of the factor: entries of L, supernodes, tree levels, the largest supernode's columns. */
int gt_schur_stats(const gt_stack* stack, int64_t* nnz_L, int32_t* n_super, int32_t* n_levels, int32_t* max_cols);
/* The dense top of the factor's tree (no counterpart in the reference, whose SuperLU solves have no levels): the tree
 * levels solved as dense products with the explicit inverse of their block of L (0: none), their columns, and the
 * inverse's stored entries (forward and backward stores). */
int gt_schur_top_stats(const gt_stack* stack, int32_t* levels, int32_t* cols, int64_t* entries);
/* B^-1 b on the host with the same factor (a check of the factorisation that needs no device). b, x: n_keep
 * complex values. */
int gt_schur_solve_host(const gt_stack* stack, const double* b, double* x);
/* B^-1 b on the device (level-scheduled triangular solves). b_dev, x_dev: n_keep complex values. */
int gt_schur_solve(gt_stack* stack, const void* b_dev, void* x_dev);

/* The same factorisation without a device: the role of solver/accelerate_shim.c:21-59 (gt_factor -> handle,
 * gt_solve(handle, x, k), gt_free) for B = Ar^T Ar given by the branches' kept end nodes (-1: a reference node).
 * gt_schur_host_solve solves k complex right-hand sides (n_keep each, consecutive) in place of b into x. */
typedef struct gt_schur_host_s gt_schur_host;
int gt_schur_host_factor(int32_t n_keep, int32_t nb, const int32_t* a_keep, const int32_t* b_keep,
                         const int32_t* node_xyz, gt_schur_host** out);
int gt_schur_host_solve(const gt_schur_host* h, const double* b, double* x, int32_t k);
int gt_schur_host_stats(const gt_schur_host* h, int64_t* nnz_L, int32_t* n_super, int32_t* n_levels,
                        int32_t* max_cols);
void gt_schur_host_free(gt_schur_host* h);

/* ---------------------------------------------------------------- per-frequency operator
 * Replaces peec.py:2150-2178 `_impedances` + :2465-2511 `_preconditioner_admittance` as uploaded by gpu.py:504-520. */
typedef struct {
    double w;                 /* angular frequency 2 pi f                                                         */
    const double* r_sheet;    /* 2 S: strip impedance per square of each sheet (surface impedance)               */
    const double* z_tail;     /* 2 nbr: branch impedances                                                         */
    const double* dinv_sheet; /* 2 S: preconditioner admittance of a strip on each sheet, 1/(R + jw L_self)      */
    const double* dinv_tail;  /* 2 nbr: 1 / z_tail                                                                */
    double s_scale[2];        /* 1 / d_s: scales B^-1 to S^-1                                                      */
    int32_t k;                /* Woodbury rank (the stack's column count, or 0)                                   */
    const double* q;          /* 2 k k: Q = C (I + G C)^-1, row-major                                             */
    const double* s_in;       /* 2 S or NULL: per-sheet scales of a stack factorised by gt_schur_factor_nodes     */
    const double* s_out;      /* (S^-1 = diag(s_out) (B + P E P^T)^-1 diag(s_in) by each kept node's sheet); NULL: */
                              /* the single s_scale                                                               */
} gt_op_desc;

int gt_op_create(gt_stack* stack, const gt_op_desc* desc, gt_op** out);
void gt_op_destroy(gt_op* op);

/* The system product y = K x = [Z Ib - Ar V ; Ar^T Ib] (peec.py:3043-3062 _HostMath.matvec, gpu.py:551-564). */
int gt_matvec(gt_op* op, const void* x_dev, void* y_dev);
/* gt_matvec run in host's scratch: op belongs to host or to a what-if stack derived from it, which then needs no
 * scratch of its own (a server's warm-start products; peec.py:2568-2576). (No counterpart in the reference.) */
int gt_matvec_on(gt_stack* host, gt_op* op, const void* x_dev, void* y_dev);
/* The preconditioner z = M^-1 r by block elimination (peec.py:3064-3083 _HostMath.precond, gpu.py:566-573),
 * with the Schur solve on the device. */
int gt_precond(gt_op* op, const void* r_dev, void* z_dev);

/* ---------------------------------------------------------------- GMRES
 * Replaces peec.py:3093-3224 `_gmres` for the right-preconditioned system (K M^-1) y = r0, x = M^-1 y
 * (peec.py:2581): CGS with a DGKS re-pass below 0.7, Givens rotations, the true-residual check with its aim and
 * miss budget, restarts and extra cycles while each falls to 0.7. The basis and every vector stay on the device;
 * per iteration only the Hessenberg column, the re-pass coefficients, the norms and the re-pass decision (taken on
 * the device) cross to the host. */
typedef struct {
    double tol;               /* stop at ||b - K x|| <= tol * bnorm                                               */
    double bnorm;             /* <= 0: ||b||                                                                      */
    int32_t restart;          /* <= 0: min(200, max(100, 4e8 / (16 n))) (peec.py:3087-3090)                      */
    int32_t max_cycles;       /* restarts (peec.py default 8)                                                     */
    int32_t extra_cycles;     /* further restarts while each still falls to 0.7 (24 for an own factorisation)   */
    int32_t max_its;          /* <= 0: no budget                                                                  */
    int32_t check_misses;     /* true-residual misses before a cycle restarts from its check (4)                 */
} gt_gmres_params;

typedef struct {
    int32_t iterations;
    int32_t converged;
    int32_t products;         /* system products made (K M^-1 applications)                                      */
    double residual;          /* ||b - K x|| / bnorm at the end                                                    */
} gt_gmres_result;

/* Solve K x = b (b, x: n = nb + n_keep complex values on the host). x holds the answer (also when unconverged:
 * the best so far, as the reference). Returns GT_OK; convergence is in result->converged. */
int gt_gmres(gt_op* op, const double* b_host, double* x_host, const gt_gmres_params* params,
             gt_gmres_result* result);
/* The same with b and x resident on the device (n complex values each, ordered on the context's stream; the
 * call returns when x is written). */
int gt_gmres_dev(gt_op* op, const void* b_dev, void* x_dev, const gt_gmres_params* params,
                 gt_gmres_result* result);
/* A batch of nr <= 16 solves on one stack, each with its own operator (frequency) and right-hand side: the many
 * frequencies and nets of one board (peec.py:2787 RouteExcitation.solve per (route, f); simulate.py:3964-4013 runs
 * them as parallel threads). Each follows the reference's GMRES exactly; the batch shares every launch, so the
 * spectra, incidence and Schur panels are read once per step. A solve's result is bit-identical alone or batched.
 * b_dev[i], x_dev[i]: device vectors of n complex values; results[i]: that solve's outcome. */
int gt_gmres_batch(gt_stack* stack, int32_t nr, gt_op* const* ops, const void* const* b_dev, void* const* x_dev,
                   const gt_gmres_params* params, gt_gmres_result* results);

/* The same batch in mixed precision: the reference's restarted GMRES (peec.py:3093-3224) whose restart cycles run in
 * FP32 (vectors, basis and operator products: the L_p FFT convolution and the incidence rows) while every restart
 * is an FP64 refinement step (peec.py:2560-2607's residual form); the Arnoldi dot products, the Givens
 * least-squares, the Schur solve of the preconditioner and every reduction stay FP64. Each round: r = b - K x in
 * FP64, one FP32 cycle of (K M^-1) y = r (params->restart iterations; it stops early only at the outer tolerance
 * relative to |r|, or at inner_tol, the smallest relative residual an FP32 cycle is asked for), x += M^-1 y in FP64;
 * until |r| <= tol * bnorm, after max_cycles rounds a round that does not bring |r| to 0.7 of the one before (the
 * reference's extra cycles, up to extra_cycles more), or max_refine rounds. results[i].iterations sums the cycles'
 * iterations; residual is the final FP64 relative residual. Not in the reference, which stays FP64 (its FP32 product
 * stalled GMRES, docs/ARCHITECTURE.md:299-301): here FP32 only carries the cycles, and each restart is exact. */
int gt_gmres_batch_mixed(gt_stack* stack, int32_t nr, gt_op* const* ops, const void* const* b_dev, void* const* x_dev,
                         const gt_gmres_params* params, double inner_tol, int32_t max_refine,
                         gt_gmres_result* results);
/* The same batch with the Krylov basis stored in FP32 and everything else FP64: the vectors w, x, r, the system
 * products and the preconditioner, the Gram-Schmidt dot products against the widened basis, the Givens
 * least-squares and the true-residual checks and restarts (each restart is a refinement step from the FP64
 * residual). The basis, which every Arnoldi step reads two or three times, moves half the bytes; the iteration
 * stays the reference's GMRES (peec.py:3093-3224) on a basis rounded to FP32. (No counterpart in the reference.) */
int gt_gmres_batch_fp32_basis(gt_stack* stack, int32_t nr, gt_op* const* ops, const void* const* b_dev,
                              void* const* x_dev, const gt_gmres_params* params, gt_gmres_result* results);

/* ---------------------------------------------------------------- serving sessions
 * Continuous batching of the solves of one stack (the device form of simulate.py:3964-4013 `run_parallel`, whose
 * threads each run a whole solve: here the solves share every launch and join or leave the batch between steps).
 * A session holds `capacity` <= 16 slots and a GMRES workspace of its own (restart: 0 for the default of
 * peec.py:3082-3090 `gmres_restart`), so sessions of several stacks, a board and its what-if stacks, interleave on
 * one context. gt_session_submit puts a solve into a free slot (*slot = -1 when all are taken); each
 * gt_session_step sets up the newly submitted solves, advances every running one by up to max_steps Arnoldi steps,
 * and writes x_dev of the solves that stopped; gt_session_poll gives a finished solve's outcome and frees its slot.
 * Each solve follows gt_gmres_batch's state machine, so its x and result are bit-identical to the same solve by
 * gt_gmres_dev, whatever shares the session with it. b_dev must stay unchanged until the poll; x_dev is written
 * by the step that finishes the solve. */
typedef struct gt_session_s gt_session;
int gt_session_create(gt_stack* stack, int32_t capacity, int32_t restart, gt_session** out);
/* The general form. basis: GT_BASIS_FP64, or GT_BASIS_FP32 for the Krylov basis stored in FP32 (as
 * gt_gmres_batch_fp32_basis). max_extra: the session's slots also take solves of what-if stacks derived from `stack`
 * (gt_stack_derive, directly or in a chain) with up to max_extra added branches, each over its own unknowns, in the
 * same launches as the stack's own solves: the board's what-ifs of different sizes batch together. gt_session_create
 * is gt_session_open(stack, capacity, restart, GT_BASIS_FP64, 0). (No counterpart in the reference, whose what-ifs
 * run as separate thread-pool jobs, simulate.py:3964-4013.) */
#define GT_BASIS_FP64 0
#define GT_BASIS_FP32 1
int gt_session_open(gt_stack* stack, int32_t capacity, int32_t restart, int32_t basis, int32_t max_extra,
                    gt_session** out);
void gt_session_destroy(gt_session* session);
int gt_session_submit(gt_session* session, gt_op* op, const void* b_dev, void* x_dev, const gt_gmres_params* params,
                      int32_t* slot);
/* *running: slots with a solve under way; *ready: finished solves waiting for their poll. */
int gt_session_step(gt_session* session, int32_t max_steps, int32_t* running, int32_t* ready);
/* *done = 1 and *result filled when the slot's solve has finished (the slot is then free), else *done = 0. */
int gt_session_poll(gt_session* session, int32_t slot, gt_gmres_result* result, int32_t* done);
/* Device bytes (the footprint admission plans with: gpu.py:593-605 `device_bytes`, made exact): what the session
 * holds, and what a session of `capacity` slots on `stack` would add, its workspace and the growth of the stack's
 * product scratch, before it is made. */
int64_t gt_session_device_bytes(const gt_session* session);
int64_t gt_session_plan_bytes(const gt_stack* stack, int32_t capacity, int32_t restart);
/* the same for gt_session_open's arguments (-1 for invalid ones) */
int64_t gt_session_open_bytes(const gt_stack* stack, int32_t capacity, int32_t restart, int32_t basis,
                              int32_t max_extra);

#ifdef __cplusplus
}
#endif

#endif /* GTPEEC_H */
