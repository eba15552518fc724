// Device vector work of GMRES for a batch of right-hand sides (krylov.hip): the Gram-Schmidt projections, norms,
// combinations and residuals of peec.py:2841-2897 `_HostVectors` / gpu.py:409-468 `TorchVectors`, each right-hand
// side with its own basis. Reductions are two-stage and in a fixed order (the same bits every run, and the same
// bits for a right-hand side alone or in any batch).
#pragma once

#include "gt_common.hpp"
#include "lp.hpp"

namespace gt {

// The vectors of a batch: basis rows V_p[0:m_p] (row stride n) and the vector w_p each pass works on. TV and TW are
// their storage types (double; float for the mixed-precision solves: the whole inner solve, or the basis alone);
// every product and sum is in double.
template <class TV, class TW = TV>
struct KrSetT {
    int nr = 0;
    int ld = 0;               // row stride of the bases (elements; a multiple of 64)
    const cpx<TV>* V[KMAX];
    int m[KMAX];
    cpx<TW>* w[KMAX];
    int n[KMAX];              // each right-hand side's vector length (0: the pass's n)
};
using KrSet = KrSetT<double>;

template <class T, class TA>
struct VecOpsT;

struct Krylov {
    int n = 0;
    int nblk = 0, nblk_d = 0;    // partial-sum blocks of the passes over w, and of the multi-dot, at the longest n
    static constexpr int chunk = 1024, chunk_d = 2048;    // their elements per block (whatever n: a right-hand
                                                          // side's sums do not depend on the batch's lengths)
    int rows_cap = 0;
    int P = 0;                // values per (right-hand side, block): 2 rows_cap (h) + 8 (norms) + 2 rows_cap (h2)
    DevBuf<double> part;      // KMAX x nblk x P partial sums
    DevBuf<double> red;       // KMAX x P reduced values

    void init(int n, int rows);
    // the partial-sum blocks the buffers hold for vectors of up to n
    static int part_blocks(int n) { return std::max(1, (n + chunk - 1) / chunk); }
    // The delayed-CGS2 passes: the multi-dot (dots2) takes twice the elements per block over an FP32 basis, so a
    // block moves the bytes of an FP64 one; the update (update2) takes `chunk` elements for either type, twice the
    // workgroups of an FP32 multi-dot to fill the device. Fixed per type, whatever the batch.
    template <class TV>
    static constexpr int chunk2() { return sizeof(TV) == sizeof(float) ? 2 * chunk : chunk; }
    template <class TV>
    int blocks2() const { return std::max(1, (n + chunk2<TV>() - 1) / chunk2<TV>()); }
    int blocks_u() const { return std::max(1, (n + chunk - 1) / chunk); }
    int slot_w() const { return 2 * rows_cap; }          // |w|^2 before the projection
    int slot_w_after() const { return 2 * rows_cap + 1; } // |w|^2 after it
    int slot_norm() const { return 2 * rows_cap + 2; }    // norm2's value
    int slot_w_after2() const { return 2 * rows_cap + 3; } // |w|^2 after the re-pass
    int slot_repass() const { return 2 * rows_cap + 4; }   // 1 where the re-pass was taken, else 0
    int slot_uu() const { return 2 * rows_cap + 5; }      // DCGS2: u^H u, then u^H v (2 values), then t at slot_h2
    int slot_h2() const { return 2 * rows_cap + 8; }      // the re-pass coefficients V^H w' (2 rows_cap values)
    // h_p = V_p[0:m_p]^H w_p and |w_p|^2
    template <class TV, class TW>
    void dots(const KrSetT<TV, TW>& k, hipStream_t s);
    // w_p -= V_p[0:m_p]^T h_p (h_p as reduced by dots, or the re-pass coefficients when `second`), |w_p|^2 after
    template <class TV, class TW>
    void project_update(const KrSetT<TV, TW>& k, hipStream_t s, bool second = false);
    // whether update_dots serves m rows
    static bool fused(int m) { return m <= 128; }
    // w_p -= V_p[0:m_p]^T h_p, |w_p|^2 after, and the re-pass coefficients V_p^H w_p of the updated vector: one
    // read of the basis for the update and the second classical Gram-Schmidt pass's products
    template <class TV, class TW>
    void update_dots(const KrSetT<TV, TW>& k, hipStream_t s);
    // after update_dots: the second pass w_p -= V_p^T h2_p where |w'_p| < 0.7 |w_p| (the DGKS test, decided on the
    // device and recorded in slot_repass), |w_p|^2 after it in slot_w_after2
    template <class TV, class TW>
    void repass(const KrSetT<TV, TW>& k, hipStream_t s);
    // |w_p|^2
    template <class TV, class TW>
    void norm2(const KrSetT<TV, TW>& k, hipStream_t s);
    // Delayed CGS2 (row m_p of V_p holds u_p, w_p holds v_p = A u_p or is null): t = V[0:m]^H u at slot_h2,
    // p = V[0:m]^H v at 0, u^H u and u^H v at slot_uu
    template <class TV, class TW>
    void dots2(const KrSetT<TV, TW>& k, hipStream_t s);
    // then q = (u - V t) ainv over row m, w = v - V p - q qv, |w|^2 at slot_w_after
    template <class TV, class TW>
    void update2(const KrSetT<TV, TW>& k, const double* ainv, const zd* qv, hipStream_t s);
    // dst_p = a_p / sqrt(red[p][slot_w_after])
    template <class T, class TA>
    void scale_by_norm(const VecOpsT<T, TA>& v, hipStream_t s);
};

// Per right-hand side: the destination (storage T) and the operands (storage TA; the same but in conversions)
template <class T, class TA = T>
struct VecOpsT {
    int nr = 0;
    cpx<T>* dst[KMAX];
    const cpx<TA>* a[KMAX];
    const cpx<TA>* b[KMAX];
    double f[KMAX];
    int k[KMAX];
    int n[KMAX];              // each right-hand side's vector length (0: the call's n, which bounds them all)
};
using VecOps = VecOpsT<double>;

template <class T, class TA>
void k_scale(const VecOpsT<T, TA>& v, int n, hipStream_t s);        // dst = f * a
template <class T>
void k_sub(const VecOpsT<T>& v, int n, hipStream_t s);              // dst = a - b
template <class T>
void k_copy(const VecOpsT<T>& v, int n, hipStream_t s);             // dst = a
template <class T>
void k_zero(const VecOpsT<T>& v, int n, hipStream_t s);             // dst = 0
// dst_p += sum_{r < k_p} a_p[r ld : r ld + n] * y[p rows + r]  (y on the device, `rows` values per right-hand side;
// ld the basis row stride)
template <class T, class TA>
void k_combine(const VecOpsT<T, TA>& v, const zd* y, int rows, int n, int ld, hipStream_t s);
// the row stride of a basis over vectors of n: whole 64-element lines, so that rows start aligned
inline size_t basis_ld(size_t n) { return (n + 63) / 64 * 64; }
// precision changes of whole vectors: dst_p = (float) a_p, and dst_p += (double) a_p
void k_to_float(int nr, cpx<float>* const* dst, const zd* const* a, int n, hipStream_t s);
void k_add_float(int nr, zd* const* dst, const cpx<float>* const* a, int n, hipStream_t s);

}  // namespace gt
