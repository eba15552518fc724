// Device vector work of GMRES (krylov.hip): the Gram-Schmidt projections, norms, combinations and residuals of
// peec.py:2841-2897 `_HostVectors` / gpu.py:409-468 `TorchVectors`, with deterministic two-stage reductions.
#pragma once

#include "gt_common.hpp"

namespace gt {

struct Krylov {
    int n = 0;
    int nblk = 0;                 // partial-sum blocks of the reductions
    DevBuf<double> part;          // nblk x (2 * (rows + 1)) partial sums
    DevBuf<double> red;           // reduced values: h (2 * rows), then |w|^2, |w_after|^2
    int rows_cap = 0;

    void init(int n, int rows);
    // h = V[0:m]^H w and |w|^2 -> red[0 : 2m], red[2m]
    void dots(const zd* V, int m, const zd* w, hipStream_t s);
    // w -= V[0:m]^T h (h = red[0:2m]); |w|^2 -> red[2m + 1]
    void project_update(const zd* V, int m, zd* w, hipStream_t s);
    // |v|^2 -> red[slot]
    void norm2(const zd* v, int slot, hipStream_t s);
};

void k_scale(zd* dst, const zd* src, double f, int n, hipStream_t s);                 // dst = f * src
void k_combine(zd* x, const zd* V, const zd* y_dev, int k, int n, hipStream_t s);     // x += V[0:k]^T y
void k_sub(zd* r, const zd* b, const zd* t, int n, hipStream_t s);                    // r = b - t
void k_add(zd* x, const zd* t, int n, hipStream_t s);                                 // x += t

}  // namespace gt
