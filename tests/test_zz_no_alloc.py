"""gt_lp_matvec_batch_noalloc (the L_p product with its intermediates in the library's fixed device scratch, no
allocation) against gt_lp_matvec_batch (the stack's own workspace): the same kernels and launch shapes, so the outputs
must be equal bit for bit. Two calls in a row on the shared scratch, the second with a smaller input (a smaller board
and fewer right-hand sides), so stale contents of the first call would show; and a batch above the scratch's size is
refused with an error, not run."""
import ctypes as C

import numpy as np
import pytest

from groundtruth_amd import _lib, boards, peec

pytestmark = pytest.mark.gpu


def _batch(st, fn, vecs):
    import torch
    X = [torch.from_numpy(v).to("cuda:0") for v in vecs]
    Y = [torch.full_like(x, float("nan")) for x in X]
    torch.cuda.synchronize()
    n = len(X)
    code = fn(st.handle, n, (C.c_void_p * n)(*[x.data_ptr() for x in X]), (C.c_void_p * n)(*[y.data_ptr() for y in Y]))
    msg = _lib.lib().gt_last_error()        # before the sync, which clears it
    st.ctx.sync()
    return code, msg, [y.cpu().numpy() for y in Y]


def _vecs(st, n, seed):
    rng = np.random.default_rng(seed)
    return [rng.standard_normal(st.ns) + 1j * rng.standard_normal(st.ns) for _ in range(n)]


def test_noalloc_product_equals_the_allocating_one_on_two_calls_in_a_row():
    L = _lib.lib()
    big = boards.scaled("two_layer", 40, n_vias=4, n_freqs=1)
    small = boards.single_plane(n=16)
    for b, n, seed in ((big, 3, 11), (small, 2, 12)):
        st = peec.PlaneStack(b.layers, b.branches)
        vecs = _vecs(st, n, seed)
        code, msg, got = _batch(st, L.gt_lp_matvec_batch_noalloc, vecs)
        assert code == _lib.GT_OK, msg
        code, msg, ref = _batch(st, L.gt_lp_matvec_batch, vecs)
        assert code == _lib.GT_OK, msg
        for g, r in zip(got, ref):
            assert np.array_equal(g, r)


def test_noalloc_product_refuses_a_batch_above_its_scratch():
    b = boards.scaled("two_layer", 256, n_vias=4, n_freqs=1)
    st = peec.PlaneStack(b.layers, b.branches)
    code, msg, _ = _batch(st, _lib.lib().gt_lp_matvec_batch_noalloc, _vecs(st, 16, 13))
    assert code == _lib.GT_ERR_ARG
    assert b"scratch" in msg


def test_noalloc_product_on_a_stack_whose_columns_exceed_one_workgroup():
    """Three sheets of 4096-point columns (the residue-class column pass): the same bits without allocation."""
    from test_gpu_parity import _strip_board
    L = _lib.lib()
    layers, branches = _strip_board(1100, 6, 3, seed=21)
    st = peec.PlaneStack(layers, branches)
    assert 4096 in tuple(st.Px)
    vecs = _vecs(st, 2, 14)
    code, msg, got = _batch(st, L.gt_lp_matvec_batch_noalloc, vecs)
    assert code == _lib.GT_OK, msg
    code, msg, ref = _batch(st, L.gt_lp_matvec_batch, vecs)
    assert code == _lib.GT_OK, msg
    for g, r in zip(got, ref):
        assert np.array_equal(g, r)
