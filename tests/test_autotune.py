"""The L_p launch-shape tuner (gt_stack_lp_autotune / gt_stack_set_lp_plan): every shape it may pick gives the same
bits, for single and batched launches, on boards whose FFTs take the register paths (256, 512 and 1024 points; at
1024 points in FP64 four rows of a workgroup exceed 64 KB of LDS, and the row pass takes a smaller register tile
rather than the LDS pass, whose radices round differently); the tuned plan is one of the allowed shapes and is
reused for the next stack of the same shape.
"""
import ctypes as C
import itertools

import numpy as np
import pytest

from groundtruth_amd import _lib, boards, peec

pytestmark = pytest.mark.gpu


def _lp_batch(st, vecs):
    import torch
    X = [torch.from_numpy(v).to("cuda:0") for v in vecs]
    Y = [torch.empty_like(v) for v in X]
    torch.cuda.synchronize()
    n = len(vecs)
    _lib.check(_lib.lib().gt_lp_matvec_batch(st.handle, n, (C.c_void_p * n)(*[v.data_ptr() for v in X]),
                                             (C.c_void_p * n)(*[v.data_ptr() for v in Y])))
    st.ctx.sync()
    return [y.cpu().numpy() for y in Y]


def _set(st, rw, cw):
    r = np.array([rw] * 3, np.int32)
    c = np.array([cw] * 3, np.int32)
    _lib.check(_lib.lib().gt_stack_set_lp_plan(st.handle, _lib.ptr(r, _lib.i32p), _lib.ptr(c, _lib.i32p)))


@pytest.mark.parametrize("name,n", [("two_layer", 150), ("four_layer", 100), ("two_layer", 300)])
def test_every_launch_shape_gives_the_same_bits(name, n):
    b = boards.scaled(name, n, n_freqs=1, **({"n_vias": 4} if name == "two_layer" else {"n_nets": 1}))
    st = peec.PlaneStack(b.layers, b.branches)
    rng = np.random.default_rng(2)
    vecs = [rng.standard_normal(st.ns) + 1j * rng.standard_normal(st.ns) for _ in range(8)]
    _set(st, 4, 1)
    ref = _lp_batch(st, vecs)
    for rw, cw in itertools.product((1, 2, 4), (1, 2, 4)):
        _set(st, rw, cw)
        for nr in (1, 3, 8):
            for a, r in zip(_lp_batch(st, vecs[:nr]), ref):
                assert np.array_equal(a, r), (rw, cw, nr)


def test_the_tuned_plan_is_allowed_and_reused():
    b = boards.scaled("two_layer", 120, n_vias=4, n_freqs=1)
    st = peec.PlaneStack(b.layers, b.branches)
    st.handle
    plan = st.lp_plan
    assert all(v in (1, 2, 4) for v in plan["rows_per_workgroup"] + plan["waves_per_workgroup"])
    st2 = peec.PlaneStack(b.layers, b.branches)
    st2.handle
    assert st2.lp_plan == plan
