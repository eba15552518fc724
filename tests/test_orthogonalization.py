"""The GMRES orthogonalisations (gt_ctx_orthogonalization): the reference's CGS with the DGKS re-pass
(peec.py:3093-3224 `_gmres`) and the delayed CGS2 (the same second pass, always taken, one step late: two reads of the
basis per Arnoldi step). Both span the same Krylov spaces, so they take the same iterations and give the same solution
to rounding, and each matches the oracle; the delayed form keeps a solve's bits the same alone and batched, with FP64
and FP32 bases.
"""
import numpy as np
import pytest

from groundtruth_amd import _lib, boards, peec
from oracle import peec as O

pytestmark = pytest.mark.gpu


def _mode(st, mode):
    st.handle
    _lib.check(_lib.lib().gt_ctx_orthogonalization(st.ctx.handle, mode))


def _cur(s):
    return np.concatenate([s.I_x.ravel(), s.I_y.ravel(), s.I_branch])


@pytest.mark.parametrize("name", ["two_layer_37", "four_layer_24"])
def test_delayed_cgs2_matches_cgs_dgks_and_the_oracle(name):
    b = (boards.scaled("two_layer", 37, n_vias=7, n_freqs=4) if name == "two_layer_37"
         else boards.scaled("four_layer", 24, n_nets=2, n_freqs=3, n_stitch=3))
    r = b.routes[0]
    st = peec.PlaneStack(b.layers, b.branches)
    ex = peec.RouteExcitation(st, r.trace, r.driver, r.receiver)
    fs = [float(f) for f in b.freqs]
    try:
        _mode(st, 0)
        ref = ex.solve_many(fs, tol=1e-12)
        _mode(st, 1)
        got = ex.solve_many(fs, tol=1e-12)
    finally:
        _mode(st, 1)
    ox = O.RouteExcitation(O.PlaneStack(b.layers, b.branches), r.trace, r.driver, r.receiver)
    for f, a, c in zip(fs, got, ref):
        assert abs(a.iterations - c.iterations) <= 1, (f, a.iterations, c.iterations)
        assert np.linalg.norm(a.x - c.x) / np.linalg.norm(c.x) < 1e-9
        o = ox.solve(f, tol=1e-13)
        assert np.linalg.norm(_cur(a) - _cur(o)) / np.linalg.norm(_cur(o)) < 1e-9
        assert abs(a.Z_loop - o.Z_loop) / abs(o.Z_loop) < 1e-9


@pytest.mark.parametrize("precision", ["fp64", "fp32-basis", "mixed"])
def test_delayed_cgs2_alone_and_batched_gives_the_same_bits(precision):
    b = boards.scaled("two_layer", 24, n_vias=5, n_freqs=5)
    r = b.routes[0]
    st = peec.PlaneStack(b.layers, b.branches)
    ex = peec.RouteExcitation(st, r.trace, r.driver, r.receiver)
    fs = [float(f) for f in b.freqs]
    _mode(st, 1)
    many = ex.solve_many(fs, tol=1e-10, precision=precision)
    for f, s in zip(fs, many):
        one = ex.solve_many([f], tol=1e-10, precision=precision)[0]
        assert np.array_equal(s.x, one.x) and s.iterations == one.iterations


def test_delayed_cgs2_restarts_short_cycles_to_the_tolerance():
    """GMRES(8): many cycle ends and restarts, each cycle finishing its last basis row without a product."""
    b = boards.scaled("two_layer", 20, n_vias=4, n_freqs=2)
    r = b.routes[0]
    st = peec.PlaneStack(b.layers, b.branches)
    ex = peec.RouteExcitation(st, r.trace, r.driver, r.receiver)
    fs = [float(f) for f in b.freqs]
    try:
        _mode(st, 0)
        ref = ex.solve_many(fs, tol=1e-11, restart=8, max_cycles=60)
        _mode(st, 1)
        got = ex.solve_many(fs, tol=1e-11, restart=8, max_cycles=60)
    finally:
        _mode(st, 1)
    for f, a, c in zip(fs, got, ref):
        rhs_top, J = ex.rhs(f)
        rhs = np.concatenate([rhs_top, np.asarray(J, float)[st.keep].astype(complex)])
        Kx = st.matvec(st.operator(f), a.x)
        assert np.linalg.norm(rhs - Kx) / np.linalg.norm(rhs) < 2e-11
        assert np.linalg.norm(a.x - c.x) / np.linalg.norm(c.x) < 1e-8
