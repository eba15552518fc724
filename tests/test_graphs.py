"""Schur solves captured in hipGraphs (gt_ctx_graphs): every replay gives the eager launches' bits. The replays are
interleaved with the program's own eager work (system products, host copies) and synchronisations, and each one is
checked, on its own and inside whole batched solves.
"""
import numpy as np
import pytest

from groundtruth_amd import _lib, boards, peec

pytestmark = pytest.mark.gpu


def _graphs(st, on):
    st.handle
    _lib.check(_lib.lib().gt_ctx_graphs(st.ctx.handle, 1 if on else 0))


def test_replayed_preconditioner_equals_the_eager_one():
    b = boards.scaled("two_layer", 48, n_vias=6, n_freqs=1)
    st = peec.PlaneStack(b.layers, b.branches)
    n = st.nb + st.n_keep
    rng = np.random.default_rng(5)
    rs = [rng.standard_normal(n) + 1j * rng.standard_normal(n) for _ in range(4)]
    op = st.operator(2e8)
    _graphs(st, False)
    eager = [st.precond(op, r) for r in rs]
    _graphs(st, True)
    try:
        for rep in range(3):
            for r, e in zip(rs, eager):
                got = st.precond(op, r)               # captured on first use, replayed after
                assert np.array_equal(got, e), rep
                st.matvec(op, r)                      # eager work between the replays
    finally:
        _graphs(st, False)


def test_batched_solves_with_graphs_equal_eager_solves():
    b = boards.scaled("two_layer", 40, n_vias=6, n_freqs=4)
    r = b.routes[0]
    st = peec.PlaneStack(b.layers, b.branches)
    ex = peec.RouteExcitation(st, r.trace, r.driver, r.receiver)
    fs = [float(f) for f in b.freqs]
    _graphs(st, False)
    eager = ex.solve_many(fs, tol=1e-10)
    one_eager = ex.solve(fs[1], tol=1e-10)
    _graphs(st, True)
    for rep in range(2):
        many = ex.solve_many(fs, tol=1e-10)
        for s, e in zip(many, eager):
            assert np.array_equal(s.x, e.x) and s.iterations == e.iterations
        one = ex.solve(fs[1], tol=1e-10)
        assert np.array_equal(one.x, one_eager.x)
    _graphs(st, False)
