"""What-if stacks (peec.py:1871-1923 `with_branches` / `map_state`, :2771-2774 `RouteExcitation.on`,
:1127-1144 `Solution.state_for`): a board plus candidate stitching vias, sharing the base's spectra and Schur
factor on the device, the added vias entering the preconditioner as Woodbury columns of the base's factor.

CPU: the derived topology equals a fresh stack's with all branches; map_state / state_for; the oracle's derived
solve equals a fresh solve; the Woodbury form with the added branches outside B equals S^-1 (host factor).
GPU: the derived device stack's products and solves against the oracle, a warm start through map_state, a via that
joins islands (a stack of its own), and the same result for a what-if alone and after other solves.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from groundtruth_amd import boards, peec
from groundtruth_amd.problem import Branch, Grid, PlaneLayer, via_impedance
from oracle import peec as O


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b))


def _board():
    return boards.scaled("two_layer", 20, n_vias=5, n_freqs=1)


def _extra_vias(b, count, seed):
    """Seeded stitching vias at cells where both layers have copper and no via stands yet."""
    rng = np.random.default_rng(seed)
    m0, m1 = b.layers[0].grid.node, b.layers[1].grid.node
    taken = {(br.a[1], br.a[2]) for br in b.branches}
    cells = [(i, j) for i, j in zip(*np.nonzero(m0 & m1)) if (i, j) not in taken]
    pick = rng.choice(len(cells), size=count, replace=False)
    length = abs(b.layers[0].z - b.layers[1].z)
    return [Branch((0, int(cells[p][0]), int(cells[p][1])), (1, int(cells[p][0]), int(cells[p][1])),
                   via_impedance(length, 0.3e-3), name=f"whatif{p}", kind="via") for p in pick]


def test_derived_topology_equals_a_fresh_stack():
    b = _board()
    extra = _extra_vias(b, 3, 7)
    st = peec.PlaneStack(b.layers, b.branches)
    d = st.with_branches(extra)
    fresh = peec.PlaneStack(b.layers, b.branches + extra)
    assert d is not st and d._base is st and d.nbr == st.nbr + 3 and d.nb == st.nb + 3
    assert np.array_equal(d.a_keep, fresh.a_keep) and np.array_equal(d.b_keep, fresh.b_keep)
    assert np.array_equal(d.keep, fresh.keep)
    assert d.nbr_factored == st.nbr
    # a second what-if on the first keeps the first base's factor
    d2 = d.with_branches(_extra_vias(b, 1, 9))
    assert d2._base is st and d2._parent is d and d2.nbr == st.nbr + 4
    # nothing new: the same stack
    assert st.with_branches([]) is st


def test_map_state_and_state_for():
    b = _board()
    st = peec.PlaneStack(b.layers, b.branches, setup="host")
    d = st.with_branches(_extra_vias(b, 2, 3))
    rng = np.random.default_rng(1)
    x = rng.standard_normal(st.nb + st.n_keep) + 1j * rng.standard_normal(st.nb + st.n_keep)
    y = d.map_state(x, st)
    assert len(y) == d.nb + d.n_keep
    assert np.array_equal(y[: st.nb], x[: st.nb]) and not np.any(y[st.nb: d.nb])
    assert np.array_equal(y[d.nb:], x[st.nb:])
    assert d.map_state(x, d) is x and d.map_state(None, st) is None
    assert st.map_state(y, d) is None
    # a Solution's unknown vector rebuilt from its currents and potentials
    r = b.routes[0]
    ex = peec.RouteExcitation(st, r.trace, r.driver, r.receiver)
    V = np.zeros(st.n_nodes, complex)
    V[st.keep] = x[st.nb:]
    sol = ex.unpack(1e8, x[: st.nb], V, 0, None)
    assert np.array_equal(sol.state_for(st), x)


def test_a_via_joining_islands_makes_a_stack_of_its_own():
    m0 = np.ones((24, 16), bool)
    m1 = np.ones((24, 16), bool)
    m1[12, :] = False                        # layer 1 cut in two
    layers = [PlaneLayer("a", 0.0, Grid.from_mask(m0, 0.4), 35e-6),
              PlaneLayer("b", 2e-4, Grid.from_mask(m1, 0.4), 35e-6)]
    via = lambda i, j: Branch((0, i, j), (1, i, j), via_impedance(2e-4, 3e-4), kind="via")
    st = peec.PlaneStack(layers, [via(3, 3)])
    d = st.with_branches([via(20, 5)])       # joins the lower piece of layer 1 to the rest
    assert d._base is None and d.nbr == 2 and d.n_keep == st.n_keep + 1
    ref = O.PlaneStack(layers, [via(3, 3)]).with_branches([via(20, 5)])
    assert ref.nbr == 2 and getattr(ref, "_base", None) is None


def test_oracle_derived_solve_equals_a_fresh_solve():
    b = boards.scaled("two_layer", 14, n_vias=3, n_freqs=1)
    extra = _extra_vias(b, 2, 5)
    r = b.routes[0]
    base = O.PlaneStack(b.layers, b.branches)
    d = base.with_branches(extra)
    fresh = O.PlaneStack(b.layers, b.branches + extra)
    s1 = O.RouteExcitation(base, r.trace, r.driver, r.receiver).on(d).solve(2e8, tol=1e-12)
    s2 = O.RouteExcitation(fresh, r.trace, r.driver, r.receiver).solve(2e8, tol=1e-12)
    assert rel(s1.x, s2.x) < 1e-10
    assert abs(s1.Z_loop - s2.Z_loop) / abs(s2.Z_loop) < 1e-10


def test_woodbury_with_added_branches_outside_the_factor():
    """The what-if preconditioner: B of the base's branches, the added vias as Woodbury columns with C = d/d_s
    (the base's with d/d_s - 1), against S^-1 of the extended incidence."""
    from test_host import HostFactor
    b = boards.scaled("two_layer", 16, n_vias=4, n_freqs=1)
    extra = _extra_vias(b, 3, 11)
    st = peec.PlaneStack(b.layers, b.branches)
    d = st.with_branches(extra)
    ref = O.PlaneStack(b.layers, b.branches).with_branches(extra)
    f = 7e7
    op = ref.operator(f)
    S = (ref.ArT @ sp.diags(op.Dinv) @ ref.Ar).tocsc()
    hf = HostFactor(st)                      # the base's factor: B of the base's branches
    k = d.nbr
    U = np.zeros((d.n_keep, k))
    for m in range(k):
        a, c = d.a_keep[d.ns + m], d.b_keep[d.ns + m]
        if a >= 0:
            U[a, m] += 1
        if c >= 0:
            U[c, m] -= 1
    W = hf.solve(U.T.astype(complex)).T.real
    Gm = U.T @ W
    ds = op.Dinv[0]
    c = op.Dinv[d.ns:] / ds - (np.arange(k) < d.nbr_factored)
    Q = np.diag(c) @ np.linalg.inv(np.eye(k) + Gm @ np.diag(c))
    rng = np.random.default_rng(4)
    s = rng.standard_normal(d.n_keep) + 1j * rng.standard_normal(d.n_keep)
    z = hf.solve(s)[0]
    y = (z - W @ (Q @ (U.T @ z))) / ds
    assert rel(y, spla.spsolve(S, s)) < 1e-10


# ---------------------------------------------------------------------------------------------------- device
@pytest.mark.gpu
def test_derived_products_match_the_oracle():
    b = _board()
    extra = _extra_vias(b, 3, 7)
    d = peec.PlaneStack(b.layers, b.branches).with_branches(extra)
    ref = O.PlaneStack(b.layers, b.branches).with_branches(extra)
    n = d.nb + d.n_keep
    rng = np.random.default_rng(6)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    for f in (1e6, 3e8):
        op, rop = d.operator(f), ref.operator(f)
        hm = O.HostMath(ref)
        assert rel(d.matvec(op, x), hm.matvec(rop, x)) < 1e-12
        assert rel(d.precond(op, x), hm.precond(rop, x)) < 1e-10


@pytest.mark.gpu
@pytest.mark.parametrize("f", [1e6, 1e9])
def test_whatif_solution_matches_the_oracle(f):
    b = _board()
    extra = _extra_vias(b, 4, 13)
    r = b.routes[0]
    st = peec.PlaneStack(b.layers, b.branches)
    d = st.with_branches(extra)
    sol = peec.RouteExcitation(st, r.trace, r.driver, r.receiver).on(d).solve(f, tol=1e-12)
    ref = O.RouteExcitation(O.PlaneStack(b.layers, b.branches + extra), r.trace, r.driver,
                            r.receiver).solve(f, tol=1e-12)
    cur = lambda s: np.concatenate([s.I_x.ravel(), s.I_y.ravel(), s.I_branch])
    assert rel(cur(sol), cur(ref)) < 1e-9
    assert abs(sol.Z_loop - ref.Z_loop) / abs(ref.Z_loop) < 1e-9
    assert abs(sol.iterations - ref.iterations) <= 2


@pytest.mark.gpu
def test_warm_start_from_the_base_solution():
    b = _board()
    r = b.routes[0]
    st = peec.PlaneStack(b.layers, b.branches)
    ex = peec.RouteExcitation(st, r.trace, r.driver, r.receiver)
    base_sol = ex.solve(1e8, tol=1e-10)
    d = st.with_branches(_extra_vias(b, 1, 21))
    exd = ex.on(d)
    cold = exd.solve(1e8, tol=1e-10)
    warm = exd.solve(1e8, tol=1e-10, guess=[d.map_state(base_sol.state_for(st), st)])
    assert rel(warm.x, cold.x) < 1e-8
    assert warm.iterations <= cold.iterations


@pytest.mark.gpu
def test_whatif_alone_and_after_other_solves_gives_the_same_bits():
    b = _board()
    r = b.routes[0]
    extra = _extra_vias(b, 2, 17)
    st1 = peec.PlaneStack(b.layers, b.branches)
    alone = peec.RouteExcitation(st1, r.trace, r.driver, r.receiver).on(st1.with_branches(extra)).solve(
        4e8, tol=1e-10)
    st2 = peec.PlaneStack(b.layers, b.branches)
    ex2 = peec.RouteExcitation(st2, r.trace, r.driver, r.receiver)
    ex2.solve(4e8, tol=1e-10)                # the base's own work on the shared factor first
    d2 = st2.with_branches(extra)
    ex2.on(d2).solve(1e6, tol=1e-10)
    mixed = ex2.on(d2).solve(4e8, tol=1e-10)
    assert np.array_equal(alone.x, mixed.x) and alone.iterations == mixed.iterations


@pytest.mark.gpu
def test_a_joined_island_solves_on_its_own_stack():
    m0 = np.ones((24, 16), bool)
    m1 = np.ones((24, 16), bool)
    m1[12, :] = False
    layers = [PlaneLayer("a", 0.0, Grid.from_mask(m0, 0.4), 35e-6),
              PlaneLayer("b", 2e-4, Grid.from_mask(m1, 0.4), 35e-6)]
    via = lambda i, j: Branch((0, i, j), (1, i, j), via_impedance(2e-4, 3e-4), kind="via")
    d = peec.PlaneStack(layers, [via(3, 3)]).with_branches([via(20, 5)])
    ref = O.PlaneStack(layers, [via(3, 3), via(20, 5)])
    from groundtruth_amd.problem import TraceSegment
    trace = [TraceSegment(np.array([0.8e-3, 2e-3, 4e-4]), np.array([8e-3, 2e-3, 4e-4]), 2e-4)]
    sol = peec.RouteExcitation(d, trace, (0, 2, 5), (0, 20, 5)).solve(1e8, tol=1e-12)
    rs = O.RouteExcitation(ref, trace, (0, 2, 5), (0, 20, 5)).solve(1e8, tol=1e-12)
    assert abs(sol.Z_loop - rs.Z_loop) / abs(rs.Z_loop) < 1e-9
