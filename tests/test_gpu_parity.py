"""Parity of the device path (libgtpeec through its C ABI) with the oracle, on the same seeded boards.

Tolerances: the L_p product, the system product and the Schur solve are one FP64 computation in a different order
(FFT sizes, reduction order): <= 1e-13 / 1e-12 of the largest entry. The product checks build their stacks with the
host's kernel values (setup="host"), the oracle's own, so they compare the device's FFTs and rows alone; the
device-made kernels are pinned in test_kernel_gen.py. The whole solves use the device-made kernels. The preconditioner uses a different exact
factorisation (Cholesky of B with the Woodbury update, against SuperLU of S): <= 1e-10. Whole solves at tol 1e-12:
strip currents and Z_loop within 1e-9 relative (the parity bar), iterations within 2 of the oracle's.
"""
import numpy as np
import pytest

from groundtruth_amd import boards, peec
from oracle import peec as O

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.max(np.abs(np.asarray(a) - np.asarray(b))) / np.max(np.abs(np.asarray(b)))


def l2(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b))


CASES = {
    "single_plane_16": lambda: boards.single_plane(n=16),
    "two_layer_20": lambda: boards.scaled("two_layer", 20, n_vias=5, n_freqs=1),
    "four_layer_24": lambda: boards.scaled("four_layer", 24, n_nets=2, n_freqs=1, n_stitch=3),
    "two_layer_37x": lambda: boards.scaled("two_layer", 37, n_vias=7, n_freqs=1),
}


@pytest.fixture(scope="module", params=sorted(CASES))
def pair(request):
    b = CASES[request.param]()
    return b, peec.PlaneStack(b.layers, b.branches, setup="host"), O.PlaneStack(b.layers, b.branches)


def test_Lp_matvec_matches_the_oracle(pair):
    b, st, ref = pair
    rng = np.random.default_rng(4)
    I = rng.standard_normal(st.ns) + 1j * rng.standard_normal(st.ns)
    assert rel(st.Lp_matvec(I), ref.Lp_matvec(I)) < 1e-13


def test_batched_Lp_launch_gives_each_vector_its_single_product(pair):
    """gt_lp_matvec_batch: one launch of 5 vectors equals 5 single products, bit for bit."""
    import ctypes as C

    import torch

    from groundtruth_amd import _lib
    b, st, ref = pair
    rng = np.random.default_rng(9)
    Is = [rng.standard_normal(st.ns) + 1j * rng.standard_normal(st.ns) for _ in range(5)]
    X = [torch.from_numpy(v).to("cuda:0") for v in Is]
    Y = [torch.empty_like(v) for v in X]
    torch.cuda.synchronize()
    _lib.check(_lib.lib().gt_lp_matvec_batch(st.handle, 5, (C.c_void_p * 5)(*[v.data_ptr() for v in X]),
                                             (C.c_void_p * 5)(*[v.data_ptr() for v in Y])))
    st.ctx.sync()
    for v, y in zip(Is, Y):
        assert np.array_equal(y.cpu().numpy(), st.Lp_matvec(v))


def test_device_schur_solve_matches_the_host_factor(pair):
    b, st, ref = pair
    rng = np.random.default_rng(5)
    s = rng.standard_normal(st.n_keep) + 1j * rng.standard_normal(st.n_keep)
    assert rel(st.schur_solve(s), st.schur_solve_host(s)) < 1e-12


@pytest.mark.parametrize("f", [1e6, 1e8, 1e9])
def test_matvec_and_precond_match_the_oracle(pair, f):
    b, st, ref = pair
    n = st.nb + st.n_keep
    rng = np.random.default_rng(6)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    op, rop = st.operator(f), ref.operator(f)
    hm = O.HostMath(ref)
    assert rel(st.matvec(op, x), hm.matvec(rop, x)) < 1e-12
    assert rel(st.precond(op, x), hm.precond(rop, x)) < 1e-10


def _solve_both(b, f, tol=1e-12, route=0, restart=None):
    r = b.routes[route]
    st = peec.PlaneStack(b.layers, b.branches)
    sol = peec.RouteExcitation(st, r.trace, r.driver, r.receiver).solve(f, tol=tol, restart=restart)
    ref = O.RouteExcitation(O.PlaneStack(b.layers, b.branches), r.trace, r.driver, r.receiver).solve(f, tol=tol)
    return sol, ref


def _currents(s):
    return np.concatenate([s.I_x.ravel(), s.I_y.ravel(), s.I_branch])


@pytest.mark.parametrize("name,f", [("single_plane_16", 1e8), ("two_layer_20", 1e6), ("two_layer_20", 1e9),
                                    ("four_layer_24", 3e8), ("two_layer_37x", 5e7)])
def test_solution_matches_the_oracle(name, f):
    sol, ref = _solve_both(CASES[name](), f)
    assert l2(_currents(sol), _currents(ref)) < 1e-9
    assert abs(sol.Z_loop - ref.Z_loop) / abs(ref.Z_loop) < 1e-9
    assert l2(sol.V, ref.V) < 1e-8
    assert abs(sol.iterations - ref.iterations) <= 2


def test_single_plane_workload_at_full_size():
    """The 64 x 64 single-plane board at 100 MHz, GMRES(50) to 1e-10, as specified; checked at tol 1e-12."""
    b = boards.single_plane()
    sol, ref = _solve_both(b, 1e8, tol=1e-12, restart=50)
    assert l2(_currents(sol), _currents(ref)) < 1e-9
    assert abs(sol.Z_loop - ref.Z_loop) / abs(ref.Z_loop) < 1e-9


def test_solves_are_bit_reproducible():
    b = CASES["two_layer_20"]()
    r = b.routes[0]
    st = peec.PlaneStack(b.layers, b.branches)
    ex = peec.RouteExcitation(st, r.trace, r.driver, r.receiver)
    a, c = ex.solve(2e8, tol=1e-10), ex.solve(2e8, tol=1e-10)
    assert np.array_equal(a.x, c.x) and a.iterations == c.iterations


def test_true_residual_meets_the_tolerance_at_full_two_layer_size():
    """The 256 x 256 two-layer board at three of its frequencies: the device's answer satisfies K x = b to the
    tolerance through an independent product (the oracle's numpy matvec), and KCL holds at every node."""
    b = boards.two_layer()
    r = b.routes[0]
    st = peec.PlaneStack(b.layers, b.branches)
    ref = O.PlaneStack(b.layers, b.branches)
    ex = peec.RouteExcitation(st, r.trace, r.driver, r.receiver)
    for f in (b.freqs[0], b.freqs[5], b.freqs[9]):
        sol = ex.solve(f, tol=1e-10, restart=50)
        rhs_top, J = ex.rhs(f)
        rhs = np.concatenate([rhs_top, J[st.keep]])
        w, R, zt = ref._impedances(O.cache_value(f))
        rop = O.Operator(w, R, zt, None, None)
        Kx = O.HostMath(ref).matvec(rop, sol.x)
        assert np.linalg.norm(Kx - rhs) <= 1.5e-10 * np.linalg.norm(rhs)


def test_zero_right_hand_side_gives_zero():
    b = CASES["single_plane_16"]()
    st = peec.PlaneStack(b.layers, b.branches)
    n = st.nb + st.n_keep
    Ib, V, its, x = st.solve(1e8, np.zeros(st.nb, complex), np.zeros(st.n_nodes))
    assert its == 0 and not np.any(x)


def test_an_iteration_budget_too_small_is_a_solver_error():
    b = CASES["two_layer_20"]()
    r = b.routes[0]
    st = peec.PlaneStack(b.layers, b.branches)
    ex = peec.RouteExcitation(st, r.trace, r.driver, r.receiver)
    with pytest.raises(peec.SolverError):
        ex.solve(1e8, tol=1e-14, max_cycles=1, restart=2)


def test_batched_solves_are_bit_identical_to_single_solves_and_match_the_oracle():
    """The frequencies of one board solved as one batch: every result bit-identical to its solve alone (the batch
    shares launches, never arithmetic), and within the parity bar of the oracle."""
    b = CASES["two_layer_20"]()
    r = b.routes[0]
    st = peec.PlaneStack(b.layers, b.branches)
    ex = peec.RouteExcitation(st, r.trace, r.driver, r.receiver)
    fs = [1e6, 3e7, 2e8, 1e9, 5e5]
    many = ex.solve_many(fs, tol=1e-12)
    for f, s in zip(fs, many):
        one = ex.solve(f, tol=1e-12)
        assert np.array_equal(s.x, one.x) and s.iterations == one.iterations
    ox = O.RouteExcitation(O.PlaneStack(b.layers, b.branches), r.trace, r.driver, r.receiver)
    for f, s in zip(fs[:3], many[:3]):
        ref = ox.solve(f, tol=1e-12)
        assert l2(_currents(s), _currents(ref)) < 1e-9
        assert abs(s.Z_loop - ref.Z_loop) / abs(ref.Z_loop) < 1e-9


def test_a_mixed_batch_of_routes_and_frequencies_gives_each_its_own_answer():
    """Different nets at different frequencies in one batch (the serving mix): each result equals its own solve."""
    b = CASES["four_layer_24"]()
    st = peec.PlaneStack(b.layers, b.branches)
    exs = [peec.RouteExcitation(st, r.trace, r.driver, r.receiver) for r in b.routes]
    jobs = [(exs[0], 1e7), (exs[1], 4e8), (exs[0], 9e8), (exs[1], 2e6)]
    rhs = [ex.rhs(f) for ex, f in jobs]
    many = st.solve_many([f for _, f in jobs], [r[0] for r in rhs], [r[1] for r in rhs], tol=1e-10)
    for (ex, f), (Ib, V, its, x) in zip(jobs, many):
        one = ex.solve(f, tol=1e-10)
        assert np.array_equal(x, one.x) and its == one.iterations


@pytest.mark.parametrize("name,n,kw", [("single_plane", 200, {}), ("two_layer", 100, {"n_vias": 4}),
                                       ("two_layer", 200, {"n_vias": 4}), ("two_layer", 300, {"n_vias": 4}),
                                       ("four_layer", 100, {"n_nets": 1, "n_stitch": 4})])
def test_Lp_matvec_with_register_transforms_matches_the_oracle(name, n, kw):
    """Column lengths 256 / 512 / 1024 and 1, 2, 4 sheets: the L_p product's column pass with its transforms in
    registers (one wave per column) against the oracle's scipy FFTs."""
    b = boards.scaled(name, n, n_freqs=1, **kw) if name != "single_plane" else boards.single_plane(n=n)
    st, ref = peec.PlaneStack(b.layers, b.branches, setup="host"), O.PlaneStack(b.layers, b.branches)
    assert st.Px[0] in (256, 512, 1024)
    rng = np.random.default_rng(8)
    I = rng.standard_normal(st.ns) + 1j * rng.standard_normal(st.ns)
    assert rel(st.Lp_matvec(I), ref.Lp_matvec(I)) < 1e-13


def test_Lp_matvec_with_register_transforms_three_sheets():
    b = boards.scaled("four_layer", 100, n_freqs=1, n_nets=1, n_stitch=4)
    layers = b.layers[:3]
    branches = [br for br in b.branches if br.a[0] < 3 and br.b[0] < 3]
    st, ref = peec.PlaneStack(layers, branches, setup="host"), O.PlaneStack(layers, branches)
    assert st.n_sheets == 3 and st.Px[0] == 256
    rng = np.random.default_rng(9)
    I = rng.standard_normal(st.ns) + 1j * rng.standard_normal(st.ns)
    assert rel(st.Lp_matvec(I), ref.Lp_matvec(I)) < 1e-13


def _strip_board(nx, ny, sheets, seed):
    """Rectangular planes of nx x ny cells (two sheets 0.2 mm apart), random voids, stitching vias."""
    from groundtruth_amd.boards import CU_T, _stitch
    from groundtruth_amd.problem import Grid, PlaneLayer
    rng = np.random.default_rng(seed)
    layers = []
    for k in range(sheets):
        mask = rng.random((nx, ny)) > 0.05
        layers.append(PlaneLayer(f"L{k}", -k * (0.2e-3 + CU_T), Grid.from_mask(mask, 0.1), CU_T))
    branches = []
    if sheets > 1:
        while len(branches) < 4:
            i, j = int(rng.integers(1, nx - 1)), int(rng.integers(1, ny - 1))
            if all(l.grid.node[i, j] for l in layers):
                branches += _stitch(layers, i, j, f"via{len(branches)}")
    return layers, branches


@pytest.mark.parametrize("nx,ny,sheets,setup", [(1100, 6, 2, "host"), (6, 1100, 2, "host"), (1100, 5, 1, "host"),
                                                  (1100, 6, 3, "host"), (6, 1100, 5, "host"), (1100, 6, 6, "host"),
                                                  (1100, 6, 6, "device"), (1100, 7, 8, "device")])
def test_Lp_matvec_with_4096_point_transforms_matches_the_oracle(nx, ny, sheets, setup):
    """Transforms of 4096 points: the LDS passes' long-transform kernels (a 4096-point row per workgroup, S x P0 = 8192
    column points: 128 KB of LDS), and beyond 8192 column points the column pass through HBM (transforms, the
    sheets' multiply-add per pixel, inverse transforms), against the oracle's scipy FFTs, with the system product.
    setup "device": the device-made kernels, one per distinct face gap (equally spaced sheets: S of S x S). The
    host's far-field closed form (the reference's) loses digits to cancellation along a board 1100 cells long, so the
    two kernels differ by ~3e-11 of the product there: the bar is 1e-10 with the device's kernels (a wrong sheet-pair
    map would be off by 1e-2 or more) and 1e-13 with the host's, which pins the transforms themselves."""
    layers, branches = _strip_board(nx, ny, sheets, seed=nx + ny + sheets)
    st, ref = peec.PlaneStack(layers, branches, setup=setup), O.PlaneStack(layers, branches)
    bar = 1e-13 if setup == "host" else 1e-10
    assert 4096 in tuple(st.Px) + tuple(st.Py)
    rng = np.random.default_rng(10)
    I = rng.standard_normal(st.ns) + 1j * rng.standard_normal(st.ns)
    assert rel(st.Lp_matvec(I), ref.Lp_matvec(I)) < bar
    f = 2e8
    x = rng.standard_normal(st.nb + st.n_keep) + 1j * rng.standard_normal(st.nb + st.n_keep)
    hm = O.HostMath(ref)
    assert rel(st.matvec(st.operator(f), x), hm.matvec(ref.operator(f), x)) < 10 * bar
