"""Mixed precision, two forms:
  "mixed": the reference's restarted GMRES whose cycles run in FP32 (vectors, basis and operator products: the L_p FFT
  convolution, the incidence rows) and whose every restart is an FP64 refinement step (r = b - K x, x += M^-1 y in
  FP64), with the Arnoldi dot products, the Givens least-squares, the Schur solve and every reduction in FP64;
  "fp32-basis": one FP64 GMRES whose Krylov basis alone is stored in FP32.
Bar: 1e-8 relative error on the strip currents and Z_loop against the FP64 oracle; the FP64 residual within the
tolerance; a solve's bits the same alone and batched. Both forms stop at their FP64 residual, as the plain method
does, and a plain FP64 solve to 1e-10 is itself only some 1e-8 from the 1e-13 solution on these boards, so both are
asked for 1e-11 to be held to 1e-8."""
import numpy as np
import pytest

from groundtruth_amd import boards, peec
from oracle import peec as O

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b))


def _currents(s):
    return np.concatenate([s.I_x.ravel(), s.I_y.ravel(), s.I_branch])


PRECISIONS = ["mixed", "fp32-basis"]
TOL_FOR_1E8 = {"mixed": 1e-11, "fp32-basis": 1e-11}


@pytest.mark.parametrize("precision", PRECISIONS)
@pytest.mark.parametrize("name", ["two_layer_20", "four_layer_24"])
def test_mixed_precision_is_within_1e8_of_the_oracle(name, precision):
    b = (boards.scaled("two_layer", 20, n_vias=5, n_freqs=1) if name == "two_layer_20"
         else boards.scaled("four_layer", 24, n_nets=2, n_freqs=1, n_stitch=3))
    r = b.routes[0]
    st = peec.PlaneStack(b.layers, b.branches)
    ex = peec.RouteExcitation(st, r.trace, r.driver, r.receiver)
    ox = O.RouteExcitation(O.PlaneStack(b.layers, b.branches), r.trace, r.driver, r.receiver)
    fs = [1e6, 1e8, 1e9]
    many = ex.solve_many(fs, tol=TOL_FOR_1E8[precision], precision=precision)
    for f, s in zip(fs, many):
        ref = ox.solve(f, tol=1e-13)
        assert rel(_currents(s), _currents(ref)) < 1e-8, f
        assert abs(s.Z_loop - ref.Z_loop) / abs(ref.Z_loop) < 1e-8, f


@pytest.mark.parametrize("precision", PRECISIONS)
def test_mixed_precision_alone_and_batched_gives_the_same_bits(precision):
    b = boards.scaled("two_layer", 20, n_vias=5, n_freqs=1)
    r = b.routes[0]
    st = peec.PlaneStack(b.layers, b.branches)
    ex = peec.RouteExcitation(st, r.trace, r.driver, r.receiver)
    fs = [3e6, 2e8, 7e8]
    many = ex.solve_many(fs, tol=1e-10, precision=precision)
    for f, s in zip(fs, many):
        one = ex.solve_many([f], tol=1e-10, precision=precision)[0]
        assert np.array_equal(s.x, one.x) and s.iterations == one.iterations


@pytest.mark.parametrize("precision", PRECISIONS)
def test_mixed_precision_residual_meets_the_tolerance_in_fp64(precision):
    b = boards.scaled("two_layer", 37, n_vias=7, n_freqs=1)
    r = b.routes[0]
    st = peec.PlaneStack(b.layers, b.branches)
    ex = peec.RouteExcitation(st, r.trace, r.driver, r.receiver)
    ref = O.PlaneStack(b.layers, b.branches)
    for f in (1e6, 5e8):
        sol = ex.solve_many([f], tol=1e-10, precision=precision)[0]
        rhs_top, J = ex.rhs(f)
        rhs = np.concatenate([rhs_top, J[st.keep]])
        w, R, zt = ref._impedances(O.cache_value(f))
        Kx = O.HostMath(ref).matvec(O.Operator(w, R, zt, None, None), sol.x)
        assert np.linalg.norm(Kx - rhs) <= 2e-10 * np.linalg.norm(rhs)


@pytest.mark.parametrize("precision", PRECISIONS)
def test_both_forms_take_about_the_fp64_iterations(precision):
    """The single-precision basis, and FP32 cycles restarted in FP64, cost a bounded number of extra iterations over
    FP64's GMRES: the restarts keep the reference's cycle structure, so no Krylov space is thrown away mid-cycle."""
    b = boards.scaled("two_layer", 37, n_vias=7, n_freqs=1)
    r = b.routes[0]
    st = peec.PlaneStack(b.layers, b.branches)
    ex = peec.RouteExcitation(st, r.trace, r.driver, r.receiver)
    fs = [1e6, 1e8, 1e9]
    d = ex.solve_many(fs, tol=1e-8)
    f32 = ex.solve_many(fs, tol=1e-8, precision=precision)
    for a, c in zip(f32, d):
        assert a.iterations <= 1.5 * c.iterations + 4


@pytest.mark.parametrize("precision", PRECISIONS)
def test_mixed_precision_with_4096_point_columns_of_six_sheets(precision):
    """Six sheets of 4096-point columns (S x P0 above one workgroup's LDS): the FP32 products take the residue-class
    column pass (four classes of bins mod 4, each S x 1024 points in LDS); the solve against the FP64 oracle."""
    from test_gpu_parity import _strip_board
    layers, branches = _strip_board(1100, 6, 6, seed=77)
    st, ref = peec.PlaneStack(layers, branches), O.PlaneStack(layers, branches)
    assert 4096 in tuple(st.Px) and st.n_sheets == 6
    rng = np.random.default_rng(11)
    rhs_top = rng.standard_normal(st.nb) + 1j * rng.standard_normal(st.nb)
    J = np.zeros(st.n_nodes)
    f = 3e8
    (Ib, V, its, x), = st.solve_many([f], [rhs_top], [J], tol=1e-11, precision=precision)
    Ib0, V0, its0, x0 = ref.solve(f, rhs_top, J, tol=1e-13)
    assert rel(Ib, Ib0) < 1e-8
