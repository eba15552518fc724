"""The rasteriser's oracle (oracle/raster.py) pinned without shapely:
- point containment against the winding number in rational arithmetic, points on edges (any slope) outside;
- line covers and cell areas against exact clipping (segment intersections in rationals, Sutherland-Hodgman),
  as the reference's TW9 holds them to shapely's intersections (tests/test_track_width.py:306-337);
- the sub-grid's fill in closed form for a box between fine samples;
- `_adopt` and `_series_factor` on hand-made arrays (the harmonic mean of a width step);
- the reference's own raster expectations: webs between antipads stay connected only on the sub-grid, a slot
  cuts (tests/test_raster.py:24-37), the fill follows perforated copper (:40-64), a track through another net's
  pour stays joined with bridges of its width (:67-85), SP-27's strip widths and centres
  (tests/test_validation_audit.py:1148-1165), touching fills are one piece and a 0.1 mm gap is not (TW12).
"""
from fractions import Fraction

import numpy as np
import pytest
import scipy.sparse as sp
from scipy.sparse.csgraph import connected_components

from oracle import raster as R
from tests import _raster_cases as C


# ---------------------------------------------------------------------------------------------- exact references
def _F(v):
    return Fraction(float(v))


def _on_segment(a, b, p) -> bool:
    ax, ay, bx, by, px, py = map(_F, (*a, *b, *p))
    if (bx - ax) * (py - ay) - (by - ay) * (px - ax) != 0:
        return False
    return min(ax, bx) <= px <= max(ax, bx) and min(ay, by) <= py <= max(ay, by)


def _winding_inside(polys, p) -> bool:
    """Interior by the winding number: odd total winding over the rings (holes wind the other way); each edge's side
    in floats where it is clear by far, else in rationals."""
    px, py = float(p[0]), float(p[1])
    total = 0
    for poly in polys:
        for r in poly:
            a, b = r, np.roll(r, -1, axis=0)
            cross = (b[:, 0] - a[:, 0]) * (py - a[:, 1]) - (b[:, 1] - a[:, 1]) * (px - a[:, 0])
            scale = np.abs(b - a).sum(1) * (abs(px) + abs(py) + np.abs(a).sum(1) + 1.0)
            for k in np.nonzero(np.abs(cross) <= 1e-9 * scale)[0]:
                cr = ((_F(b[k, 0]) - _F(a[k, 0])) * (_F(py) - _F(a[k, 1]))
                      - (_F(b[k, 1]) - _F(a[k, 1])) * (_F(px) - _F(a[k, 0])))
                if cr == 0 and _on_segment(a[k], b[k], p):
                    return False
                cross[k] = float((cr > 0) - (cr < 0))
            up = (a[:, 1] <= py) & (py < b[:, 1]) & (cross > 0)
            down = (b[:, 1] <= py) & (py < a[:, 1]) & (cross < 0)
            total += int(up.sum()) - int(down.sum())
    return total % 2 != 0


def _clip(ring, x0, y0, x1, y1):
    """Sutherland-Hodgman: a ring clipped to the rectangle (floats)."""
    pts = [tuple(p) for p in ring]
    for inside, cut in (
            (lambda p: p[0] >= x0, lambda a, b: (x0, a[1] + (x0 - a[0]) * (b[1] - a[1]) / (b[0] - a[0]))),
            (lambda p: p[0] <= x1, lambda a, b: (x1, a[1] + (x1 - a[0]) * (b[1] - a[1]) / (b[0] - a[0]))),
            (lambda p: p[1] >= y0, lambda a, b: (a[0] + (y0 - a[1]) * (b[0] - a[0]) / (b[1] - a[1]), y0)),
            (lambda p: p[1] <= y1, lambda a, b: (a[0] + (y1 - a[1]) * (b[0] - a[0]) / (b[1] - a[1]), y1))):
        out = []
        for k in range(len(pts)):
            a, b = pts[k - 1], pts[k]
            if inside(b):
                if not inside(a):
                    out.append(cut(a, b))
                out.append(b)
            elif inside(a):
                out.append(cut(a, b))
        pts = out
        if not pts:
            break
    return pts


def _area(pts) -> float:
    if len(pts) < 3:
        return 0.0
    x, y = np.array(pts).T
    return 0.5 * float(np.sum(x * np.roll(y, -1) - np.roll(x, -1) * y))


def _pieces(g) -> int:
    n = g.nx * g.ny
    idx = np.arange(n).reshape(g.nx, g.ny)
    rows = np.concatenate([idx[:-1, :][g.xb], idx[:, :-1][g.yb]] + [np.array([idx[b[0], b[1]] for b in g.bridges],
                                                                                int)])
    cols = np.concatenate([idx[1:, :][g.xb], idx[:, 1:][g.yb]] + [np.array([idx[b[2], b[3]] for b in g.bridges],
                                                                               int)])
    _, lab = connected_components(sp.coo_matrix((np.ones(len(rows)), (rows, cols)), shape=(n, n)), directed=False)
    return lab


SHAPES = {
    "webbed": C.webbed_plane,
    "diagonal": lambda: list(C.diagonal_lattice().values()),
    "necked": lambda: C.necked_track()["N"],
    "angled": lambda: C.angled_track(30.0)[1]["T"],
}


def _polys(shape):
    if isinstance(shape, list):
        return [p for g in shape for p in R.polygons(g)]
    return R.polygons(shape)


# ---------------------------------------------------------------------------------------------- containment
@pytest.mark.parametrize("name", sorted(SHAPES))
def test_contains_matches_the_winding_number(name):
    polys = _polys(SHAPES[name]())
    rng = np.random.default_rng(3)
    x0, y0, x1, y1 = R.bounds(polys)
    pts = [tuple(p) for p in rng.uniform((x0 - 1, y0 - 1), (x1 + 1, y1 + 1), size=(200, 2))]
    # points on the boundary: vertices and edge midpoints (exact for axis-parallel and 45 degree edges), and a
    # lattice through them
    for poly in polys:
        for r in poly:
            for k in range(0, len(r), max(1, len(r) // 8)):
                a, b = r[k], r[(k + 1) % len(r)]
                pts += [tuple(a), ((a[0] + b[0]) / 2, (a[1] + b[1]) / 2)]
    X, Y = np.meshgrid(np.arange(np.floor(x0), np.ceil(x1) + 1, 0.5), np.arange(np.floor(y0), np.ceil(y1) + 1, 0.5),
                       indexing="ij")
    pts += list(zip(X.ravel()[::7], Y.ravel()[::7]))
    P = np.array(pts)
    got = R.contains(polys, P[:, 0], P[:, 1])
    want = np.array([_winding_inside(polys, p) for p in pts])
    assert np.array_equal(got, want)
    on_edge = [p for p in pts[200:] if any(_on_segment(r[k], r[(k + 1) % len(r)], p) for poly in polys for r in poly
                                           for k in range(len(r)))]
    assert on_edge and not R.contains(polys, *np.array(on_edge).T).any()


# ---------------------------------------------------------------------------------------------- covers and areas
@pytest.mark.parametrize("name", ["webbed", "diagonal", "angled"])
def test_line_cover_matches_exact_clipping(name):
    polys = _polys(SHAPES[name]())
    pa, pb = R.ring_edges(polys)
    x0, y0, x1, y1 = R.bounds(polys)
    d = 0.37
    u0, n_lines = x0 - 0.3, int((x1 - x0) / 0.29) + 2
    v0, n_cells = y0 - 0.2, int((y1 - y0) / d) + 3
    got = R.line_cover(pa[:, 0], pa[:, 1], pb[:, 0], pb[:, 1], u0, 0.29, n_lines, v0, d, n_cells)
    for m in range(0, n_lines, 3):
        u = u0 + m * 0.29
        ts = []
        for (ax, ay), (bx, by) in zip(pa, pb):
            if min(ax, bx) <= u <= max(ax, bx) and ax != bx:
                ts.append(float(_F(ay) + (_F(u) - _F(ax)) * (_F(by) - _F(ay)) / (_F(bx) - _F(ax))))
        for c in range(n_cells):
            lo, hi = v0 + (c - 0.5) * d, v0 + (c + 0.5) * d
            cuts = sorted({lo, hi} | {t for t in ts if lo < t < hi})
            want = sum(b - a for a, b in zip(cuts[:-1], cuts[1:]) if _winding_inside(polys, (u, (a + b) / 2)))
            assert abs(got[m, c] - want) <= 1e-12, (m, c)


@pytest.mark.parametrize("name", ["webbed", "diagonal", "necked"])
def test_cell_areas_match_polygon_clipping(name):
    polys = _polys(SHAPES[name]())
    pa, pb = R.ring_edges(polys)
    x0, y0, x1, y1 = R.bounds(polys)
    d = 0.45
    xl, yl = x0 - 0.71, y0 - 0.61          # no grid line on an edge (there the reference's formula drops it)
    wi, wj = int((x1 - xl) / d) + 2, int((y1 - yl) / d) + 2
    right = R.line_cover(pa[:, 0], pa[:, 1], pb[:, 0], pb[:, 1], xl + d, d, wi, yl + d / 2, d, wj)
    got = R.cell_areas(pa, pb, xl, yl, d, wi, wj, right)
    rings = [r if (R._signed_area(r) > 0) == (h == 0) else r[::-1] for p in polys for h, r in enumerate(p)]
    want = np.zeros((wi, wj))
    for i in range(wi):
        for j in range(wj):
            want[i, j] = sum(_area(_clip(r, xl + i * d, yl + j * d, xl + (i + 1) * d, yl + (j + 1) * d)) for r in rings)
    assert np.max(np.abs(got - want)) <= 1e-12
    total = sum(R._signed_area(r) for r in rings)
    assert abs(got.sum() - total) <= 1e-11 * abs(total)


# ---------------------------------------------------------------------------------------------- the sub-grid
def test_fine_fill_of_a_box_between_samples():
    """A box whose edges fall between fine samples: each cell keeps its samples inside the box, counted in closed
    form."""
    d, s = 1.0, 4
    bx0, by0, bx1, by1 = 1.05, 2.2, 6.3, 4.9
    g = R.build_grid({"A": C.box(bx0, by0, bx1, by1)}, (0, 0, 8, 7), d, sub=s)
    xf = -d / 2 + (np.arange(g.nx * s) + 0.5) * d / s
    yf = -d / 2 + (np.arange(g.ny * s) + 0.5) * d / s
    inx = ((xf > bx0) & (xf < bx1)).reshape(g.nx, s).sum(1)
    iny = ((yf > by0) & (yf < by1)).reshape(g.ny, s).sum(1)
    count = np.outer(inx, iny)
    keep = (count >= s) & (inx > 0)[:, None] & (iny > 0)[None, :]
    assert np.array_equal(g.node, keep)
    assert np.array_equal(g.fill, np.where(keep, count / (s * s), 0.0))


def test_adopt_and_series_factor_by_hand():
    cov = np.array([[0.0, 0.3, 0.2, 0.0, 0.5]])
    has = np.array([[False, True, False, False, True]])
    out = R.adopt(cov, has, ((0, 1), (0, -1)), 3)
    # 0.2 at column 2: columns 1 (own 0.3) and 3 (none) at one step; it goes to column 1
    assert np.allclose(out, [[0.0, 0.5, 0.0, 0.0, 0.5]])
    # a one-strip band whose copper is 0.4 wide on three sample lines and 0.1 on three, 0.1 at its border
    face = np.array([[0.0] * 8 + [0.1] + [0.0] * 8])
    has = face > 0
    samples = [np.where(has, v, 0.0) for v in (0.4, 0.4, 0.4, 0.1, 0.1, 0.1)]
    f = R.series_factor(face, samples, has, 1)
    assert f[0, 8] == pytest.approx((6 / (3 / 0.4 + 3 / 0.1)) / 0.1, rel=1e-14)
    # another strip within 2 ADOPT_FACES lanes: not alone, factor 1
    has2 = has.copy()
    has2[0, 12] = True
    assert R.series_factor(np.where(has2, 0.1, 0.0), [np.where(has2, v, 0.0) for v in (0.4,) * 3 + (0.1,) * 3],
                           has2, 1)[0, 8] == 1.0


# ---------------------------------------------------------------------------------------------- the reference's tests
def test_antipad_webs_and_slot():
    g1 = R.build_grid(C.webbed_plane(), (0, 0, 20, 10), 0.5)
    lab = _pieces(g1)
    assert len(set(lab[g1.node.ravel()])) == 2
    g5 = R.build_grid(C.webbed_plane(), (0, 0, 20, 10), 0.5, sub=5)
    assert len(set(_pieces(g5)[g5.node.ravel()])) == 1
    assert 0.2 < g5.fill[20, 4] < 1.0 and g5.fill[5, 5] == 1.0
    gs = R.build_grid(C.slotted_plane(), (0, 0, 20, 10), 0.5, sub=5)
    assert len(set(_pieces(gs)[gs.node.ravel()])) == 2


def test_fill_follows_perforated_copper():
    solid = R.build_grid(C.box(*C.PLANE), (0, 0, 20, 10), 0.25, sub=4)
    holes = R.build_grid(C.perforated_plane(), (0, 0, 20, 10), 0.25, sub=4)
    expected = C.perforated_plane().area / C.box(*C.PLANE).area
    assert abs(holes.fill.sum() / solid.fill.sum() - expected) < 0.03


@pytest.mark.parametrize("d", [0.86, 0.61, 0.5, 0.43, 0.3])
def test_track_through_another_nets_pour_stays_joined(d):
    pts, geom = C.track_through_pour()
    sub = max(1, min(6, int(np.ceil(d / 0.15 - 1e-9))))
    g = R.build_grid(geom, (0, 0, 20, 10), d, seeds={"A": [pts[0], pts[-1]]}, lines={"A": [(pts, 0.4)]}, sub=sub)
    ends = []
    for x, y in (pts[0], pts[-1]):
        ii, jj = np.nonzero(g.label == g.nets.index("A"))
        k = int(np.argmin((g.x0 + ii * d - x) ** 2 + (g.y0 + jj * d - y) ** 2))
        ends.append((ii[k], jj[k]))
    lab = _pieces(g).reshape(g.nx, g.ny)
    assert lab[ends[0]] == lab[ends[1]], f"track cut at d = {d}"
    assert all(w == 0.4 and length >= d for *_, length, w in g.bridges)
    # B's pour keeps one piece
    b = g.label == g.nets.index("B")
    assert len(set(lab[b])) == 1


@pytest.mark.parametrize("d, y", [(1.0, 12.0), (1.0, 12.4), (0.5, 12.1), (0.35, 12.0)])
def test_sp27_a_narrow_strip_keeps_its_width_and_centre(d, y):
    w = 0.35
    sub = int(np.ceil(d / 0.15))
    g = R.build_grid(C.narrow_strip(y, w), (0.0, 0.0, 20.0, 24.0), d, sub=sub)
    for i in range(int(4 / d), int(16 / d)):
        jj = np.nonzero(g.xb[i])[0]
        assert abs(g.wx[i, jj].sum() * d - w) <= d / sub + 1e-9, (i, g.wx[i, jj])
        centre = g.y0 + np.sum((jj + g.ox[i, jj]) * g.wx[i, jj]) / g.wx[i, jj].sum() * d
        assert centre == pytest.approx(y, abs=d / sub)


def test_touching_fills_are_one_piece():
    g = R.build_grid(C.touching_fills(), (0, 0, 15, 5), 0.5, sub=4)
    left, mid, right = g.part[4, 4], g.part[16, 4], g.part[26, 4]
    assert left == mid and mid != right and right >= 0


def test_areas_sum_to_the_copper():
    """TW9: the nodes' areas add up to the copper's area."""
    for geom, bb, d in ((C.perforated_plane(), (0, 0, 20, 10), 0.25), (C.necked_track(), (0, 0, 20, 10), 0.2)):
        g = R.build_grid(geom, bb, d, sub=4)
        polys = R.polygons(geom if not isinstance(geom, dict) else list(geom.values())[0])
        total = sum(abs(R._signed_area(p[0])) - sum(abs(R._signed_area(h)) for h in p[1:]) for p in polys)
        assert float(np.sum(np.where(g.node, g.area, 0.0))) * d * d == pytest.approx(total, rel=0.005)


def test_sp27_two_touching_partial_strips_are_one_strip():
    """tests/test_validation_audit.py SP-27: two side-by-side strips phi d wide, touching (centres phi d apart),
    carrying half the current each, have the inductance of one 2 phi d strip: (self + dL + M) / 2 against the flat
    strip's integral, within 1 % (the oracle's narrow_self and side_mutual, zero thickness)."""
    from oracle import kernel as K
    from oracle import peec as O
    d = 1e-3
    full = K.strip_kernel(3, 3, d)[2, 2]
    assert full == pytest.approx(K.strip_self_numeric(d, d), rel=1e-6)
    for phi in (0.1, 0.15, 0.3, 0.5):
        self_ = full + O.narrow_self(d, phi, 0.0)
        m = float(O.side_mutual(d, phi, phi, phi)[0])
        assert (self_ + m) / 2 == pytest.approx(K.strip_self_numeric(d, 2 * phi * d), rel=0.01), phi
