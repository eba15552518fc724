"""Oracle restatement of the copper rasterisation in `src/groundtruth/solver/peec.py:329-983` (`build_grid` and its
helpers), numpy + scipy, no shapely.

TEST INFRASTRUCTURE (see oracle/__init__.py): only tests, smoke() and bench.py's cpu_baseline use it.

Geometry arrives as GeoJSON-like mappings or objects with `__geo_interface__` (a shapely geometry has one): Polygon,
MultiPolygon and GeometryCollection; lines and points are left out, as `_polygons` (peec.py:873-883) leaves them.
Where the reference asks shapely, this module answers the same question itself:
- `contains_xy(g, x, y)` (peec.py:379-385, :791-797): `contains`, the point in the interior of the polygons'
  even-odd region, decided by exact orientation signs (a float filter, then rational arithmetic), so a point on an
  edge is outside whatever the edge's slope;
- `_pieces`' `dwithin` grouping (peec.py:852-870): polygon distance by segment distances and containment;
- `covers` of a point (peec.py:395-405): inside or on the boundary;
- `shortest_line` to the copper (peec.py:472-474): the nearest point of the rings' edges;
- `orient_polygons` (peec.py:522-527): signed ring areas.
Every other step restates the reference's arithmetic: the sample coordinates, the windows, `_coarse_from_fine`'s
pieces and crossings, `_line_cover`, `_cell_areas`, `_adopt`, `_series_factor`, `_strip_widths`, `_exact_copper`,
`_heat_points` and `_rasterise_lines`.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from fractions import Fraction
from typing import Optional

import numpy as np
import scipy.sparse as sp
from scipy.sparse.csgraph import connected_components

PIECE_TOL_MM = 0.005        # peec.py:106
STRIP_SAMPLES = 6           # peec.py:107
BAND_STRIPS = 4             # peec.py:108
ADOPT_FACES = 3             # peec.py:110
ADOPT_CELLS = 2             # peec.py:111


# ---------------------------------------------------------------------------------------------------- geometry
def polygons(geom) -> list:
    """The polygons of a geometry, each a list of rings (exterior first) as (n, 2) float arrays without the closing
    point; lines and points left out (peec.py:873-883)."""
    if geom is None:
        return []
    if isinstance(geom, list) and all(isinstance(p, list) for p in geom):
        return geom                                 # already this form
    m = geom.__geo_interface__ if hasattr(geom, "__geo_interface__") else geom
    t = m.get("type")
    if t == "Polygon":
        rings = [_ring(r) for r in m["coordinates"]]
        return [rings] if rings and len(rings[0]) >= 3 else []
    if t == "MultiPolygon":
        return [p for c in m["coordinates"] for p in polygons({"type": "Polygon", "coordinates": c})]
    if t == "GeometryCollection":
        return [p for g in m["geometries"] for p in polygons(g)]
    return []


def _ring(coords) -> np.ndarray:
    r = np.asarray(coords, float).reshape(-1, 2)
    if len(r) > 1 and r[0, 0] == r[-1, 0] and r[0, 1] == r[-1, 1]:
        r = r[:-1]
    return r


def bounds(polys: list) -> tuple:
    pts = np.concatenate([p[0] for p in polys])
    return float(pts[:, 0].min()), float(pts[:, 1].min()), float(pts[:, 0].max()), float(pts[:, 1].max())


def _signed_area(r: np.ndarray) -> float:
    x, y = r[:, 0], r[:, 1]
    return 0.5 * float(np.sum(x * np.roll(y, -1) - np.roll(x, -1) * y))


def ring_edges(polys: list) -> tuple:
    """Oriented edges (peec.py:514-530): each exterior counter-clockwise, each hole clockwise, the copper on the left;
    ring by ring, polygon by polygon, as start and end points (n, 2)."""
    pa, pb = [], []
    for poly in polys:
        for h, r in enumerate(poly):
            ccw = _signed_area(r) > 0
            if ccw != (h == 0):
                r = np.concatenate([r[:1], r[1:][::-1]])
            pa.append(r)
            pb.append(np.roll(r, -1, axis=0))
    if not pa:
        return np.zeros((0, 2)), np.zeros((0, 2))
    return np.concatenate(pa), np.concatenate(pb)


def _edges(polys: list) -> tuple:
    """All ring edges, as they are stored (orientation irrelevant): ax, ay, bx, by."""
    pa, pb = [], []
    for poly in polys:
        for r in poly:
            pa.append(r)
            pb.append(np.roll(r, -1, axis=0))
    a, b = np.concatenate(pa), np.concatenate(pb)
    return a[:, 0], a[:, 1], b[:, 0], b[:, 1]


_ERRB = 3.3306690738754716e-16          # (3 + 16 eps) eps: the float orientation's error bound (Shewchuk 1997)


def orient(ax, ay, bx, by, px, py) -> np.ndarray:
    """Exact sign of (b - a) x (p - a) for float inputs: the float value where it clears its error bound, else the
    rational value."""
    ax, ay, bx, by, px, py = np.broadcast_arrays(*(np.atleast_1d(np.asarray(v, float))
                                                   for v in (ax, ay, bx, by, px, py)))
    l = (bx - ax) * (py - ay)
    r = (by - ay) * (px - ax)
    det = l - r
    out = np.sign(det).astype(np.int64)
    unsure = ~(np.abs(det) > _ERRB * (np.abs(l) + np.abs(r)))
    for i in zip(*np.nonzero(unsure)):
        F = [Fraction(float(v[i])) for v in (ax, ay, bx, by, px, py)]
        e = (F[2] - F[0]) * (F[5] - F[1]) - (F[3] - F[1]) * (F[4] - F[0])
        out[i] = (e > 0) - (e < 0)
    return out


def locate(polys: list, px, py) -> tuple:
    """(inside, on_boundary) for points: the even-odd crossing parity of a ray towards +x (half-open in y) and the
    exact on-edge test, both from orientation signs."""
    px, py = np.asarray(px, float), np.asarray(py, float)
    shape = px.shape
    px, py = px.ravel(), py.ravel()
    par = np.zeros(px.shape, bool)
    onb = np.zeros(px.shape, bool)
    if not polys or len(px) == 0:
        return par.reshape(shape), onb.reshape(shape)
    AX, AY, BX, BY = _edges(polys)
    for ax, ay, bx, by in zip(AX, AY, BX, BY):
        up = (ay <= py) & (py < by)
        down = (by <= py) & (py < ay)
        box_ = ((min(ax, bx) <= px) & (px <= max(ax, bx)) & (min(ay, by) <= py) & (py <= max(ay, by)))
        sel = np.nonzero(up | down | box_)[0]
        if len(sel) == 0:
            continue
        o = orient(ax, ay, bx, by, px[sel], py[sel])
        par[sel] ^= (up[sel] & (o > 0)) | (down[sel] & (o < 0))
        onb[sel] |= box_[sel] & (o == 0)
    return par.reshape(shape), onb.reshape(shape)


def contains(polys: list, px, py) -> np.ndarray:
    """Points in the interior (peec.py:379-385, shapely.contains_xy)."""
    inside, onb = locate(polys, px, py)
    return inside & ~onb


def covers(polys: list, px: float, py: float) -> bool:
    """A point in the region or on its boundary (peec.py:400, Geometry.covers)."""
    inside, onb = locate(polys, [px], [py])
    return bool(inside[0] or onb[0])


def _seg_dist(p, q, a, b) -> float:
    """Distance between segments p-q and a-b (0 where they cross)."""
    def side(u, v, w):
        return int(orient(u[0], u[1], v[0], v[1], w[0], w[1])[0])
    d1, d2, d3, d4 = side(a, b, p), side(a, b, q), side(p, q, a), side(p, q, b)
    if d1 * d2 < 0 and d3 * d4 < 0:
        return 0.0

    def pt(x, u, v):
        uv = v - u
        L = float(uv @ uv)
        t = 0.0 if L == 0 else min(max(float((x - u) @ uv) / L, 0.0), 1.0)
        return float(np.hypot(*(u + t * uv - x)))
    return min(pt(p, a, b), pt(q, a, b), pt(a, p, q), pt(b, p, q))


def _within(A: list, B: list, tol: float) -> bool:
    """Polygon distance <= tol (peec.py:863, dwithin): 0 when one holds a point of the other."""
    ax0, ay0, ax1, ay1 = bounds([A])
    bx0, by0, bx1, by1 = bounds([B])
    if ax0 > bx1 + tol or bx0 > ax1 + tol or ay0 > by1 + tol or by0 > ay1 + tol:
        return False
    if covers([B], *A[0][0]) or covers([A], *B[0][0]):
        return True
    ea = [(r[i], r[(i + 1) % len(r)]) for r in A for i in range(len(r))]
    eb = [(r[i], r[(i + 1) % len(r)]) for r in B for i in range(len(r))]
    return min(_seg_dist(p, q, a, b) for p, q in ea for a, b in eb) <= tol


def pieces(geom, tol: float = PIECE_TOL_MM) -> list:
    """One net's copper in connected pieces (peec.py:852-870): polygons within `tol` of each other grouped, groups
    numbered by their first polygon (connected_components' order), each a list of polygons."""
    polys = polygons(geom)
    if len(polys) <= 1:
        return [[p] for p in polys]
    n = len(polys)
    rows, cols = [], []
    for i in range(n):
        for j in range(n):
            if i == j or _within(polys[i], polys[j], tol):
                rows.append(i)
                cols.append(j)
    _, grp = connected_components(sp.coo_matrix((np.ones(len(rows)), (rows, cols)), shape=(n, n)), directed=False)
    return [[polys[i] for i in np.nonzero(grp == g)[0]] for g in range(int(grp.max()) + 1)]


# ---------------------------------------------------------------------------------------------------- the grid
@dataclass
class OGrid:
    """The fields of the reference's Grid (peec.py:197-223)."""
    x0: float
    y0: float
    d: float
    nx: int
    ny: int
    node: np.ndarray
    xb: np.ndarray
    yb: np.ndarray
    label: np.ndarray
    nets: list
    fill: Optional[np.ndarray] = None
    bridges: list = field(default_factory=list)
    wx: Optional[np.ndarray] = None
    wy: Optional[np.ndarray] = None
    ox: Optional[np.ndarray] = None
    oy: Optional[np.ndarray] = None
    area: Optional[np.ndarray] = None
    rx: Optional[np.ndarray] = None
    ry: Optional[np.ndarray] = None
    hx: Optional[np.ndarray] = None
    hy: Optional[np.ndarray] = None
    part: Optional[np.ndarray] = None       # the piece of each cell (build_grid's `part`, kept for the tests)


def coarse_from_fine(piece_polys: list, x0: float, y0: float, d: float, nx: int, ny: int, s: int):
    """peec.py:774-849: fine labels by piece, the pieces of each cell's fine copper (strips across cell borders left
    out), the largest kept where it spans s samples, crossings and offsets of the kept pieces' fine strips."""
    df = d / s
    mx, my = nx * s, ny * s
    xf = x0 - d / 2 + (np.arange(mx) + 0.5) * df
    yf = y0 - d / 2 + (np.arange(my) + 0.5) * df
    lab = -np.ones((mx, my), np.int32)
    for k, polys in enumerate(piece_polys):
        if not polys:
            continue
        gx0, gy0, gx1, gy1 = bounds(polys)
        p0, p1 = max(int(math.floor((gx0 - xf[0]) / df)) - 1, 0), min(int(math.ceil((gx1 - xf[0]) / df)) + 2, mx)
        q0, q1 = max(int(math.floor((gy0 - yf[0]) / df)) - 1, 0), min(int(math.ceil((gy1 - yf[0]) / df)) + 2, my)
        if p0 >= p1 or q0 >= q1:
            continue
        X, Y = np.meshgrid(xf[p0:p1], yf[q0:q1], indexing="ij")
        win = lab[p0:p1, q0:q1]
        win[contains(polys, X, Y) & (win < 0)] = k
    sx = (lab[:-1, :] >= 0) & (lab[:-1, :] == lab[1:, :])
    sy = (lab[:, :-1] >= 0) & (lab[:, :-1] == lab[:, 1:])
    bx = ((np.arange(mx - 1) + 1) % s == 0)[:, None]
    by = ((np.arange(my - 1) + 1) % s == 0)[None, :]
    n = mx * my
    idx = np.arange(n).reshape(mx, my)
    ix, iy = sx & ~bx, sy & ~by
    rows = np.concatenate([idx[:-1, :][ix], idx[:, :-1][iy]])
    cols = np.concatenate([idx[1:, :][ix], idx[:, 1:][iy]])
    _, comp = connected_components(sp.coo_matrix((np.ones(len(rows), np.int8), (rows, cols)), shape=(n, n)),
                                   directed=False)
    comp = comp.reshape(mx, my)
    cell = (np.arange(mx) // s)[:, None] * ny + (np.arange(my) // s)[None, :]
    has = lab >= 0
    # per cell: its pieces by size, ties to the piece met first in row-major fine order
    best = {}
    size = np.bincount(comp[has], minlength=n)
    for c in np.unique(comp[has]):
        ps, qs = np.nonzero((comp == c) & has)
        ce = int(cell[ps[0], qs[0]])
        key = (-int(size[c]), int(idx[ps[0], qs[0]]))
        if ce not in best or key < best[ce][0]:
            best[ce] = (key, int(c), int(lab[ps[0], qs[0]]))
    label = -np.ones(nx * ny, int)
    fill = np.zeros(nx * ny)
    kept = np.zeros(n, bool)
    for ce, (key, c, piece) in best.items():
        if -key[0] >= s:
            label[ce] = piece
            fill[ce] = -key[0] / (s * s)
            kept[c] = True
    on = has & kept[comp]
    cross_x, off_x = np.zeros((nx - 1, ny)), np.zeros((nx - 1, ny))
    pi, qi = np.nonzero(sx & bx & on[:-1, :] & on[1:, :])
    np.add.at(cross_x, (pi // s, qi // s), 1.0 / s)
    np.add.at(off_x, (pi // s, qi // s), ((qi % s + 0.5) / s - 0.5) / s)
    cross_y, off_y = np.zeros((nx, ny - 1)), np.zeros((nx, ny - 1))
    pj, qj = np.nonzero(sy & by & on[:, :-1] & on[:, 1:])
    np.add.at(cross_y, (pj // s, qj // s), 1.0 / s)
    np.add.at(off_y, (pj // s, qj // s), ((pj % s + 0.5) / s - 0.5) / s)
    with np.errstate(invalid="ignore", divide="ignore"):
        off_x = np.where(cross_x > 0, off_x / cross_x, 0.0)
        off_y = np.where(cross_y > 0, off_y / cross_y, 0.0)
    return label.reshape(nx, ny), cross_x, cross_y, fill.reshape(nx, ny), off_x, off_y, lab


def line_cover(ua, va, ub, vb, u0: float, du: float, n_lines: int, v0: float, d: float, n_cells: int):
    """peec.py:533-583: on each line u = u0 + m du, the covered length of each interval [v0 + (c - 1/2) d,
    v0 + (c + 1/2) d], from the even-odd spans between the line's sorted crossings (a line with an odd count none).
    Written line by line and span by span; the arithmetic of each span's pieces is the reference's."""
    out = np.zeros((max(n_lines, 0), max(n_cells, 0)))
    if len(ua) == 0 or n_lines <= 0 or n_cells <= 0:
        return out
    lo, hi = np.minimum(ua, ub), np.maximum(ua, ub)
    m0 = np.clip(np.ceil((lo - u0) / du), 0, n_lines).astype(np.int64)
    m1 = np.clip(np.ceil((hi - u0) / du), 0, n_lines).astype(np.int64)
    cross = [[] for _ in range(n_lines)]
    for e in np.nonzero(m1 > m0)[0]:
        for m in range(m0[e], m1[e]):
            u = u0 + m * du
            cross[m].append(va[e] + (u - ua[e]) * (vb[e] - va[e]) / (ub[e] - ua[e]))
    vlo, top = v0 - d / 2, n_cells * d
    for m in range(n_lines):
        v = sorted(cross[m])
        if len(v) % 2:
            continue
        full = np.zeros(n_cells + 1)
        for a, b in zip(v[0::2], v[1::2]):
            a, b = min(max(a - vlo, 0.0), top), min(max(b - vlo, 0.0), top)
            if not b > a:
                continue
            ca, cb = min(int(a / d), n_cells - 1), min(int(b / d), n_cells - 1)
            if ca == cb:
                out[m, ca] += b - a
            else:
                out[m, ca] += (ca + 1) * d - a
                out[m, cb] += b - cb * d
                full[ca + 1] += d
                full[cb] -= d
        out[m] += np.cumsum(full)[:n_cells]
    return out


def cell_areas(pa, pb, xl: float, yl: float, d: float, wi: int, wj: int, right):
    """peec.py:586-625: each cell's copper area by Green's theorem, the oriented edges cut at the grid lines, each
    piece adding (mean x - the cell's left) dy to the cell holding its midpoint, plus d times the right border's cover."""
    out = np.zeros((wi, wj))
    for (xa, ya), (xb, yb) in zip(pa, pb):
        ts = [0.0, 1.0]
        for a, b, origin, n in ((xa, xb, xl, wi + 1), (ya, yb, yl, wj + 1)):
            lo, hi = min(a, b), max(a, b)
            m0 = int(np.clip(np.ceil((lo - origin) / d), 0, n))
            m1 = int(np.clip(np.ceil((hi - origin) / d), 0, n))
            ts += [(origin + m * d - a) / (b - a) for m in range(m0, m1)]
        ts.sort()
        for t0, t1 in zip(ts[:-1], ts[1:]):
            x0p, y0p = xa + t0 * (xb - xa), ya + t0 * (yb - ya)
            x1p, y1p = xa + t1 * (xb - xa), ya + t1 * (yb - ya)
            i = int(np.floor(((x0p + x1p) / 2 - xl) / d))
            j = int(np.floor(((y0p + y1p) / 2 - yl) / d))
            if 0 <= i < wi and 0 <= j < wj:
                out[i, j] += ((x0p + x1p) / 2 - (xl + i * d)) * (y1p - y0p)
    return out + d * right


def adopt(cov, has, steps, reach: int = 1):
    """peec.py:628-656, cell by cell: copper where its net has no strip (node) moved to the nearest that has one along
    `steps`, up to `reach` steps away (ties: the neighbour with more copper of its own; then the first step)."""
    own = np.where(has, cov, 0.0)
    orphan = np.where(has, 0.0, cov)
    out = own.copy()
    ni, nj = cov.shape
    for k in range(1, reach + 1):
        if not np.any(orphan > 0):
            break
        moved = []
        for i in range(ni):
            for j in range(nj):
                if not orphan[i, j] > 0:
                    continue
                best, pick = -1.0, None
                for di, dj in steps:
                    a, b = i + di * k, j + dj * k
                    if 0 <= a < ni and 0 <= b < nj and has[a, b] and own[a, b] > best:
                        best, pick = own[a, b], (a, b)
                if pick is not None:
                    moved.append((i, j, pick))
        for i, j, (a, b) in moved:
            out[a, b] += orphan[i, j]
        for i, j, _ in moved:
            orphan[i, j] = 0.0
    return out


def series_factor(face, samples, has, axis: int):
    """peec.py:659-705, run by run: a band of at most BAND_STRIPS side-by-side strips with no other strip of its net
    within 2 ADOPT_FACES lanes takes the harmonic mean of its total width at the sample lines over its total at the
    border (1 where a sample line has none of its copper)."""
    h = has if axis == 1 else has.T
    f = face if axis == 1 else face.T
    rows = [s if axis == 1 else s.T for s in samples]
    out = np.ones(h.shape)
    n0, n1 = h.shape
    far = n1 + 2 * ADOPT_FACES + 2
    for r in range(n0):
        runs, c = [], 0
        while c < n1:
            if h[r, c]:
                c1 = c
                while c1 < n1 and h[r, c1]:
                    c1 += 1
                runs.append((c, c1))
                c = c1
            else:
                c += 1
        for q, (c0, c1) in enumerate(runs):
            before = c0 - runs[q - 1][1] if q > 0 else far
            after = runs[q + 1][0] - c1 if q + 1 < len(runs) else far
            total = float(np.sum(f[r, c0:c1]))
            alone = (c1 - c0 <= BAND_STRIPS) and before > 2 * ADOPT_FACES and after > 2 * ADOPT_FACES
            t = [float(np.sum(s[r, c0:c1])) for s in rows]
            ok = all(v > 0 for v in t)
            if alone and ok and total > 0:
                harm = len(rows) / sum(1.0 / v for v in t)
                out[r, c0:c1] = harm / total
    return out if axis == 1 else out.T


def strip_widths(exists, cross, offset, forced, d: float, exact=None):
    """peec.py:482-511."""
    w = np.where(forced > 0, np.maximum(cross, np.minimum(forced / d, 1.0)), cross)
    w = np.where(w > 0, np.minimum(w, 1.0), 1.0)
    if exact is not None:
        w = np.where(exact > 0, np.minimum(exact, 3.0), w)
    o = np.where(cross > 0, offset, 0.0)
    lim = np.maximum((1.0 - w) / 2, 0.0)
    return np.where(exists, w, 0.0), np.where(exists, np.clip(o, -lim, lim), 0.0)


def exact_copper(geoms: list, label, xb, yb, x0: float, y0: float, d: float):
    """peec.py:708-771."""
    nx, ny = label.shape
    ex_x, ex_y, area = np.zeros((nx - 1, ny)), np.zeros((nx, ny - 1)), np.zeros((nx, ny))
    kx, ky = np.ones((nx - 1, ny)), np.ones((nx, ny - 1))
    for k, g in enumerate(geoms):
        polys = polygons(g)
        if not polys:
            continue
        pa, pb = ring_edges(polys)
        gx0, gy0, gx1, gy1 = bounds(polys)
        i0, i1 = max(int(math.floor((gx0 - x0) / d)) - 2, 0), min(int(math.ceil((gx1 - x0) / d)) + 3, nx)
        j0, j1 = max(int(math.floor((gy0 - y0) / d)) - 2, 0), min(int(math.ceil((gy1 - y0) / d)) + 3, ny)
        wi, wj = i1 - i0, j1 - j0
        if wi < 2 or wj < 2:
            continue
        xa, ya, xe, ye = pa[:, 0], pa[:, 1], pb[:, 0], pb[:, 1]
        cv = line_cover(xa, ya, xe, ye, x0 + (i0 - 0.5) * d, d, wi + 1, y0 + j0 * d, d, wj)
        ch = line_cover(ya, xa, ye, xe, y0 + (j0 + 0.5) * d, d, wj - 1, x0 + i0 * d, d, wi).T
        own = label[i0:i1, j0:j1] == k
        has_x = xb[i0:i1 - 1, j0:j1] & own[:-1, :]
        has_y = yb[i0:i1, j0:j1 - 1] & own[:, :-1]
        face_x = adopt(cv[1:wi, :] / d, has_x, ((0, 1), (0, -1)), ADOPT_FACES)
        face_y = adopt(ch / d, has_y, ((1, 0), (-1, 0)), ADOPT_FACES)
        ex_x[i0:i1 - 1, j0:j1] += face_x
        ex_y[i0:i1, j0:j1 - 1] += face_y
        q_n = STRIP_SAMPLES
        sx = (line_cover(xa, ya, xe, ye, x0 + (i0 + 0.5 / q_n) * d, d / q_n, (wi - 1) * q_n, y0 + j0 * d, d, wj)
              / d).reshape(wi - 1, q_n, wj)
        sy = (line_cover(ya, xa, ye, xe, y0 + (j0 + 0.5 / q_n) * d, d / q_n, (wj - 1) * q_n, x0 + i0 * d, d, wi)
              / d).reshape(wj - 1, q_n, wi)
        along_x = [adopt(sx[:, q, :], has_x, ((0, 1), (0, -1)), ADOPT_FACES) for q in range(q_n)]
        along_y = [adopt(sy[:, q, :].T, has_y, ((1, 0), (-1, 0)), ADOPT_FACES) for q in range(q_n)]
        win_x, win_y = kx[i0:i1 - 1, j0:j1], ky[i0:i1, j0:j1 - 1]
        win_x[has_x] = series_factor(face_x, along_x, has_x, 1)[has_x]
        win_y[has_y] = series_factor(face_y, along_y, has_y, 0)[has_y]
        cell = cell_areas(pa, pb, x0 + (i0 - 0.5) * d, y0 + (j0 - 0.5) * d, d, wi, wj, cv[1:wi + 1, :]) / (d * d)
        area[i0:i1, j0:j1] += adopt(np.clip(cell, 0.0, 1.0), own, ((1, 0), (-1, 0), (0, 1), (0, -1)), ADOPT_CELLS)
    return ex_x, ex_y, area, kx, ky


def heat_points(geoms: list, label, x0: float, y0: float, d: float):
    """peec.py:452-479: from each node outside its net's copper (not in it, not on its border) to the nearest point
    of that copper, a hair (1e-4 mm) past it."""
    nx, ny = label.shape
    hx, hy = np.zeros((nx, ny)), np.zeros((nx, ny))
    for k, g in enumerate(geoms):
        polys = polygons(g)
        if not polys:
            continue
        ii, jj = np.nonzero(label == k)
        if len(ii) == 0:
            continue
        px, py = x0 + ii * d, y0 + jj * d
        inside, onb = locate(polys, px, py)
        out = ~(inside | onb)
        if not out.any():
            continue
        ii, jj, px, py = ii[out], jj[out], px[out], py[out]
        AX, AY, BX, BY = _edges(polys)
        ex, ey = np.zeros(len(px)), np.zeros(len(px))
        for n in range(len(px)):
            ux, uy = BX - AX, BY - AY
            L = ux * ux + uy * uy
            t = np.clip(np.where(L > 0, ((px[n] - AX) * ux + (py[n] - AY) * uy) / np.where(L > 0, L, 1.0), 0.0),
                        0.0, 1.0)
            qx, qy = AX + t * ux, AY + t * uy
            e = int(np.argmin((qx - px[n]) ** 2 + (qy - py[n]) ** 2))
            ex[n], ey[n] = qx[e], qy[e]
        vx, vy = ex - px, ey - py
        nrm = np.hypot(vx, vy)
        step = np.where(nrm > 0, 1e-4 / np.maximum(nrm, 1e-300), 0.0)
        hx[ii, jj] = vx * (1 + step)
        hy[ii, jj] = vy * (1 + step)
    return hx, hy


def line_point(pts) -> tuple:
    """peec.py:887-890."""
    (ax, ay), (bx, by) = pts[0], pts[1] if len(pts) > 1 else pts[0]
    return (ax + bx) / 2, (ay + by) / 2


def rasterise_lines(k: int, polylines, x0: float, y0: float, d: float, label, seeded, forced_x, forced_y, part=None,
                    owners=None) -> list:
    """peec.py:893-983: each centreline walked through the grid cell by cell (the nearest own cell of the four around
    each sample 0.4 pitch apart, else the nearest free one), consecutive cells joined by the L path through fewer
    foreign cells; free cells claimed, strips forced to the widest track, gaps bridged."""
    nx, ny = label.shape
    bridges = []
    for n_line, (pts, width) in enumerate(polylines):
        pid = owners[n_line] if owners is not None else None

        def on_grid(c):
            return 0 <= c[0] < nx and 0 <= c[1] < ny

        def own(c):
            return on_grid(c) and label[c] == k and (part is None or pid is None or part[c] == pid)

        def free(c):
            return on_grid(c) and (label[c] == -1 or own(c))

        def l_path(a, b):
            sx, sy = (b[0] > a[0]) - (b[0] < a[0]), (b[1] > a[1]) - (b[1] < a[1])
            xs_ = [(a[0] + sx * t, a[1]) for t in range(1, abs(b[0] - a[0]) + 1)]
            ys_ = [(a[0], a[1] + sy * t) for t in range(1, abs(b[1] - a[1]) + 1)]
            xf = xs_ + [(b[0], c[1]) for c in ys_]
            yf = ys_ + [(c[0], b[1]) for c in xs_]
            bad_x, bad_y = sum(not free(c) for c in xf), sum(not free(c) for c in yf)
            return xf if bad_x <= bad_y else yf

        cells = []
        s0 = 0.0
        for (ax, ay), (bx, by) in zip(pts, pts[1:]):
            seg = math.hypot(bx - ax, by - ay)
            n = max(int(math.ceil(seg / (0.4 * d))), 1)
            for tt in np.linspace(0.0, 1.0, n + 1):
                u, v = (ax + tt * (bx - ax) - x0) / d, (ay + tt * (by - ay) - y0) / d
                i0, j0 = math.floor(u), math.floor(v)
                near = sorted([(i0, j0), (i0, j0 + 1), (i0 + 1, j0), (i0 + 1, j0 + 1)],
                              key=lambda c: (c[0] - u) ** 2 + (c[1] - v) ** 2)
                c = next((c for c in near if own(c)), None) or next((c for c in near if free(c)), near[0])
                s = s0 + tt * seg
                if not cells:
                    cells.append((c[0], c[1], s))
                elif cells[-1][:2] != c:
                    cells += [(q[0], q[1], s) for q in l_path(cells[-1][:2], c)]
            s0 += seg
        prev = last = None
        for i, j, s in cells:
            if not (0 <= i < nx and 0 <= j < ny):
                prev = last = None
                continue
            if label[i, j] < 0:
                label[i, j] = k
                if part is not None and pid is not None:
                    part[i, j] = pid
                seeded[i, j] = True
            if not own((i, j)):
                prev = None
                continue
            if prev is not None:
                pi, pj = prev
                tw = max(float(width), 1e-6)
                if pj == j and abs(pi - i) == 1:
                    forced_x[min(pi, i), j] = max(forced_x[min(pi, i), j], tw)
                elif pi == i and abs(pj - j) == 1:
                    forced_y[i, min(pj, j)] = max(forced_y[i, min(pj, j)], tw)
            elif last is not None and last[:2] != (i, j):
                bridges.append((last[0], last[1], i, j, max(s - last[2], d), float(width)))
            prev = (i, j)
            last = (i, j, s)
    return bridges


def build_grid(geom, bbox, d: float, seeds=None, lines=None, sub: int = 1) -> OGrid:
    """peec.py:329-449."""
    x0, y0, x1, y1 = bbox
    nx = max(int(math.ceil((x1 - x0) / d)) + 1, 3)
    ny = max(int(math.ceil((y1 - y0) / d)) + 1, 3)
    xs = x0 + np.arange(nx) * d
    ys = y0 + np.arange(ny) * d
    by_net = isinstance(geom, dict) and not isinstance(geom.get("type"), str)    # {net: geometry}, not GeoJSON
    nets = list(geom.keys()) if by_net else [""]
    geoms = [polygons(g) for g in (geom.values() if by_net else [geom])]
    geom_list = list(geoms)
    fill = None
    fine = sub > 1
    plist: list = []
    if fine:
        for k, g in enumerate(geoms):
            for p in pieces(g):
                plist.append((k, p))
        part, cross_x, cross_y, fill, off_x, off_y, _ = coarse_from_fine([p for _, p in plist], x0, y0, d, nx, ny,
                                                                         sub)
        net_of = np.array([k for k, _ in plist] + [-1], int)
        label = net_of[part]
        mids_x = np.where(cross_x > 0, label[:-1, :], -1)
        mids_y = np.where(cross_y > 0, label[:, :-1], -1)
        geoms = []
    else:
        label = -np.ones((nx, ny), int)
        mids_x = np.full((nx - 1, ny), -1)
        mids_y = np.full((nx, ny - 1), -1)
        part = None
    for k, g in enumerate(geoms):
        if not g:
            continue
        gx0, gy0, gx1, gy1 = bounds(g)
        i0, i1 = max(int(math.floor((gx0 - x0) / d)) - 1, 0), min(int(math.ceil((gx1 - x0) / d)) + 2, nx)
        j0, j1 = max(int(math.floor((gy0 - y0) / d)) - 1, 0), min(int(math.ceil((gy1 - y0) / d)) + 2, ny)
        if i0 >= i1 or j0 >= j1:
            continue
        X, Y = np.meshgrid(xs[i0:i1], ys[j0:j1], indexing="ij")
        w = label[i0:i1, j0:j1]
        w[contains(g, X, Y) & (w < 0)] = k
        mx = mids_x[i0:min(i1, nx - 1), j0:j1]
        mx[contains(g, X[: mx.shape[0]] + d / 2, Y[: mx.shape[0]]) & (mx < 0)] = k
        my = mids_y[i0:i1, j0:min(j1, ny - 1)]
        my[contains(g, X[:, : my.shape[1]], Y[:, : my.shape[1]] + d / 2) & (my < 0)] = k
    if part is None:
        part = label

    def piece_at(k, px, py):
        if not plist:
            return k
        for c, (kk, p) in enumerate(plist):
            if kk == k and p is not None and covers(p, px, py):
                return c
        plist.append((k, None))
        return len(plist) - 1

    seeded = np.zeros((nx, ny), bool)
    for net, points in (seeds or {}).items():
        if net not in nets:
            continue
        k = nets.index(net)
        for (px, py) in points:
            i, j = int(round((px - x0) / d)), int(round((py - y0) / d))
            if 0 <= i < nx and 0 <= j < ny and label[i, j] < 0:
                label[i, j] = k
                part[i, j] = piece_at(k, px, py)
                seeded[i, j] = True
    forced_x = np.zeros((nx - 1, ny))
    forced_y = np.zeros((nx, ny - 1))
    bridges = []
    for net, polylines in (lines or {}).items():
        if net not in nets:
            continue
        k = nets.index(net)
        owners = [piece_at(k, *line_point(pts)) if plist else None for pts, _ in polylines]
        bridges += rasterise_lines(k, polylines, x0, y0, d, label, seeded, forced_x, forced_y, part, owners)
    xb = (label[:-1, :] >= 0) & (label[:-1, :] == label[1:, :]) & (part[:-1, :] == part[1:, :]) & (
        (mids_x == label[:-1, :]) | seeded[:-1, :] | seeded[1:, :] | (forced_x > 0))
    yb = (label[:, :-1] >= 0) & (label[:, :-1] == label[:, 1:]) & (part[:, :-1] == part[:, 1:]) & (
        (mids_y == label[:, :-1]) | seeded[:, :-1] | seeded[:, 1:] | (forced_y > 0))
    g = OGrid(x0, y0, d, nx, ny, label >= 0, xb, yb, label, nets, fill, bridges, part=part)
    if fine:
        ex_x, ex_y, area, kx, ky = exact_copper(geom_list, label, xb, yb, x0, y0, d)
        g.wx, g.ox = strip_widths(xb, cross_x, off_x, forced_x, d, ex_x)
        g.wy, g.oy = strip_widths(yb, cross_y, off_y, forced_y, d, ex_y)
        g.area = area
        g.rx, g.ry = g.wx * kx, g.wy * ky
        g.hx, g.hy = heat_points(geom_list, label, x0, y0, d)
    return g


def copper_area(g) -> np.ndarray:
    """peec.py:259-313 (Grid.copper_area), node by node."""
    out = g.node.astype(float)
    if g.wx is None or g.wy is None:
        return out
    nx, ny = g.node.shape
    sx, nxs, sy, nys = (np.zeros((nx, ny)) for _ in range(4))
    for i in range(nx - 1):
        for j in range(ny):
            if g.xb[i, j]:
                sx[i, j] += g.wx[i, j]
                sx[i + 1, j] += g.wx[i, j]
                nxs[i, j] += 1
                nxs[i + 1, j] += 1
    for i in range(nx):
        for j in range(ny - 1):
            if g.yb[i, j]:
                sy[i, j] += g.wy[i, j]
                sy[i, j + 1] += g.wy[i, j]
                nys[i, j] += 1
                nys[i, j + 1] += 1
    n = nxs + nys
    mean_width = np.where(n > 0, (sx + sy) / np.maximum(n, 1), out)
    if g.area is None:
        return mean_width
    x_run, y_run = nxs == 2, nys == 2
    run_width = np.where(x_run & ~y_run, sx / np.maximum(nxs, 1),
                         np.where(y_run & ~x_run, sy / np.maximum(nys, 1), mean_width))
    area = np.where(g.node, g.area, 0.0)
    width = np.where(g.node, run_width, 0.0)
    lab = np.where(g.node, g.label if g.label is not None else 0, -1)
    sum_a, sum_w = np.zeros((nx, ny)), np.zeros((nx, ny))
    for i in range(nx):
        for j in range(ny):
            if lab[i, j] < 0:
                continue
            for a in range(max(i - 1, 0), min(i + 2, nx)):
                for b in range(max(j - 1, 0), min(j + 2, ny)):
                    if lab[a, b] == lab[i, j]:
                        sum_a[i, j] += area[a, b]
                        sum_w[i, j] += width[a, b]
    shared = np.where(sum_w > 0, width * sum_a / np.maximum(sum_w, 1e-12), 0.0)
    return np.where(g.node, np.where(shared > 0, shared, mean_width), 0.0)
