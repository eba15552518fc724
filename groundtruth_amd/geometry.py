"""Copper geometry for the rasteriser: polygons with holes, read from GeoJSON-like mappings or any object with
`__geo_interface__` (a shapely geometry has one), so boards drawn with shapely pass straight in and boards built
without it use the small constructors here (`box`, `disc`, `Polygon`, `MultiPolygon`).

A geometry becomes a list of polygons, each a list of rings (exterior first, then holes) as (n, 2) float arrays
without the closing point. Lines and points are left out, as the reference's `_polygons` leaves them
(`src/groundtruth/solver/peec.py:873-883`).
"""
from __future__ import annotations

import math
from typing import Iterable, Sequence

import numpy as np

PIECE_TOL_MM = 0.005        # peec.py:106: one net's copper within this of itself is one piece


class Polygon:
    """A polygon (mm): an exterior ring and holes, each a sequence of (x, y)."""

    def __init__(self, exterior: Sequence, holes: Iterable[Sequence] = ()):
        self.rings = [_ring(exterior)] + [_ring(h) for h in holes]

    @property
    def __geo_interface__(self) -> dict:
        return {"type": "Polygon", "coordinates": [_closed(r) for r in self.rings]}

    @property
    def area(self) -> float:
        return abs(_signed_area(self.rings[0])) - sum(abs(_signed_area(h)) for h in self.rings[1:])


class MultiPolygon:
    """Several polygons (mm) of one net."""

    def __init__(self, polys: Iterable):
        self.polys = [p if isinstance(p, Polygon) else Polygon(*p) for p in polys]

    @property
    def __geo_interface__(self) -> dict:
        return {"type": "MultiPolygon", "coordinates": [[_closed(r) for r in p.rings] for p in self.polys]}

    @property
    def area(self) -> float:
        return sum(p.area for p in self.polys)


def box(x0: float, y0: float, x1: float, y1: float, holes: Iterable[Sequence] = ()) -> Polygon:
    """The rectangle [x0, x1] x [y0, y1] (mm), optionally with holes."""
    return Polygon([(x0, y0), (x1, y0), (x1, y1), (x0, y1)], holes)


def disc(x: float, y: float, r: float, quad_segs: int = 16) -> np.ndarray:
    """A circle of radius r about (x, y) as a ring of 4 quad_segs vertices (shapely's `buffer` resolution), counter-
    clockwise from angle 0: an outline, or a hole (antipad) for `box` / `Polygon`."""
    a = 2 * math.pi * np.arange(4 * quad_segs) / (4 * quad_segs)
    return np.column_stack([x + r * np.cos(a), y + r * np.sin(a)])


def _ring(coords) -> np.ndarray:
    r = np.asarray(coords, float).reshape(-1, 2)
    if len(r) > 1 and r[0, 0] == r[-1, 0] and r[0, 1] == r[-1, 1]:
        r = r[:-1]
    return r


def _closed(r: np.ndarray) -> list:
    return [tuple(p) for p in r] + [tuple(r[0])]


def _signed_area(r: np.ndarray) -> float:
    x, y = r[:, 0], r[:, 1]
    return 0.5 * float(np.dot(x, np.roll(y, -1)) - np.dot(np.roll(x, -1), y))


def polygons(geom) -> list:
    """The polygons of a geometry as lists of rings (module docstring)."""
    if geom is None:
        return []
    if isinstance(geom, Polygon):
        return [list(geom.rings)] if len(geom.rings[0]) >= 3 else []
    if isinstance(geom, MultiPolygon):
        return [p for q in geom.polys for p in polygons(q)]
    m = geom.__geo_interface__ if hasattr(geom, "__geo_interface__") else geom
    if not isinstance(m, dict):
        raise TypeError(f"not a geometry: {type(geom).__name__}")
    kind = m.get("type")
    if kind == "Polygon":
        rings = [_ring(r) for r in m.get("coordinates", [])]
        return [rings] if rings and len(rings[0]) >= 3 else []
    if kind == "MultiPolygon":
        return [p for c in m.get("coordinates", []) for p in polygons({"type": "Polygon", "coordinates": c})]
    if kind == "GeometryCollection":
        return [p for g in m.get("geometries", []) for p in polygons(g)]
    return []                                   # lines, points


def is_geometry(obj) -> bool:
    """A single geometry (not a {net: geometry} mapping)."""
    if isinstance(obj, (Polygon, MultiPolygon)) or hasattr(obj, "__geo_interface__"):
        return True
    return isinstance(obj, dict) and isinstance(obj.get("type"), str)


def bounds(polys: list) -> tuple:
    """(x0, y0, x1, y1) of the polygons' exteriors."""
    pts = np.concatenate([p[0] for p in polys])
    return float(pts[:, 0].min()), float(pts[:, 1].min()), float(pts[:, 0].max()), float(pts[:, 1].max())


def _ring_rows(rings: list) -> np.ndarray:
    """(ax, ay, bx, by) rows of rings' edges, ring after ring, each ring closed back to its first point."""
    if not rings:
        return np.zeros((0, 4))
    pts = np.concatenate(rings)
    n = np.array([len(r) for r in rings])
    start = np.repeat(np.cumsum(n) - n, n)
    nxt = np.arange(len(pts)) + 1
    ends = np.cumsum(n) - 1
    nxt[ends] = start[ends]
    return np.ascontiguousarray(np.hstack([pts, pts[nxt]]))


def edges(polys: list) -> np.ndarray:
    """Every ring edge as (ax, ay, bx, by) rows, ring by ring, as stored."""
    return _ring_rows([r for p in polys for r in p])


def oriented_edges(polys: list) -> np.ndarray:
    """Ring edges with each exterior counter-clockwise and each hole clockwise (the copper on every edge's left;
    peec.py:514-530), as (ax, ay, bx, by) rows."""
    rings = []
    for p in polys:
        for h, r in enumerate(p):
            if (_signed_area(r) > 0) != (h == 0):
                r = np.concatenate([r[:1], r[:0:-1]])
            rings.append(r)
    return _ring_rows(rings)


def _segment_distances(P: np.ndarray, Q: np.ndarray) -> float:
    """The least distance between any edge of P and any edge of Q ((n, 4) rows), 0 where two cross."""
    best = math.inf
    for s in range(0, len(P), 512):
        p = P[s:s + 512, None, :]
        q = Q[None, :, :]
        # crossing: each segment's ends on opposite sides of the other
        def side(a, b, c):
            return (b[..., 0] - a[..., 0]) * (c[..., 1] - a[..., 1]) - (b[..., 1] - a[..., 1]) * (c[..., 0] - a[..., 0])
        pa, pb, qa, qb = p[..., :2], p[..., 2:], q[..., :2], q[..., 2:]
        cross = (side(qa, qb, pa) * side(qa, qb, pb) < 0) & (side(pa, pb, qa) * side(pa, pb, qb) < 0)
        if cross.any():
            return 0.0

        def point_seg(x, a, b):
            u = b - a
            L = np.sum(u * u, axis=-1)
            t = np.clip(np.sum((x - a) * u, axis=-1) / np.where(L > 0, L, 1.0), 0.0, 1.0)
            return np.hypot(*np.moveaxis(a + t[..., None] * u - x, -1, 0))
        d = np.minimum.reduce([point_seg(pa, qa, qb), point_seg(pb, qa, qb), point_seg(qa, pa, pb),
                               point_seg(qb, pa, pb)])
        best = min(best, float(d.min()))
    return best


def pieces(polys: list, inside, tol: float = PIECE_TOL_MM) -> list:
    """One net's copper in connected pieces (peec.py:852-870): polygons grouped where their distance is at most `tol`
    (touching or overlapping polygons, or one holding the other, are at distance 0); groups numbered by their first
    polygon, each a list of polygons. `inside(polys, points)` says which points lie in polygons' copper or on its
    border (the device's point test)."""
    n = len(polys)
    if n <= 1:
        return [[p] for p in polys]
    box_ = np.array([bounds([p]) for p in polys])
    parent = list(range(n))

    def find(i):
        while parent[i] != i:
            parent[i] = parent[parent[i]]
            i = parent[i]
        return i

    E = [edges([p]) for p in polys]
    for i in range(n):
        near = np.nonzero((box_[:, 0] <= box_[i, 2] + tol) & (box_[i, 0] <= box_[:, 2] + tol)
                          & (box_[:, 1] <= box_[i, 3] + tol) & (box_[i, 1] <= box_[:, 3] + tol))[0]
        for j in near[near > i]:
            if find(i) == find(j):
                continue
            hit = (inside([polys[j]], polys[i][0][:1]).any() or inside([polys[i]], polys[j][0][:1]).any()
                   or _segment_distances(E[i], E[j]) <= tol)
            if hit:
                parent[find(j)] = find(i)
    order: dict = {}
    for i in range(n):
        order.setdefault(find(i), []).append(polys[i])
    return list(order.values())
