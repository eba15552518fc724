"""Copper rasterisation: `build_grid` turns each net's copper polygons into a `Grid` (nodes, strips, nets, sub-grid
fill, strip widths and offsets, copper areas, series factors, heat points, track bridges), the reference's
`build_grid` (`src/groundtruth/solver/peec.py:329-449`) with its helpers (`peec.py:452-983`).

The work over the board's area runs on the device (libgtpeec, `gt_raster_*`): point labels of the node, midpoint and
sub-grid lattices by exact scanline tests, the sub-grid's cells and crossings, the line covers that measure the
strips' widths, and the point queries of seeds, tracks and heat points. The host keeps what is sequential or
proportional to the copper's outline: the pieces of each net, seeds, the track walk, the cell areas by Green's
theorem, and the adoption of slivers and the series factors over the measured arrays.
"""
from __future__ import annotations

import math
from typing import Optional

import numpy as np

from . import _lib
from . import geometry as geo
from .problem import Grid

STRIP_SAMPLES = 6           # peec.py:107: lines along each strip where its copper's width is measured
BAND_STRIPS = 4             # peec.py:108: a narrow conductor is a band of at most this many side-by-side strips
ADOPT_FACES = 3             # peec.py:110: sliver copper on a border line goes to a strip this many cells along it
ADOPT_CELLS = 2             # peec.py:111: a cell's copper without a node goes to a node this many cells away


class _Device:
    """The device calls of one build (a context of the library)."""

    def __init__(self, device: int):
        self.ctx = _lib.Context.get(device)

    def label(self, groups: list, windows: np.ndarray, xs: np.ndarray, ys: np.ndarray, lab: np.ndarray):
        eptr, E = _pack(groups)
        win = np.ascontiguousarray(windows, np.int32)
        _lib.check(_lib.lib().gt_raster_label(self.ctx.handle, len(groups), _p(eptr, np.int64), _lib.ptr(E),
                                              _p(win, np.int32), len(xs), _lib.ptr(xs), len(ys), _lib.ptr(ys),
                                              _p(lab, np.int32)), "gt_raster_label")

    def fine(self, groups: list, windows: np.ndarray, nx: int, ny: int, s: int, xf: np.ndarray, yf: np.ndarray):
        eptr, E = _pack(groups)
        win = np.ascontiguousarray(windows, np.int32)
        part = np.empty((nx, ny), np.int32)
        fill = np.empty((nx, ny))
        cx, ox = np.empty((nx - 1, ny)), np.empty((nx - 1, ny))
        cy, oy = np.empty((nx, ny - 1)), np.empty((nx, ny - 1))
        _lib.check(_lib.lib().gt_raster_fine(self.ctx.handle, len(groups), _p(eptr, np.int64), _lib.ptr(E),
                                             _p(win, np.int32), nx, ny, s, _lib.ptr(xf), _lib.ptr(yf),
                                             _p(part, np.int32), _lib.ptr(fill), _lib.ptr(cx), _lib.ptr(ox),
                                             _lib.ptr(cy), _lib.ptr(oy), None), "gt_raster_fine")
        return part.astype(int), cx, cy, fill, ox, oy

    def cover(self, E: np.ndarray, u0: float, du: float, n_lines: int, v0: float, d: float, n_cells: int):
        out = np.zeros((max(n_lines, 0), max(n_cells, 0)))
        if len(E) == 0 or n_lines <= 0 or n_cells <= 0:
            return out
        E = np.ascontiguousarray(E, np.float64)
        _lib.check(_lib.lib().gt_raster_cover(self.ctx.handle, len(E), _lib.ptr(E), float(u0), float(du),
                                              int(n_lines), float(v0), float(d), int(n_cells), _lib.ptr(out)),
                   "gt_raster_cover")
        return out

    def locate(self, polys: list, pts: np.ndarray):
        """(inside, on the border) of points against the polygons' copper."""
        E = geo.edges(polys)
        pts = np.ascontiguousarray(np.asarray(pts, float).reshape(-1, 2))
        ins, onb = np.zeros(len(pts), np.int8), np.zeros(len(pts), np.int8)
        _lib.check(_lib.lib().gt_raster_points(self.ctx.handle, len(E), _lib.ptr(E), len(pts), _lib.ptr(pts),
                                               _p(ins, np.int8), _p(onb, np.int8), None), "gt_raster_points")
        return ins.astype(bool), onb.astype(bool)

    def state(self, polys: list, window, xs: np.ndarray, ys: np.ndarray) -> np.ndarray:
        """Lattice points of a window against the polygons: 1 inside, 2 on an edge, 0 outside."""
        E = geo.edges(polys)
        win = np.ascontiguousarray(window, np.int32)
        out = np.zeros((len(xs), len(ys)), np.int8)
        _lib.check(_lib.lib().gt_raster_state(self.ctx.handle, len(E), _lib.ptr(E), _p(win, np.int32), len(xs),
                                              _lib.ptr(xs), len(ys), _lib.ptr(ys), _p(out, np.int8)),
                   "gt_raster_state")
        return out

    def nearest(self, polys: list, pts: np.ndarray) -> np.ndarray:
        E = geo.edges(polys)
        pts = np.ascontiguousarray(np.asarray(pts, float).reshape(-1, 2))
        out = np.zeros_like(pts)
        _lib.check(_lib.lib().gt_raster_points(self.ctx.handle, len(E), _lib.ptr(E), len(pts), _lib.ptr(pts), None,
                                               None, _lib.ptr(out)), "gt_raster_points")
        return out


def _p(a: np.ndarray, dtype):
    import ctypes as C
    ct = {np.int32: C.c_int32, np.int64: C.c_int64, np.int8: C.c_int8}[dtype]
    assert a.dtype == dtype and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.POINTER(ct))


def _pack(groups: list):
    """Pieces' edges in one array with offsets."""
    Es = [geo.edges(g) for g in groups]
    eptr = np.zeros(len(groups) + 1, np.int64)
    eptr[1:] = np.cumsum([len(e) for e in Es])
    E = np.ascontiguousarray(np.concatenate(Es) if Es and eptr[-1] else np.zeros((1, 4)))
    return eptr, E


def _adopt(cov: np.ndarray, has: np.ndarray, steps, reach: int = 1) -> np.ndarray:
    """peec.py:628-656: `cov` where `has`, and each value where not (orphan copper) added to the nearest cell along
    `steps` that has, up to `reach` steps away (of equals, the one with more copper of its own, then the first step);
    orphans with none in reach are dropped. Worked over the orphan cells only; each cell receives the moved values in
    step order, one per step at most."""
    own = np.where(has, cov, 0.0)
    out = own.copy()
    oi, oj = np.nonzero(~has & (cov > 0))
    val = cov[oi, oj]
    ni, nj = cov.shape
    for k in range(1, reach + 1):
        if len(oi) == 0:
            break
        best = np.full(len(oi), -1.0)
        pick = np.full(len(oi), -1)
        for s, (di, dj) in enumerate(steps):
            ti, tj = oi + di * k, oj + dj * k
            ok = (ti >= 0) & (ti < ni) & (tj >= 0) & (tj < nj)
            nb = np.full(len(oi), -1.0)
            sel = np.nonzero(ok)[0]
            hs = has[ti[sel], tj[sel]]
            nb[sel[hs]] = own[ti[sel[hs]], tj[sel[hs]]]
            better = nb > best                  # a neighbour with copper of the net (own >= 0) beats none (-1)
            best = np.where(better, nb, best)
            pick = np.where(better, s, pick)
        for s, (di, dj) in enumerate(steps):
            m = pick == s
            out[oi[m] + di * k, oj[m] + dj * k] += val[m]
        keep = pick < 0
        oi, oj, val = oi[keep], oj[keep], val[keep]
    return out


def _series_factor(face: np.ndarray, samples: list, has: np.ndarray, axis: int) -> np.ndarray:
    """peec.py:659-705: per strip of a lone narrow band (at most BAND_STRIPS side by side, no other strip of the net
    within 2 ADOPT_FACES lanes), the harmonic mean of the band's width at the sample lines over its width at the
    border; 1 elsewhere."""
    h = has if axis == 1 else has.T
    if not h.any():
        return np.ones(has.shape)
    f = face if axis == 1 else face.T
    rows = [s if axis == 1 else s.T for s in samples]
    n0, n1 = h.shape
    starts = h & ~np.pad(h, ((0, 0), (1, 0)))[:, :-1]
    run = np.cumsum(starts, axis=1) - 1
    ri, ci = np.nonzero(h)
    key = ri * n1 + run[ri, ci]
    nk = n0 * n1
    total = np.bincount(key, f[h], nk)
    size = np.bincount(key, minlength=nk)
    # lanes between a run and its neighbours in the same row (far: none)
    far = n1 + 2 * ADOPT_FACES + 2
    new_row = np.r_[True, ri[1:] != ri[:-1]]
    gap = np.where(new_row, far, np.r_[far, ci[1:] - ci[:-1]]) - 1
    first = gap != 0
    before = np.full(nk, far)
    before[key[first]] = gap[first]
    after = np.where(np.r_[size[1:] > 0, False], np.r_[before[1:], far], far)
    alone = (size <= BAND_STRIPS) & (before > 2 * ADOPT_FACES) & (after > 2 * ADOPT_FACES) & (total > 0)
    out = np.ones(h.shape)
    lone = np.nonzero(alone[key])[0]            # the strips of lone bands: only they can take a factor
    if len(lone):
        kl = key[lone]
        inv, ok = np.zeros(nk), np.ones(nk, bool)
        for s in rows:
            t = np.bincount(kl, s[ri[lone], ci[lone]], nk)
            ok &= t > 0
            inv += 1.0 / np.maximum(t, 1e-300)
        harmonic = len(rows) / np.maximum(inv, 1e-300)
        factor = np.where(alone & ok, harmonic / np.maximum(total, 1e-300), 1.0)
        out[ri[lone], ci[lone]] = factor[kl]
    return out if axis == 1 else out.T


def _cell_areas(E: np.ndarray, xl: float, yl: float, d: float, wi: int, wj: int, right: np.ndarray) -> np.ndarray:
    """peec.py:586-625: each cell's copper by Green's theorem, the oriented edges E cut where they cross the grid
    lines, each piece adding (its mean x - the cell's left) dy to the cell holding its midpoint, plus d times the
    covered length of the cell's right border."""
    out = np.zeros((wi, wj))
    if len(E):
        xa, ya, xb, yb = E.T
        ts, es = [np.zeros(len(E)), np.ones(len(E))], [np.arange(len(E))] * 2
        for a, b, origin, n in ((xa, xb, xl, wi + 1), (ya, yb, yl, wj + 1)):
            lo, hi = np.minimum(a, b), np.maximum(a, b)
            m0 = np.clip(np.ceil((lo - origin) / d), 0, n).astype(np.int64)
            m1 = np.clip(np.ceil((hi - origin) / d), 0, n).astype(np.int64)
            cnt = np.maximum(m1 - m0, 0)
            e = np.repeat(np.arange(len(E)), cnt)
            m = np.repeat(m0, cnt) + (np.arange(int(cnt.sum())) - np.repeat(np.cumsum(cnt) - cnt, cnt))
            ts.append((origin + m * d - a[e]) / (b[e] - a[e]))
            es.append(e)
        t, e = np.concatenate(ts), np.concatenate(es)
        o = np.lexsort((t, e))
        t, e = t[o], e[o]
        same = e[1:] == e[:-1]
        e, t0, t1 = e[:-1][same], t[:-1][same], t[1:][same]
        x0p, y0p = xa[e] + t0 * (xb[e] - xa[e]), ya[e] + t0 * (yb[e] - ya[e])
        x1p, y1p = xa[e] + t1 * (xb[e] - xa[e]), ya[e] + t1 * (yb[e] - ya[e])
        i = np.floor(((x0p + x1p) / 2 - xl) / d).astype(np.int64)
        j = np.floor(((y0p + y1p) / 2 - yl) / d).astype(np.int64)
        ok = (i >= 0) & (i < wi) & (j >= 0) & (j < wj)
        add = ((x0p + x1p) / 2 - (xl + i * d)) * (y1p - y0p)
        out.reshape(-1)[:] += np.bincount(i[ok] * wj + j[ok], add[ok], wi * wj)
    return out + d * right


def _strip_widths(exists, cross, offset, forced, d: float, exact=None):
    """peec.py:482-511: a strip's width (fraction of d): the measured border where there is copper on it, else the
    sub-grid's crossing share (at least a forced track's width), else 1; its centre's offset kept inside the cell."""
    w = np.where(forced > 0, np.maximum(cross, np.minimum(forced / d, 1.0)), cross)
    w = np.where(w > 0, np.minimum(w, 1.0), 1.0)
    if exact is not None:
        w = np.where(exact > 0, np.minimum(exact, 3.0), w)
    o = np.where(cross > 0, offset, 0.0)
    lim = np.maximum((1.0 - w) / 2, 0.0)
    return np.where(exists, w, 0.0), np.where(exists, np.clip(o, -lim, lim), 0.0)


def _exact_copper(dev: _Device, nets_polys: list, label, xb, yb, x0: float, y0: float, d: float):
    """peec.py:708-771: the strips' widths at their borders and the nodes' areas measured on the copper (line covers
    and Green's theorem), slivers adopted, and the narrow bands' series factors from STRIP_SAMPLES lines along each
    strip."""
    nx, ny = label.shape
    ex_x, ex_y, area = np.zeros((nx - 1, ny)), np.zeros((nx, ny - 1)), np.zeros((nx, ny))
    kx, ky = np.ones((nx - 1, ny)), np.ones((nx, ny - 1))
    for k, polys in enumerate(nets_polys):
        if not polys:
            continue
        E = geo.oriented_edges(polys)
        gx0, gy0, gx1, gy1 = geo.bounds(polys)
        i0, i1 = max(int(math.floor((gx0 - x0) / d)) - 2, 0), min(int(math.ceil((gx1 - x0) / d)) + 3, nx)
        j0, j1 = max(int(math.floor((gy0 - y0) / d)) - 2, 0), min(int(math.ceil((gy1 - y0) / d)) + 3, ny)
        wi, wj = i1 - i0, j1 - j0
        if wi < 2 or wj < 2:
            continue
        Ex = E                                   # lines x = const: u = x, v = y
        Ey = E[:, [1, 0, 3, 2]]                  # lines y = const: u = y, v = x
        cv = dev.cover(Ex, x0 + (i0 - 0.5) * d, d, wi + 1, y0 + j0 * d, d, wj)
        ch = dev.cover(Ey, y0 + (j0 + 0.5) * d, d, wj - 1, x0 + i0 * d, d, wi).T
        own = label[i0:i1, j0:j1] == k
        has_x = xb[i0:i1 - 1, j0:j1] & own[:-1, :]
        has_y = yb[i0:i1, j0:j1 - 1] & own[:, :-1]
        face_x = _adopt(cv[1:wi, :] / d, has_x, ((0, 1), (0, -1)), ADOPT_FACES)
        face_y = _adopt(ch / d, has_y, ((1, 0), (-1, 0)), ADOPT_FACES)
        ex_x[i0:i1 - 1, j0:j1] += face_x
        ex_y[i0:i1, j0:j1 - 1] += face_y
        q_n = STRIP_SAMPLES
        sx = (dev.cover(Ex, x0 + (i0 + 0.5 / q_n) * d, d / q_n, (wi - 1) * q_n, y0 + j0 * d, d, wj)
              / d).reshape(wi - 1, q_n, wj)
        sy = (dev.cover(Ey, y0 + (j0 + 0.5 / q_n) * d, d / q_n, (wj - 1) * q_n, x0 + i0 * d, d, wi)
              / d).reshape(wj - 1, q_n, wi)
        along_x = [_adopt(sx[:, q, :], has_x, ((0, 1), (0, -1)), ADOPT_FACES) for q in range(q_n)]
        along_y = [_adopt(sy[:, q, :].T, has_y, ((1, 0), (-1, 0)), ADOPT_FACES) for q in range(q_n)]
        win_x, win_y = kx[i0:i1 - 1, j0:j1], ky[i0:i1, j0:j1 - 1]
        win_x[has_x] = _series_factor(face_x, along_x, has_x, 1)[has_x]
        win_y[has_y] = _series_factor(face_y, along_y, has_y, 0)[has_y]
        cell = _cell_areas(E, x0 + (i0 - 0.5) * d, y0 + (j0 - 0.5) * d, d, wi, wj, cv[1:wi + 1, :]) / (d * d)
        area[i0:i1, j0:j1] += _adopt(np.clip(cell, 0.0, 1.0), own, ((1, 0), (-1, 0), (0, 1), (0, -1)), ADOPT_CELLS)
    return ex_x, ex_y, area, kx, ky


def _heat_points(dev: _Device, nets_polys: list, label, x0: float, y0: float, d: float):
    """peec.py:452-479: from each node lying outside its net's copper (not in it, not on its border) to the nearest
    point of that copper, a hair (1e-4 mm) past it; 0 elsewhere."""
    nx, ny = label.shape
    hx, hy = np.zeros((nx, ny)), np.zeros((nx, ny))
    for k, polys in enumerate(nets_polys):
        if not polys:
            continue
        ii, jj = np.nonzero(label == k)
        if len(ii) == 0:
            continue
        pts = np.column_stack([x0 + ii * d, y0 + jj * d])
        # the net's nodes against its copper, over the nodes' window of the node lattice
        w = (int(ii.min()), int(ii.max()) + 1, int(jj.min()), int(jj.max()) + 1)
        xs = np.ascontiguousarray(x0 + np.arange(w[0], w[1]) * d)
        ys = np.ascontiguousarray(y0 + np.arange(w[2], w[3]) * d)
        st = dev.state(polys, (0, w[1] - w[0], 0, w[3] - w[2]), xs, ys)
        out = st[ii - w[0], jj - w[2]] == 0
        if not out.any():
            continue
        q = dev.nearest(polys, pts[out])
        v = q - pts[out]
        n = np.hypot(v[:, 0], v[:, 1])
        step = np.where(n > 0, 1e-4 / np.maximum(n, 1e-300), 0.0)
        hx[ii[out], jj[out]] = v[:, 0] * (1 + step)
        hy[ii[out], jj[out]] = v[:, 1] * (1 + step)
    return hx, hy


def _walk_tracks(k: int, polylines, x0: float, y0: float, d: float, label, seeded, forced_x, forced_y, part,
                 owners) -> list:
    """peec.py:893-983: each track centreline (polyline, width mm) walked cell by cell, 4-connected: at samples 0.4
    pitch apart the nearest of the four cells around it that is already this track's copper (its net and piece),
    else the nearest free one; consecutive cells joined by the L-shaped path through fewer foreign cells. Free cells
    on the walk are claimed (seeded), strips between consecutive own cells are forced to the widest track's width,
    and a gap across foreign cells is returned as a bridge (i0, j0, i1, j1, length along the track mm, width mm)."""
    nx, ny = label.shape
    bridges = []
    for n_line, (pts, width) in enumerate(polylines):
        pid = owners[n_line] if owners is not None else None

        def own(c):
            return 0 <= c[0] < nx and 0 <= c[1] < ny and label[c] == k and (pid is None or part[c] == pid)

        def free(c):
            return 0 <= c[0] < nx and 0 <= c[1] < ny and (label[c] == -1 or own(c))

        def l_path(a, b):
            di, dj = int(np.sign(b[0] - a[0])), int(np.sign(b[1] - a[1]))
            along_x = [(a[0] + di * t, a[1]) for t in range(1, abs(b[0] - a[0]) + 1)]
            along_y = [(a[0], a[1] + dj * t) for t in range(1, abs(b[1] - a[1]) + 1)]
            xy = along_x + [(b[0], c[1]) for c in along_y]
            yx = along_y + [(c[0], b[1]) for c in along_x]
            return xy if sum(not free(c) for c in xy) <= sum(not free(c) for c in yx) else yx

        cells = []
        walked = 0.0
        for (ax, ay), (bx, by) in zip(pts, pts[1:]):
            seg = math.hypot(bx - ax, by - ay)
            for t in np.linspace(0.0, 1.0, max(int(math.ceil(seg / (0.4 * d))), 1) + 1):
                u, v = (ax + t * (bx - ax) - x0) / d, (ay + t * (by - ay) - y0) / d
                fi, fj = math.floor(u), math.floor(v)
                around = sorted([(fi, fj), (fi, fj + 1), (fi + 1, fj), (fi + 1, fj + 1)],
                                key=lambda c: (c[0] - u) ** 2 + (c[1] - v) ** 2)
                c = next((c for c in around if own(c)), None) or next((c for c in around if free(c)), around[0])
                s = walked + t * seg
                if not cells:
                    cells.append((c[0], c[1], s))
                elif cells[-1][:2] != c:
                    cells.extend((q[0], q[1], s) for q in l_path(cells[-1][:2], c))
            walked += seg
        prev = last = None
        for i, j, s in cells:
            if not (0 <= i < nx and 0 <= j < ny):
                prev = last = None
                continue
            if label[i, j] < 0:
                label[i, j] = k
                if pid is not None:
                    part[i, j] = pid
                seeded[i, j] = True
            if not own((i, j)):
                prev = None
                continue
            if prev is not None:
                tw = max(float(width), 1e-6)
                if prev[1] == j and abs(prev[0] - i) == 1:
                    forced_x[min(prev[0], i), j] = max(forced_x[min(prev[0], i), j], tw)
                elif prev[0] == i and abs(prev[1] - j) == 1:
                    forced_y[i, min(prev[1], j)] = max(forced_y[i, min(prev[1], j)], tw)
            elif last is not None and last[:2] != (i, j):
                bridges.append((last[0], last[1], i, j, max(s - last[2], d), float(width)))
            prev = (i, j)
            last = (i, j, s)
    return bridges


def _windows(groups: list, origin_x: float, origin_y: float, step: float, n_u: int, n_v: int, pad_lo: int,
             pad_hi: int) -> np.ndarray:
    """Per group the lattice columns and samples around its bounds (peec.py:371-373, :788-789)."""
    w = np.zeros((len(groups), 4), np.int32)
    for k, g in enumerate(groups):
        if not g:
            continue
        gx0, gy0, gx1, gy1 = geo.bounds(g)
        w[k] = (max(int(math.floor((gx0 - origin_x) / step)) - pad_lo, 0),
                min(int(math.ceil((gx1 - origin_x) / step)) + pad_hi, n_u),
                max(int(math.floor((gy0 - origin_y) / step)) - pad_lo, 0),
                min(int(math.ceil((gy1 - origin_y) / step)) + pad_hi, n_v))
    return w


def build_grid(geom, bbox: tuple, d: float, seeds: Optional[dict] = None, lines: Optional[dict] = None,
               sub: int = 1, device: int = 0) -> Grid:
    """Sample copper (mm) on a grid of pitch d covering bbox (x0, y0, x1, y1) (peec.py:329-449).

    `geom`: one geometry, or {net: geometry} (copper of different nets is never joined by a strip). `seeds`
    {net: [(x, y), ...]} makes the nearest empty node that net's copper (small via rings and pads). `lines`
    {net: [(polyline, width mm), ...]}: track centrelines whose nodes and strips are forced; a stretch across
    another net's cells becomes a bridge (`Grid.bridges`). `sub` > 1 derives nodes and strips from a sub-grid of
    pitch d/sub (2 .. 8) and measures the strips' widths, the nodes' copper areas and heat points on the copper."""
    x0, y0, x1, y1 = bbox
    nx = max(int(math.ceil((x1 - x0) / d)) + 1, 3)
    ny = max(int(math.ceil((y1 - y0) / d)) + 1, 3)
    xs = x0 + np.arange(nx) * d
    ys = y0 + np.arange(ny) * d
    by_net = not geo.is_geometry(geom)
    nets = list(geom.keys()) if by_net else [""]
    net_polys = [geo.polygons(g) for g in (geom.values() if by_net else [geom])]
    dev = _Device(device)
    fine = sub > 1
    plist: list = []                             # (net, piece polygons or None)
    if fine:
        if sub > 8:
            raise ValueError("sub-grid factor above 8")

        def inside(polys, pts):
            a, b = dev.locate(polys, pts)
            return a | b
        for k, polys in enumerate(net_polys):
            plist.extend((k, p) for p in geo.pieces(polys, inside))
        groups = [p for _, p in plist]
        df = d / sub
        mx, my = nx * sub, ny * sub
        xf = x0 - d / 2 + (np.arange(mx) + 0.5) * df
        yf = y0 - d / 2 + (np.arange(my) + 0.5) * df
        win = _windows(groups, xf[0], yf[0], df, mx, my, 1, 2)
        part, cross_x, cross_y, fill, off_x, off_y = dev.fine(groups, win, nx, ny, sub, xf, yf)
        net_of = np.array([k for k, _ in plist] + [-1], int)
        label = net_of[part]
        mids_x = np.where(cross_x > 0, label[:-1, :], -1)
        mids_y = np.where(cross_y > 0, label[:, :-1], -1)
    else:
        fill = None
        win = _windows(net_polys, x0, y0, d, nx, ny, 1, 2)
        lab = np.full((nx, ny), -1, np.int32)
        dev.label(net_polys, win, xs, ys, lab)
        label = lab.astype(int)
        mid = np.full((nx - 1, ny), -1, np.int32)
        wmx = win.copy()
        wmx[:, 1] = np.minimum(wmx[:, 1], nx - 1)
        dev.label(net_polys, wmx, np.ascontiguousarray(xs[:-1] + d / 2), ys, mid)
        mids_x = mid.astype(int)
        mid = np.full((nx, ny - 1), -1, np.int32)
        wmy = win.copy()
        wmy[:, 3] = np.minimum(wmy[:, 3], ny - 1)
        dev.label(net_polys, wmy, xs, np.ascontiguousarray(ys[:-1] + d / 2), mid)
        mids_y = mid.astype(int)
        part = label

    def pieces_at(k: int, pts: list) -> list:
        """For each point, the piece of net k holding it (on its border counts), else a new piece of its own, in
        order (peec.py:395-405)."""
        if not plist:
            return [k] * len(pts)
        P = np.asarray(pts, float).reshape(-1, 2)
        found = np.full(len(P), -1)
        for c, (kk, p) in enumerate(list(plist)):
            if kk != k or p is None or not (found < 0).any():
                continue
            bx0, by0, bx1, by1 = geo.bounds(p)
            near = np.nonzero((found < 0) & (P[:, 0] >= bx0) & (P[:, 0] <= bx1) & (P[:, 1] >= by0)
                              & (P[:, 1] <= by1))[0]
            if len(near):
                a, b = dev.locate(p, P[near])
                found[near[a | b]] = c
        out = []
        for f in found:
            if f < 0:
                plist.append((k, None))
                f = len(plist) - 1
            out.append(int(f))
        return out

    seeded = np.zeros((nx, ny), bool)
    for net, points in (seeds or {}).items():
        if net not in nets:
            continue
        k = nets.index(net)
        todo = []
        for (px, py) in points:
            i, j = int(round((px - x0) / d)), int(round((py - y0) / d))
            if 0 <= i < nx and 0 <= j < ny and label[i, j] < 0:
                label[i, j] = k
                seeded[i, j] = True
                todo.append((i, j, px, py))
        for (i, j, _, _), c in zip(todo, pieces_at(k, [(px, py) for _, _, px, py in todo])):
            part[i, j] = c
    forced_x = np.zeros((nx - 1, ny))
    forced_y = np.zeros((nx, ny - 1))
    bridges: list = []
    for net, polylines in (lines or {}).items():
        if net not in nets:
            continue
        k = nets.index(net)
        owners = None
        if plist:
            mids = []
            for pts, _ in polylines:                # the middle of each track's first segment (peec.py:887-890)
                (ax, ay), (bx, by) = pts[0], (pts[1] if len(pts) > 1 else pts[0])
                mids.append(((ax + bx) / 2, (ay + by) / 2))
            owners = pieces_at(k, mids)
        bridges += _walk_tracks(k, polylines, x0, y0, d, label, seeded, forced_x, forced_y, part, owners)
    # a strip: both nodes the same net and piece, and its midpoint that net's copper (or a seeded end, or a track)
    xb = (label[:-1, :] >= 0) & (label[:-1, :] == label[1:, :]) & (part[:-1, :] == part[1:, :]) & (
        (mids_x == label[:-1, :]) | seeded[:-1, :] | seeded[1:, :] | (forced_x > 0))
    yb = (label[:, :-1] >= 0) & (label[:, :-1] == label[:, 1:]) & (part[:, :-1] == part[:, 1:]) & (
        (mids_y == label[:, :-1]) | seeded[:, :-1] | seeded[:, 1:] | (forced_y > 0))
    g = Grid(x0, y0, d, nx, ny, label >= 0, xb, yb, label=label, nets=nets, fill=fill, bridges=bridges)
    if fine:
        ex_x, ex_y, area, kx, ky = _exact_copper(dev, net_polys, label, xb, yb, x0, y0, d)
        g.wx, g.ox = _strip_widths(xb, cross_x, off_x, forced_x, d, ex_x)
        g.wy, g.oy = _strip_widths(yb, cross_y, off_y, forced_y, d, ex_y)
        g.area = area
        g.rx, g.ry = g.wx * kx, g.wy * ky
        g.hx, g.hy = _heat_points(dev, net_polys, label, x0, y0, d)
    return g
