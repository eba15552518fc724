"""Seeded synthetic boards for the plane-current solve: copper masks, stitching vias and signal routes built
directly from a seed (no board files, no rasterisation). Geometry seeds are numpy `default_rng(1000 + k)` for
workload k, so every run (CPU and GPU) sees the same board.

Workloads (sizes are per copper layer):
  single_plane   one 64 x 64 plane at 0.5 mm, a straight microstrip 0.2 mm above it between two seeded cells,
                 100 MHz, GMRES(50) to 1e-10.
  two_layer      two 256 x 256 planes at 0.25 mm, 0.2 mm dielectric; a seeded 0.5 mm slot 20-60 mm long at 0 or
                 90 degrees in the lower plane; 20 stitching vias; one Manhattan route of 5-15 segments; 10
                 frequencies log-spaced 1 MHz - 1 GHz.
  four_layer     four 512 x 512 planes at 0.2 mm with seeded rectangular voids (0-15 % of each plane); 64 nets of
                 5-40 segments on random signal layers with 0-4 layer changes (each with a return via beside it);
                 4 frequencies per net in 10-500 MHz.
  large          six 2048 x 2048 planes at 0.05 mm, two of them split into 3-5 Voronoi regions, 200 stitching vias,
                 one 1 A source at 1 GHz.
`scaled(name, n)` gives the same construction at n x n cells (the tests' small cases).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from .problem import Branch, Grid, LidNode, PlaneLayer, TraceSegment, via_impedance

CU_T = 35e-6
DRILL = 0.3e-3


@dataclass
class Route:
    """One signal loop: its trace filaments and the return injection nodes (layer, i, j)."""
    trace: list
    driver: tuple
    receiver: tuple


@dataclass
class Board:
    """A synthetic board: plane layers, via branches, routes, the frequencies and the solver settings."""
    name: str
    layers: list
    branches: list
    routes: list
    freqs: list
    tol: float = 1e-10
    restart: Optional[int] = 50
    seed: int = 0
    meta: dict = field(default_factory=dict)


def _stitch(layers: list, i: int, j: int, name: str) -> list:
    """Via branches joining consecutive plane layers at cell (i, j) where both have copper."""
    out = []
    ks = [k for k, l in enumerate(layers) if l.grid.node[i, j]]
    for a, b in zip(ks[:-1], ks[1:]):
        length = abs(layers[a].z - layers[b].z)
        out.append(Branch((a, i, j), (b, i, j), via_impedance(length, DRILL), name=name, kind="via",
                          radius=DRILL / 2))
    return out


def decoupling_cap(layers: list, a: tuple, b: tuple, lid: int, C: float = 100e-9, esr: float = 5e-3,
                   esl: float = 0.5e-9, pad: int = 2, name: str = "decap") -> list:
    """A capacitor between two planes as the reference models a part (peec.py:1009-1038): a lid (LidNode) over cell
    a = (layer, i, j) and one over b, each joined by `bond` branches to the copper nodes of a pad x pad block of
    cells at its corner, and the part itself (ESR + jw ESL + 1/(jw C)) from lid to lid. Lids numbered lid, lid + 1."""
    la, lb = LidNode(*a, lid), LidNode(*b, lid + 1)
    out = []
    for lnode, (k, i, j) in ((la, a), (lb, b)):
        g = layers[k].grid.node
        for di in range(pad):
            for dj in range(pad):
                ii, jj = i + di, j + dj
                if 0 <= ii < g.shape[0] and 0 <= jj < g.shape[1] and g[ii, jj]:
                    out.append(Branch(lnode, (k, ii, jj), lambda f: complex(1e-4, 2 * math.pi * f * 5e-11),
                                      name=f"{name}_bond", kind="bond"))
    out.append(Branch(la, lb, lambda f, C=C, R=esr, L=esl: complex(R, 2 * math.pi * f * L - 1 / (2 * math.pi * f * C)),
                      name=name, kind="part"))
    return out


def _P(i: float, j: float, z: float, d: float) -> np.ndarray:
    """A point at grid position (i, j) (cells) and height z (m), in metres from node (0, 0)."""
    return np.array([i * d, j * d, z])


def _manhattan(rng, nx: int, ny: int, nseg: int, margin: int, step: tuple) -> list:
    """Vertices of a seeded Manhattan polyline of nseg segments inside the margin, alternating x and y legs."""
    i, j = int(rng.integers(margin, nx - margin)), int(rng.integers(margin, ny - margin))
    pts = [(i, j)]
    horiz = bool(rng.integers(2))
    for _ in range(nseg):
        for _try in range(20):
            L = int(rng.integers(step[0], step[1] + 1)) * (1 if rng.integers(2) else -1)
            ni, nj = (i + L, j) if horiz else (i, j + L)
            if margin <= ni < nx - margin and margin <= nj < ny - margin:
                break
        else:
            L = -L
            ni, nj = (i + L, j) if horiz else (i, j + L)
            ni, nj = min(max(ni, margin), nx - margin - 1), min(max(nj, margin), ny - margin - 1)
        i, j = ni, nj
        pts.append((i, j))
        horiz = not horiz
    return pts


def _route(pts: list, z_trace: float, z_plane: float, d: float, width: float, layer: int) -> Route:
    """A route along the vertices at z_trace, with vertical legs down to the plane at both ends."""
    trace = [TraceSegment(_P(*pts[0], z_plane, d), _P(*pts[0], z_trace, d), width)]
    for a, b in zip(pts[:-1], pts[1:]):
        if a != b:
            trace.append(TraceSegment(_P(*a, z_trace, d), _P(*b, z_trace, d), width))
    trace.append(TraceSegment(_P(*pts[-1], z_trace, d), _P(*pts[-1], z_plane, d), width))
    return Route(trace, (layer, *pts[0]), (layer, *pts[-1]))


def single_plane(n: int = 64, pitch_mm: float = 0.5, seed: int = 1001) -> Board:
    """One plane, a straight microstrip 0.2 mm above it between two seeded cells, 100 MHz."""
    rng = np.random.default_rng(seed)
    d = pitch_mm * 1e-3
    g = Grid.from_mask(np.ones((n, n), bool), pitch_mm)
    layer = PlaneLayer("plane", 0.0, g, CU_T)
    m = max(2, n // 16)
    j = int(rng.integers(m, n - m))
    i0 = int(rng.integers(m, n // 2 - 1))
    i1 = int(rng.integers(n // 2 + 1, n - m))
    route = _route([(i0, j), (i1, j)], 0.2e-3, 0.0, d, 0.2e-3, 0)
    return Board("single_plane", [layer], [], [route], [100e6], tol=1e-10, restart=50, seed=seed)


def two_layer(n: int = 256, pitch_mm: float = 0.25, seed: int = 1002, n_vias: int = 20,
              n_freqs: int = 10) -> Board:
    """Two planes 0.2 mm apart, a seeded slot in the lower one, stitching vias, one Manhattan route."""
    rng = np.random.default_rng(seed)
    d = pitch_mm * 1e-3
    scale = n / 256.0                               # the slot and route scale with the board
    top = np.ones((n, n), bool)
    low = np.ones((n, n), bool)
    width = max(1, int(round(0.5 / pitch_mm)))
    length = int(round(rng.uniform(20.0, 60.0) * scale / pitch_mm))
    length = min(length, n - 2)
    vertical = bool(rng.integers(2))
    a0 = int(rng.integers(1, n - length)) if n - length > 1 else 0
    c0 = int(rng.integers(n // 4, 3 * n // 4))
    if vertical:
        low[c0:c0 + width, a0:a0 + length] = False
    else:
        low[a0:a0 + length, c0:c0 + width] = False
    z_top, z_low = 0.0, -(0.2e-3 + CU_T)
    layers = [PlaneLayer("L1", z_top, Grid.from_mask(top, pitch_mm), CU_T),
              PlaneLayer("L2", z_low, Grid.from_mask(low, pitch_mm), CU_T)]
    branches = []
    while len(branches) < n_vias:
        i, j = (int(v) for v in rng.integers(1, n - 1, 2))
        if low[i, j]:
            branches += _stitch(layers, i, j, f"via{len(branches)}")
    nseg = int(rng.integers(5, 16))
    pts = _manhattan(rng, n, n, nseg, max(2, n // 16), (max(2, int(8 * scale)), max(3, int(40 * scale))))
    route = _route(pts, 0.2e-3, z_top, d, 0.15e-3, 0)
    freqs = list(np.logspace(6, 9, n_freqs))
    return Board("two_layer", layers, branches, [route], freqs, tol=1e-10, restart=50, seed=seed,
                 meta={"slot": (vertical, a0, c0, length, width)})


def four_layer(n: int = 512, pitch_mm: float = 0.2, seed: int = 1003, n_nets: int = 64, n_freqs: int = 4,
               n_stitch: int = 32) -> Board:
    """Four planes with seeded voids, signal nets on the layers between them, 4 frequencies per net."""
    rng = np.random.default_rng(seed)
    d = pitch_mm * 1e-3
    zs = [0.0, -0.3e-3, -1.2e-3, -1.5e-3]
    masks = []
    for k in range(4):
        m = np.ones((n, n), bool)
        target = rng.uniform(0.0, 0.15) * n * n
        cut = 0
        while cut < target:
            w, h = (int(v) for v in rng.integers(max(2, n // 64), max(3, n // 10), 2))
            i, j = int(rng.integers(0, n - w)), int(rng.integers(0, n - h))
            cut += np.count_nonzero(m[i:i + w, j:j + h])
            m[i:i + w, j:j + h] = False
        masks.append(m)
    layers = [PlaneLayer(f"P{k}", zs[k], Grid.from_mask(masks[k], pitch_mm), CU_T) for k in range(4)]
    branches = []
    for v in range(n_stitch):
        i, j = (int(t) for t in rng.integers(1, n - 1, 2))
        branches += _stitch(layers, i, j, f"stitch{v}")
    # signal layers: above P0, between P1 and P2, below P3; each with the plane its return uses
    sig = [(0.2e-3, 0), (-0.75e-3, 1), (-1.7e-3, 3)]
    routes = []
    scale = n / 512.0
    for net in range(n_nets):
        lay = int(rng.integers(0, 3))
        nseg = int(rng.integers(5, 41))
        pts = _manhattan(rng, n, n, nseg, max(2, n // 32), (max(2, int(4 * scale)), max(3, int(24 * scale))))
        changes = int(rng.integers(0, 5))
        where = sorted(rng.choice(np.arange(1, len(pts) - 1), size=min(changes, len(pts) - 2), replace=False))
        trace = []
        z, plane = sig[lay]
        start_plane = plane
        trace.append(TraceSegment(_P(*pts[0], zs[plane], d), _P(*pts[0], z, d), 0.15e-3))
        for k, (a, b) in enumerate(zip(pts[:-1], pts[1:])):
            if a != b:
                trace.append(TraceSegment(_P(*a, z, d), _P(*b, z, d), 0.15e-3))
            if k + 1 in where:
                nlay = (lay + 1 + int(rng.integers(0, 2))) % 3
                nz, nplane = sig[nlay]
                trace.append(TraceSegment(_P(*b, z, d), _P(*b, nz, d), 0.15e-3))
                ri, rj = min(b[0] + 1, n - 2), b[1]
                branches += _stitch(layers, ri, rj, f"ret{net}_{k}")
                lay, z, plane = nlay, nz, nplane
        trace.append(TraceSegment(_P(*pts[-1], z, d), _P(*pts[-1], zs[plane], d), 0.15e-3))
        drv = _nearest_copper(layers[start_plane].grid.node, start_plane, pts[0])
        rcv = _nearest_copper(layers[plane].grid.node, plane, pts[-1])
        routes.append(Route(trace, drv, rcv))
    freqs = list(np.exp(rng.uniform(math.log(10e6), math.log(500e6), n_freqs)))
    return Board("four_layer", layers, branches, routes, freqs, tol=1e-8, restart=50, seed=seed)


def _nearest_copper(mask: np.ndarray, layer: int, p: tuple) -> tuple:
    """The copper node of `mask` nearest to cell p, as (layer, i, j)."""
    if mask[p]:
        return (layer, int(p[0]), int(p[1]))
    ii, jj = np.nonzero(mask)
    k = int(np.argmin((ii - p[0]) ** 2 + (jj - p[1]) ** 2))
    return (layer, int(ii[k]), int(jj[k]))


def large(n: int = 2048, pitch_mm: float = 0.05, seed: int = 1005, n_vias: int = 200) -> Board:
    """Six planes, two split into Voronoi regions, stitching vias, one switching-node source at 1 GHz."""
    rng = np.random.default_rng(seed)
    d = pitch_mm * 1e-3
    zs = [-k * 0.2e-3 for k in range(6)]
    masks = [np.ones((n, n), bool) for _ in range(6)]
    for k in (2, 3):
        nreg = int(rng.integers(3, 6))
        seeds = rng.uniform(0, n, (nreg, 2))
        ii, jj = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
        dist = (ii[..., None] - seeds[:, 0]) ** 2 + (jj[..., None] - seeds[:, 1]) ** 2
        lab = np.argmin(dist, axis=-1)
        edge = np.zeros((n, n), bool)
        edge[:-1, :] |= lab[:-1, :] != lab[1:, :]
        edge[:, :-1] |= lab[:, :-1] != lab[:, 1:]
        masks[k] &= ~edge
    layers = [PlaneLayer(f"P{k}", zs[k], Grid.from_mask(masks[k], pitch_mm), CU_T) for k in range(6)]
    branches = []
    for v in range(n_vias):
        i, j = (int(t) for t in rng.integers(1, n - 1, 2))
        branches += _stitch(layers, i, j, f"via{v}")
    pts = _manhattan(rng, n, n, 8, n // 16, (n // 32, n // 8))
    route = _route(pts, 0.2e-3, zs[0], d, 0.15e-3, 0)
    # GMRES(100): the reference's default restart at this size (min(200, max(100, 4e8 / 16 n)), peec.py:3087-3090)
    return Board("large", layers, branches, [route], [1e9], tol=1e-8, restart=100, seed=seed)


def moved_slot(board: Board, rng, protect=()) -> tuple:
    """A topology-changing what-if of `board`: its seeded slot (the two-layer board's, in its lower plane) moved to a
    seeded new place, or on a board without one a new slot of the same kind (0.5 mm wide, 20-60 mm scaled to the
    board) cut into a seeded plane. The slot avoids the cells in `protect` ((layer, i, j): the routes' injection
    nodes) and leaves a margin of two cells at the board's edge. (new layers, (layer, vertical, a0, c0, length,
    width))."""
    g0 = board.layers[0].grid
    n0, n1, pitch = g0.nx, g0.ny, g0.d
    masks = [np.array(l.grid.node, bool) for l in board.layers]
    old = board.meta.get("slot")
    k = 1 if old is not None else int(rng.integers(len(board.layers)))
    if old is not None:
        vertical, a0, c0, length, width = old
        if vertical:
            masks[k][c0:c0 + width, a0:a0 + length] = True
        else:
            masks[k][a0:a0 + length, c0:c0 + width] = True
    if old is not None:                     # the same slot, moved
        length, width = old[3], old[4]
    else:                                   # 20-60 mm of the two-layer board's 64 mm side, 0.5 mm wide
        width = max(1, int(round(0.5 / pitch)))
        length = int(round(rng.uniform(20.0, 60.0) / 64.0 * min(n0, n1)))
    length = max(2, min(length, min(n0, n1) - 6))
    keep_out = {(i, j) for (l, i, j) in protect if l == k}
    for _ in range(100):
        vertical = bool(rng.integers(2))
        a0 = int(rng.integers(2, min(n0, n1) - length - 2)) if min(n0, n1) - length - 4 > 0 else 2
        c0 = int(rng.integers(2, (n0 if vertical else n1) - width - 2))
        cells = ({(c0 + w, a0 + t) for w in range(width) for t in range(length)} if vertical
                 else {(a0 + t, c0 + w) for w in range(width) for t in range(length)})
        if not cells & keep_out:
            break
    m = masks[k].copy()
    for (i, j) in cells:
        m[i, j] = False
    masks[k] = m
    layers = [PlaneLayer(l.name, l.z, Grid.from_mask(mk, pitch), l.thickness) for l, mk in zip(board.layers, masks)]
    return layers, (k, vertical, a0, c0, length, width)


WORKLOADS = {"single_plane": single_plane, "two_layer": two_layer, "four_layer": four_layer, "large": large}


def scaled(name: str, n: int, **kw) -> Board:
    """Workload `name` built at n x n cells per layer (pitch scaled to keep the board's physical size)."""
    full = {"single_plane": (64, 0.5), "two_layer": (256, 0.25), "four_layer": (512, 0.2), "large": (2048, 0.05)}
    n0, p0 = full[name]
    return WORKLOADS[name](n=n, pitch_mm=p0 * n0 / n, **kw)


def antipad_pour(size_mm: float = 200.0, pitch_mm: float = 2.0, radius_mm: float = 0.3, seed: int = 1006) -> dict:
    """Copper for the rasteriser: a size x size mm ground pour 0.5 mm in from the board edge, with an antipad (a
    16-segment-a-quadrant circle of radius_mm) every pitch_mm, jittered by up to 0.3 mm: {"GND": geometry}."""
    from .geometry import box, disc
    rng = np.random.default_rng(seed)
    holes = []
    for x in np.arange(2.0, size_mm - 1.5, pitch_mm):
        for y in np.arange(2.0, size_mm - 1.5, pitch_mm):
            holes.append(disc(x + rng.uniform(-0.3, 0.3), y + rng.uniform(-0.3, 0.3), radius_mm, 16)[::-1])
    return {"GND": box(0.5, 0.5, size_mm - 0.5, size_mm - 0.5, holes=holes)}
