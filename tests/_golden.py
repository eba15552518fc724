"""Loads the reference-solve fixtures under tests/golden/ (written by tests/golden/make_solve_fixtures.py) and
rebuilds their boards as this library's problem objects, which the oracle's classes also take."""
from __future__ import annotations

import math
from dataclasses import dataclass
from pathlib import Path

import numpy as np

from groundtruth_amd.problem import Branch, Grid, LidNode, PlaneLayer, TraceSegment

GOLDEN = Path(__file__).resolve().parent / "golden"
NAMES = ["single_plane_64", "two_layer_48", "four_layer_24", "large_32", "mixed_copper_32", "decap_lid_32",
         "narrow_sub4"]
KINDS = ["via", "part", "bond"]


@dataclass
class Fixture:
    name: str
    data: dict
    layers: list
    branches: list
    routes: list            # [(trace, driver, receiver)]

    @property
    def solves(self):
        return list(enumerate(zip(self.data["solve_route"].tolist(), self.data["solve_f"].tolist())))


def _z(R: float, L: float, C: float = 0.0):
    """R + jwL (+ 1/(jwC) when C > 0)."""
    if C > 0:
        return lambda f, R=R, L=L, C=C: complex(R, 2 * math.pi * f * L - 1.0 / (2 * math.pi * f * C))
    return lambda f, R=R, L=L: complex(R, 2 * math.pi * f * L)


def _end(p, lid: int):
    """A branch end: a grid node (layer, i, j), or the lid numbered `lid` over that cell."""
    return LidNode(*(int(v) for v in p), int(lid)) if lid >= 0 else tuple(int(v) for v in p)


def load(name: str) -> Fixture:
    with np.load(GOLDEN / f"solve_{name}.npz") as z:
        a = {k: z[k] for k in z.files}
    pitch = float(a["pitch_mm"])
    layers = []
    for k, m in enumerate(a["nodes"]):
        if "wx" in a:            # a rasterised board: its strips and measured widths
            g = Grid(0.0, 0.0, pitch, m.shape[0], m.shape[1], m.copy(), a["xb"][k].copy(), a["yb"][k].copy(),
                     **{f: a[f][k].copy() for f in ("wx", "wy", "ox", "oy", "rx", "ry")})
        else:
            g = Grid.from_mask(m, pitch)
        layers.append(PlaneLayer(f"L{k}", float(a["z"][k]), g, float(a["t"][k])))
    nbr = len(a["br_R"])
    lid_a = a.get("br_lid_a", -np.ones(nbr, np.int64))
    lid_b = a.get("br_lid_b", -np.ones(nbr, np.int64))
    caps = a.get("br_C", np.zeros(nbr))
    kinds = a.get("br_kind", np.zeros(nbr, np.int64))
    branches = [Branch(_end(pa, la), _end(pb, lb), _z(float(R), float(L), float(Cc)), f"br{n}", KINDS[int(kd)],
                       float(rad))
                for n, (pa, pb, la, lb, R, L, Cc, kd, rad) in enumerate(zip(a["br_a"], a["br_b"], lid_a, lid_b,
                                                                          a["br_R"], a["br_L"], caps, kinds,
                                                                          a["br_radius"]))]
    routes = []
    for k in range(len(a["drv"])):
        sel = np.nonzero(a["tr_route"] == k)[0]
        trace = [TraceSegment(a["tr_a"][s].copy(), a["tr_b"][s].copy(), float(a["tr_w"][s]), float(a["tr_i"][s]))
                 for s in sel]
        routes.append((trace, tuple(int(v) for v in a["drv"][k]), tuple(int(v) for v in a["rcv"][k])))
    return Fixture(name, a, layers, branches, routes)


def l2(a, b) -> float:
    a, b = np.asarray(a).ravel(), np.asarray(b).ravel()
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def currents(I_x, I_y, I_branch) -> np.ndarray:
    return np.concatenate([np.ravel(I_x), np.ravel(I_y), np.ravel(I_branch)])
