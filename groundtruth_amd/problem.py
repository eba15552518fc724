"""Inputs and outputs of the plane-current PEEC solve, with the reference's names and meanings.

Grid / PlaneLayer / Branch / LidNode / TraceSegment / Solution / SolverError mirror `src/groundtruth/solver/peec.py`
(Grid :197-223, PlaneLayer :987-993, Branch :996-1006, LidNode :1009-1038, TraceSegment :1051-1058,
via_impedance :1061-1072, track_impedance :1075-1089, Solution :1092-1110, SolverError :124-127) so a caller of the
reference's solver passes the same objects.
Only the fields the solve reads are kept on Grid (copper masks and pitch); rasterisation is not part of this path.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Callable, Optional

import numpy as np

MU0 = 4e-7 * math.pi
SIGMA_CU = 5.8e7                # S/m, annealed copper at 20 C
VIA_PLATING = 25e-6             # m, a via barrel's plating


class SolverError(RuntimeError):
    """A model that cannot be solved (no copper where current must go, unbalanced injections, GMRES not
    converging): reported by the caller, not a bug (peec.py:124-127)."""


class DeviceTooSmall(RuntimeError):
    """The device's free memory cannot hold the model: raised before anything is uploaded (gpu.py:576-577, raised
    at gpu.py:483-488 when the model needs more than 0.8 of the free memory)."""


class DeviceError(RuntimeError):
    """The HIP library failed (missing library, out of device memory, a failed launch)."""


@dataclass
class Grid:
    """One layer's copper on a square grid of pitch d (mm): copper nodes and the strips joining neighbours."""
    x0: float
    y0: float
    d: float                    # mm
    nx: int
    ny: int
    node: np.ndarray            # (nx, ny) bool
    xb: np.ndarray              # (nx-1, ny) bool: x strip between (i, j) and (i+1, j)
    yb: np.ndarray              # (nx, ny-1) bool
    # what the rasteriser measures (raster.build_grid; peec.py:197-223), None where a grid has no such field
    label: Optional[np.ndarray] = None      # (nx, ny) index into nets, -1 no copper
    nets: list = field(default_factory=list)
    fill: Optional[np.ndarray] = None       # (nx, ny) copper share of each node's cell (the sub-grid's)
    bridges: list = field(default_factory=list)     # [(i0, j0, i1, j1, length mm, width mm)] tracks across
    wx: Optional[np.ndarray] = None         # (nx-1, ny) strip copper width, a fraction of d (None: full)
    wy: Optional[np.ndarray] = None         # (nx, ny-1)
    ox: Optional[np.ndarray] = None         # offset of that copper's centre across the strip, a fraction of d
    oy: Optional[np.ndarray] = None
    area: Optional[np.ndarray] = None       # (nx, ny) the node's own copper in its cell, a fraction of d^2
    rx: Optional[np.ndarray] = None         # the strip's width for its resistance (a change of width in series)
    ry: Optional[np.ndarray] = None
    hx: Optional[np.ndarray] = None         # (nx, ny) mm from a node outside its copper to the copper (else 0)
    hy: Optional[np.ndarray] = None

    def xy(self, i, j):
        """Board coordinates (mm) of node (i, j) (arrays too)."""
        return self.x0 + np.asarray(i) * self.d, self.y0 + np.asarray(j) * self.d

    def nearest_node(self, x: float, y: float, max_dist: Optional[float] = None,
                     net: Optional[str] = None) -> Optional[tuple]:
        """The copper node nearest (x, y) mm, of `net` only when given; None when there is none (within max_dist)
        (peec.py:229-255)."""
        mask = self.node
        if net is not None and self.label is not None:
            if net not in self.nets:
                return None
            mask = self.label == self.nets.index(net)
        ci, cj = int(round((x - self.x0) / self.d)), int(round((y - self.y0) / self.d))
        if max_dist is not None:
            r = max(int(math.ceil(max_dist / self.d)) + 1, 2)
            i0, j0 = max(ci - r, 0), max(cj - r, 0)
            ii, jj = np.nonzero(mask[i0:ci + r + 1, j0:cj + r + 1])
            ii, jj = ii + i0, jj + j0
        else:
            ii, jj = np.nonzero(mask)
        if len(ii) == 0:
            return None
        px, py = self.xy(ii, jj)
        d2 = (px - x) ** 2 + (py - y) ** 2
        k = int(np.argmin(d2))
        if max_dist is not None and d2[k] > max_dist ** 2:
            return None
        return int(ii[k]), int(jj[k])

    def net_at(self, node: tuple) -> str:
        """The net of a node's copper ("" for none or a grid without nets)."""
        if self.label is None or self.label[node] < 0:
            return ""
        return self.nets[self.label[node]]

    @property
    def extent(self) -> tuple:
        """(x0, x1, y0, y1) mm the cells cover, each node at its cell's centre."""
        return (self.x0 - self.d / 2, self.x0 + (self.nx - 0.5) * self.d,
                self.y0 - self.d / 2, self.y0 + (self.ny - 0.5) * self.d)

    def copper_area(self) -> np.ndarray:
        """Each node's copper as a fraction of its cell (peec.py:259-313): the measured area shared over the nodes of
        a 3 x 3 neighbourhood of one net in proportion to their strips' widths along their run; the mean width of
        its strips where nothing was measured; 1 at full strips, 0 off copper."""
        out = self.node.astype(float)
        if self.wx is None or self.wy is None:
            return out
        xw, yw = np.where(self.xb, self.wx, 0.0), np.where(self.yb, self.wy, 0.0)
        sx, sy = np.zeros(out.shape), np.zeros(out.shape)
        nxs, nys = np.zeros(out.shape), np.zeros(out.shape)
        sx[:-1] += xw
        sx[1:] += xw
        nxs[:-1] += self.xb
        nxs[1:] += self.xb
        sy[:, :-1] += yw
        sy[:, 1:] += yw
        nys[:, :-1] += self.yb
        nys[:, 1:] += self.yb
        n = nxs + nys
        mean_width = np.where(n > 0, (sx + sy) / np.maximum(n, 1), out)
        if self.area is None:
            return mean_width
        xr, yr = nxs == 2, nys == 2
        run_width = np.where(xr & ~yr, sx / np.maximum(nxs, 1), np.where(yr & ~xr, sy / np.maximum(nys, 1),
                                                                         mean_width))
        area = np.where(self.node, self.area, 0.0)
        width = np.where(self.node, run_width, 0.0)
        lab = np.where(self.node, self.label if self.label is not None else 0, -1)
        sum_a, sum_w = np.zeros(out.shape), np.zeros(out.shape)
        ni, nj = out.shape
        for di in (-1, 0, 1):
            for dj in (-1, 0, 1):
                src = (slice(max(di, 0), ni + min(di, 0)), slice(max(dj, 0), nj + min(dj, 0)))
                dst = (slice(max(-di, 0), ni + min(-di, 0)), slice(max(-dj, 0), nj + min(-dj, 0)))
                same = (lab[src] == lab[dst]) & (lab[dst] >= 0)
                sum_a[dst] += np.where(same, area[src], 0.0)
                sum_w[dst] += np.where(same, width[src], 0.0)
        shared = np.where(sum_w > 0, width * sum_a / np.maximum(sum_w, 1e-12), 0.0)
        return np.where(self.node, np.where(shared > 0, shared, mean_width), 0.0)

    @classmethod
    def from_mask(cls, node: np.ndarray, d_mm: float, x0: float = 0.0, y0: float = 0.0) -> "Grid":
        """A grid whose strips join every pair of neighbouring copper nodes."""
        node = np.asarray(node, bool)
        return cls(x0, y0, d_mm, node.shape[0], node.shape[1], node,
                   node[:-1, :] & node[1:, :], node[:, :-1] & node[:, 1:])


@dataclass
class PlaneLayer:
    """A copper layer: name, height z (m), rasterised copper, copper thickness (m)."""
    name: str
    z: float
    grid: Grid
    thickness: float = 35e-6


@dataclass
class Branch:
    """A lumped connection between two grid nodes (layer, i, j), impedance z(f)."""
    a: tuple
    b: tuple
    z: Callable[[float], complex]
    name: str = ""
    kind: str = "part"
    radius: float = 1.5e-4


@dataclass
class TraceSegment:
    """A straight piece of the known current path from a to b (m), `i` amps per amp of loop current."""
    a: np.ndarray
    b: np.ndarray
    width: float
    i: float = 1.0


def filament_self(length: float, rho: float) -> float:
    """Partial self inductance of a straight filament of radius rho (the mutual of two parallel filaments rho
    apart, Neumann's integral in closed form [Grover 1946])."""
    def g(u):
        return u * math.asinh(u / rho) - math.sqrt(u * u + rho * rho)
    return MU0 / (4 * math.pi) * (g(length) - 2 * g(0.0) + g(-length))


def via_impedance(length_m: float, drill_m: float, sigma: float = SIGMA_CU, filled: bool = False,
                  plating: float = VIA_PLATING) -> Callable[[float], complex]:
    """A plated barrel: R of its plating ring (or the whole hole when filled) + jw L of a filament at its radius
    (peec.py:1061-1072)."""
    r = max(drill_m / 2, 5e-5)
    L = filament_self(max(length_m, 1e-6), r)
    area = math.pi * r * r if filled else math.pi * 2 * r * plating
    R = max(length_m, 1e-6) / (sigma * area)
    return lambda f, R=R, L=L: complex(R, 2 * math.pi * f * L)


class LidNode(tuple):
    """A circuit node that is no grid node: the metal of one side of a lumped part (a capacitor's terminal), which
    `bond` branches join to the grid nodes under its pads (peec.py:1009-1038). As a tuple it is the (layer, i, j) of
    the cell it sits over (its position for the loop moment); `lid` numbers it within its model. Equality and hashing
    go by the number, so two lids over one cell stay two nodes, and a grid node never equals a lid."""

    def __new__(cls, layer: int, i: int, j: int, lid: int):
        t = super().__new__(cls, (int(layer), int(i), int(j)))
        t.lid = int(lid)
        return t

    def __getnewargs__(self):
        return (self[0], self[1], self[2], self.lid)

    def __eq__(self, other):
        return isinstance(other, LidNode) and other.lid == self.lid and tuple.__eq__(self, other)

    def __ne__(self, other):
        return not self.__eq__(other)

    def __hash__(self):
        return hash(("lid", self.lid))


def track_impedance(length_m: float, width_m: float, thickness_m: float = 35e-6, sigma: float = SIGMA_CU
                    ) -> Callable[[float], complex]:
    """A piece of track as a lumped branch (a bridge across another net's cells, peec.py:1075-1089): its resistance
    with the copper's effective thickness at f for the skin effect, plus jw its partial self inductance as a flat
    strip (the filament mutual at the strip's geometric mean distance 0.2235 (w + t) [Grover 1946])."""
    from .greens import skin_thickness, strip_self_gmd
    length_m = max(length_m, 1e-6)
    L = strip_self_gmd(length_m, width_m, thickness_m)

    def z(f: float) -> complex:
        return complex(length_m / (sigma * width_m * skin_thickness(thickness_m, f, sigma)), 2 * math.pi * f * L)
    return z


@dataclass
class Solution:
    """One solve per amp of loop current (peec.py:1092-1110): strip currents, node potentials, branch currents,
    loop impedance, dipole moment, iterations; `error` is the refinement's stated relative error bound where a
    mixed-precision solve reports one."""
    f: float
    I_x: np.ndarray
    I_y: np.ndarray
    V: np.ndarray
    I_branch: np.ndarray
    Z_loop: complex
    m: np.ndarray
    iterations: int
    x: Optional[np.ndarray] = field(default=None, repr=False)
    residual: float = float("nan")
    V_lid: Optional[np.ndarray] = None      # the potentials of the stack's lids, in their numbers' order
    lid_at: Optional[dict] = None           # a lid's number -> its place in V_lid

    def state_for(self, stack) -> np.ndarray:
        """The solver's unknown vector [branch currents; kept node potentials] of this solution on `stack`, the
        stack it was solved on (a warm start), rebuilt from the currents and potentials held here
        (peec.py:1127-1144)."""
        if self.x is not None and len(self.x) == stack.nb + stack.n_keep:
            return self.x
        strips = np.concatenate([self.I_x[tuple(stack.xa.T)], self.I_y[tuple(stack.ya.T)]])
        Ib = np.concatenate([strips, np.asarray(self.I_branch, complex)])
        n_lid = stack.n_nodes - stack.n_grid
        lids = self.V_lid if self.V_lid is not None and len(self.V_lid) == n_lid else np.zeros(n_lid, complex)
        return np.concatenate([Ib, np.concatenate([self.V[np.nonzero(stack.node)], lids])[stack.keep]])

    @property
    def L_loop(self) -> float:
        """Loop inductance (H): Im(Z_loop) / w."""
        return self.Z_loop.imag / (2 * math.pi * self.f)

    def at(self, node) -> complex:
        """A node's potential (peec.py:1118-1125): a grid node's from V, a LidNode's from V_lid by its place in the
        stack; KeyError for a lid the stack does not have."""
        if isinstance(node, LidNode):
            if self.V_lid is None or self.lid_at is None:
                raise KeyError(f"no potential kept for lid {node.lid}")
            return complex(self.V_lid[self.lid_at[node.lid]])
        return complex(self.V[tuple(node)])

    def node_field(self, k: int, d_mm: float, copper: bool = False) -> tuple:
        """Complex (Kx, Ky) at the nodes of layer k, in A/m per A of signal (peec.py:1151-1199): at each node, the
        mean of the strips that touch it in each direction and carry current, over the pitch. copper=True asks for
        the current per width of the copper itself, which needs the strips' copper widths and the nodes' copper
        areas from the reference's rasterisation; this library's strips are full cells, so it is refused."""
        if copper:
            raise NotImplementedError("copper-width fields need the rasterised copper widths and areas")
        d = d_mm * 1e-3

        def mean_of_touching(I: np.ndarray, axis: int) -> np.ndarray:
            # strip s along `axis` joins nodes s and s + 1: pad one empty strip at each end, so node n sees the
            # strips n - 1 and n
            pad = [(0, 0), (0, 0)]
            pad[axis] = (1, 1)
            P = np.pad(I, pad)
            lo, hi = np.take(P, range(P.shape[axis] - 1), axis), np.take(P, range(1, P.shape[axis]), axis)
            n = (lo != 0).astype(float) + (hi != 0)
            return (lo + hi) / np.maximum(n, 1.0) / d

        return mean_of_touching(self.I_x[k], 0), mean_of_touching(self.I_y[k], 1)

    def sheet_current(self, k: int, d_mm: float, copper: bool = False) -> np.ndarray:
        """|K| at the nodes of layer k, A/m per A of signal (peec.py:1201-1204)."""
        kx, ky = self.node_field(k, d_mm, copper)
        return np.sqrt(np.abs(kx) ** 2 + np.abs(ky) ** 2)

    def dominant_phase(self) -> float:
        """The phase phi at which the strip currents are largest (peec.py:1206-1223). Over all strips z = I, the
        in-phase energy sum Re(z e^-j phi)^2 = (sum |z|^2 + Re(e^-2j phi sum z^2)) / 2 peaks at phi = arg(sum z^2) / 2
        (modulo pi); of the two ends, the one along which Re(z e^-j phi) follows Re(z) (the 1 A excitation is real).
        0 when sum z^2 is 0. Computed once per solution."""
        if getattr(self, "_phase", None) is None:
            z = np.concatenate([np.ravel(self.I_x), np.ravel(self.I_y)])
            q = complex(np.sum(z * z))
            phi = 0.5 * math.atan2(q.imag, q.real) if q != 0 else 0.0
            if np.dot((z * complex(math.cos(phi), -math.sin(phi))).real, z.real) < 0:
                phi += math.pi
            self._phase = phi
        return self._phase

    def sheet_vectors(self, k: int, d_mm: float, copper: bool = False) -> tuple:
        """Real (Kx, Ky) at the nodes of layer k at the dominant phase (peec.py:1225-1230)."""
        kx, ky = self.node_field(k, d_mm, copper)
        turn = complex(math.cos(self.dominant_phase()), -math.sin(self.dominant_phase()))
        return (kx * turn).real, (ky * turn).real

    def layer_share(self, masks: Optional[list] = None) -> np.ndarray:
        """Each layer's fraction of the summed strip-current magnitude (peec.py:1232-1245); masks: per layer a pair
        of boolean arrays (x strips, y strips) of the strips to count."""
        ax, ay = np.abs(self.I_x), np.abs(self.I_y)
        if masks is not None:
            ax = ax * np.stack([m[0] for m in masks])
            ay = ay * np.stack([m[1] for m in masks])
        per = ax.reshape(len(ax), -1).sum(axis=1) + ay.reshape(len(ay), -1).sum(axis=1)
        tot = per.sum()
        return per / tot if tot > 0 else per
