"""Oracle restatement of the plane-current PEEC solve in `src/groundtruth/solver/peec.py` (numpy + scipy).

TEST INFRASTRUCTURE (see oracle/__init__.py): only tests, smoke() and bench.py's cpu_baseline use it.

Scope: one sheet per layer (PlaneStack(..., screen=False), the Simulate default, peec.py:1561 / simulate.py:318),
"diag" preconditioner, strips of any measured width (`_widths`), lumped branches; no plane-pair capacitance. The classes take the same duck
types the reference's do: layers with `.z`, `.thickness`, `.grid` (`.nx`, `.ny`, `.d` in mm, `.node`, `.xb`, `.yb`),
branches with `.a`, `.b` ((layer, i, j) tuples), `.z(f)`, `.kind`, `.radius`, trace segments with `.a`, `.b` (m),
`.width`, `.i`.

The FFT product uses the reference's numpy multiply-add (its numba kernel is absent here, peec.py:2069-2074) and
scipy's `fft2` / `ifft2` on `next_fast_len` sizes; the Krylov basis is complex128 with BLAS `zgemv` CGS
(peec.py:2889-2894); the Schur complement is factorised by SuperLU with the reference's options (peec.py:2423-2424).
Each function cites the reference lines it follows.
"""
from __future__ import annotations

import copy
import math
from dataclasses import dataclass, field
from typing import Callable, Optional

import numpy as np
import scipy.sparse as sp
from scipy.fft import fft2, ifft2, next_fast_len
from scipy.linalg import solve_triangular
from scipy.linalg.blas import zgemv
from scipy.sparse.csgraph import connected_components
from scipy.sparse.linalg import splu

from .kernel import (GMD_FLAT, filament_mutual, leggauss, segment_mutual, segment_vector_potential,
                     skin_thickness, slab_impedance, strip_kernel, strip_self, strip_self_numeric)

SIGMA_CU = 5.8e7                # peec.py:67
VIA_PLATING = 25e-6             # peec.py:68
GMRES_MEMORY = 4e8              # peec.py:92
GMRES_RESTART_MIN = 100         # peec.py:93
GMRES_FIRST_ROWS = 16           # peec.py:97
GMRES_EXTRA_CYCLES = 24         # peec.py:98
GMRES_PROGRESS = 0.7            # peec.py:99
GMRES_CHECK_MISSES = 4          # peec.py:100
LU_PIVOT_THRESHOLD = 0.1        # peec.py:105
KERNEL_DIGITS = 12              # peec.py:3259


class SolverError(RuntimeError):
    """A model that cannot be solved (peec.py:124-127)."""


def cache_value(v: float) -> float:
    """v to 12 significant digits (peec.py:3264-3267)."""
    return float(f"{float(v):.{KERNEL_DIGITS}g}")


def via_impedance(length_m: float, drill_m: float, sigma: float = SIGMA_CU, filled: bool = False,
                  plating: float = VIA_PLATING) -> Callable[[float], complex]:
    """A barrel: R of the plating ring (or the filled hole) + jw times its filament self inductance at its radius
    (peec.py:1061-1072)."""
    r = max(drill_m / 2, 5e-5)
    L = float(filament_mutual(max(length_m, 1e-6), np.array(0.0), np.array(r)))
    area = math.pi * r * r if filled else math.pi * 2 * r * plating
    R = max(length_m, 1e-6) / (sigma * area)
    return lambda f, R=R, L=L: complex(R, 2 * math.pi * f * L)


# --------------------------------------------------------------------------- kernels
_KERNELS: dict = {}


def _kernel_spectrum(K: np.ndarray, P: tuple) -> np.ndarray:
    """Embed the centred kernel circularly (centre rolled to (0, 0)) in a P array and keep the real quadrant
    [0, P0//2] x [0, P1//2] of its FFT (peec.py:3362-3371)."""
    Z = np.zeros(P)
    Z[: K.shape[0], : K.shape[1]] = K
    Z = np.roll(Z, (-(K.shape[0] // 2), -(K.shape[1] // 2)), axis=(0, 1))
    return np.ascontiguousarray(fft2(Z, workers=-1).real[: P[0] // 2 + 1, : P[1] // 2 + 1])


def kernel_pair(nx: int, ny: int, d: float, dz: float) -> tuple:
    """(x spectrum quadrant, y spectrum quadrant, zero-offset mutual, sum of x kernel, sum of y kernel) at layer
    distance dz, for d and dz rounded to 12 digits (peec.py:3276-3306). The y kernel is the x kernel of the
    transposed grid, transposed back."""
    d, dz = cache_value(d), cache_value(dz)
    key = (nx, ny, d, dz)
    hit = _KERNELS.get(key)
    if hit is None:
        Px = (next_fast_len(2 * (nx - 1) - 1), next_fast_len(2 * ny - 1))
        Py = (next_fast_len(2 * nx - 1), next_fast_len(2 * (ny - 1) - 1))
        kx = strip_kernel(nx - 1, ny, d, dz)
        ky = strip_kernel(ny - 1, nx, d, dz).T
        hit = (_kernel_spectrum(kx, Px), _kernel_spectrum(ky, Py), float(kx[nx - 2, ny - 1]), float(kx.sum()),
               float(ky.sum()))
        _KERNELS[key] = hit
    return hit


# --------------------------------------------------------------------------- the stack
@dataclass
class Operator:
    """The system and preconditioner at one frequency (peec.py:1505-1516; Dinv kept as its diagonal)."""
    w: float
    R_strip: np.ndarray
    z_tail: np.ndarray
    Dinv: np.ndarray
    solve_S: Callable


NEAR_AXIAL = 2      # peec.py:3325


def narrow_self(d: float, phi, t):
    """What a strip d long and phi d wide adds to the full square strip's self inductance (peec.py:3309-3322)."""
    phi, t = np.asarray(phi, float), np.asarray(t, float)

    def thick(w):
        return filament_mutual(d, 0.0, GMD_FLAT * w) - filament_mutual(d, 0.0, GMD_FLAT * (w + t))
    out = strip_self_numeric(d, phi * d) - strip_self_numeric(d, d) - thick(phi * d) + thick(d)
    return float(out) if np.ndim(out) == 0 else out


def side_mutual(length: float, wa, wb, gap, steps=0.0, n: int = 8) -> np.ndarray:
    """Two flat coplanar parallel strips' mutual: n filaments across each, the mean filament mutual
    (peec.py:3328-3341)."""
    wa, wb, gap, steps = (np.atleast_1d(v) for v in np.broadcast_arrays(*(np.asarray(v, float) for v in
                                                                            (wa, wb, gap, steps))))
    wa, wb, gap = wa * length, wb * length, gap * length
    u = (np.arange(n) + 0.5) / n - 0.5
    ya = wa[:, None, None] * u[None, :, None]
    yb = gap[:, None, None] + wb[:, None, None] * u[None, None, :]
    rho = np.maximum(np.abs(yb - ya), GMD_FLAT * np.maximum(wa, wb)[:, None, None] / n)
    axial = np.broadcast_to((steps * length)[:, None, None], rho.shape)
    return filament_mutual(length, axial, rho).mean(axis=(1, 2))


class PlaneStack:
    """Plane layers plus lumped branches: topology, inductance kernels, per-frequency operators and the solve
    (peec.py:1555-2700 restricted to one sheet per layer, diag preconditioner, no capacitance)."""

    def __init__(self, layers: list, branches: Optional[list] = None, sigma: float = SIGMA_CU):
        if not layers:
            raise SolverError("no plane layers")
        g0 = layers[0].grid
        for l in layers:
            if (l.grid.nx, l.grid.ny, l.grid.d) != (g0.nx, g0.ny, g0.d):
                raise SolverError("all layers must share one grid")
        self.layers = list(layers)
        self.N = len(layers)
        self.n_sheets = self.N
        self.nx, self.ny = g0.nx, g0.ny
        self.d = g0.d * 1e-3
        self.zs = np.array([l.z for l in layers], float)
        self.sigma = sigma
        self.sheet_z = self.zs.copy()                                   # peec.py:1758-1765 without splits
        self.sheet_t = np.array([l.thickness for l in layers], float)
        self._topology(list(branches or []))
        self._kernels()
        self._widths()
        self._ops: dict = {}

    # ---------------------------------------------------------------- narrow strips (peec.py:1732-1822)
    def _widths(self):
        """strip_width, r_width, dL and M_near from the grids' wx / wy, rx / ry, ox / oy (one sheet per layer: no
        face twins, peec.py:1823-1843 does nothing)."""
        def gather(attr, idx, shape, default):
            if len(idx) == 0:
                return np.zeros(0)
            if all(getattr(l.grid, attr, None) is None for l in self.layers):
                return np.full(len(idx), default)
            w = np.stack([getattr(l.grid, attr) if getattr(l.grid, attr, None) is not None
                          else np.full(shape, default) for l in self.layers])
            return w[tuple(idx.T)]
        sx, sy = (self.nx - 1, self.ny), (self.nx, self.ny - 1)
        lx, ly = self.xa, self.ya
        w = np.concatenate([gather("wx", lx, sx, 1.0), gather("wy", ly, sy, 1.0)])
        ok = w > 0
        self.strip_width = np.where(ok, w, 1.0)
        rw = np.concatenate([gather("rx", lx, sx, 0.0), gather("ry", ly, sy, 0.0)])
        self.r_width = np.where(ok & (rw > 0), rw, self.strip_width)
        off = np.where(ok, np.concatenate([gather("ox", lx, sx, 0.0), gather("oy", ly, sy, 0.0)]), 0.0)
        self.dL = self.M_near = None
        if not np.any(self.strip_width != 1.0) and not np.any(self.r_width != 1.0):
            return
        narrow = self.strip_width < 1.0
        if not np.any(narrow):
            return
        d = self.d
        t = np.array([l.thickness for l in self.layers])[np.concatenate([lx[:, 0], ly[:, 0]])]
        dL = np.zeros(self.ns)
        keys, inv = np.unique(np.round(np.column_stack([self.strip_width[narrow], t[narrow]]), 12), axis=0,
                              return_inverse=True)
        dL[narrow] = narrow_self(d, keys[:, 0], keys[:, 1])[np.ravel(inv)]
        self.dL = dL
        rows, cols, steps = [], [], []
        for idx, shape, lo, axis in ((self.xa, sx, 0, 2), (self.ya, sy, self.nxb, 1)):
            if len(idx) == 0:
                continue
            at = -np.ones((self.n_sheets, *shape), int)
            at[tuple(idx.T)] = lo + np.arange(len(idx))
            n_ax = shape[0] if axis == 2 else shape[1]
            for k in range(-NEAR_AXIAL, NEAR_AXIAL + 1):
                if abs(k) >= n_ax:
                    continue
                ax_a = slice(max(-k, 0), n_ax - max(k, 0))
                ax_b = slice(max(k, 0), n_ax - max(-k, 0))
                a = at[:, ax_a, :-1] if axis == 2 else at[:, :-1, ax_a]
                b = at[:, ax_b, 1:] if axis == 2 else at[:, 1:, ax_b]
                both = (a >= 0) & (b >= 0)
                a, b = a[both], b[both]
                part = narrow[a] | narrow[b]
                rows.append(a[part])
                cols.append(b[part])
                steps.append(np.full(int(part.sum()), abs(k)))
        a, b = np.concatenate(rows), np.concatenate(cols)
        if len(a) == 0:
            return
        k = np.concatenate(steps)
        wa, wb = self.strip_width[a], self.strip_width[b]
        gap = 1.0 + off[b] - off[a]
        keys, inv = np.unique(np.round(np.column_stack([wa, wb, gap, k]), 9), axis=0, return_inverse=True)
        dM = (side_mutual(d, keys[:, 0], keys[:, 1], keys[:, 2], keys[:, 3])
              - side_mutual(d, np.ones(len(keys)), np.ones(len(keys)), np.ones(len(keys)), keys[:, 3]))
        M = sp.coo_matrix((dM[np.ravel(inv)], (a, b)), shape=(self.ns, self.ns))
        self.M_near = (M + M.T).tocsr()

    # ---------------------------------------------------------------- topology (peec.py:1596-1656, 1845-1862)
    def _topology(self, branches: list):
        node = np.stack([l.grid.node for l in self.layers]).astype(bool)
        self.xb = np.stack([l.grid.xb for l in self.layers]).astype(bool) & node[:, :-1, :] & node[:, 1:, :]
        self.yb = np.stack([l.grid.yb for l in self.layers]).astype(bool) & node[:, :, :-1] & node[:, :, 1:]
        self.node = node
        self.nid = -np.ones(node.shape, int)
        at = np.nonzero(node)
        self.nid[at] = np.arange(len(at[0]))
        self.n_grid = len(at[0])
        # the parts' lids (nodes with a `lid` number: no grid node) after the grid's, by their numbers
        # (peec.py:1613-1616)
        lids = sorted({e.lid for b in branches for e in (b.a, b.b) if _is_lid(e)})
        self.lid_nid = {k: self.n_grid + n for n, k in enumerate(lids)}
        self.n_nodes = self.n_grid + len(lids)
        self.xa, self.ya = np.argwhere(self.xb), np.argwhere(self.yb)
        self.nxb, self.nyb = len(self.xa), len(self.ya)
        self.ns = self.nxb + self.nyb
        self.branches = [b for b in branches if self.node_id(b.a) >= 0 and self.node_id(b.b) >= 0
                         and self.node_id(b.a) != self.node_id(b.b)]
        self.nbr = len(self.branches)
        self.ncap = 0
        self.nb = self.ns + self.nbr
        xa, ya = self.xa, self.ya
        self.a_nodes = np.concatenate([self.nid[tuple(xa.T)], self.nid[tuple(ya.T)],
                                       np.array([self.node_id(b.a) for b in self.branches], int)]).astype(int)
        self.b_nodes = np.concatenate([self.nid[xa[:, 0], xa[:, 1] + 1, xa[:, 2]],
                                       self.nid[ya[:, 0], ya[:, 1], ya[:, 2] + 1],
                                       np.array([self.node_id(b.b) for b in self.branches], int)]).astype(int)
        r = np.arange(self.nb)
        self.A = sp.csr_matrix((np.concatenate([np.ones(self.nb), -np.ones(self.nb)]),
                                (np.concatenate([r, r]), np.concatenate([self.a_nodes, self.b_nodes]))),
                               shape=(self.nb, self.n_nodes))
        n = self.n_nodes
        g = sp.coo_matrix((np.ones(self.nb), (self.a_nodes, self.b_nodes)), shape=(n, n))
        _, self.full_label = connected_components(g, directed=False)
        self.cond_label = self.full_label
        _, first = np.unique(self.full_label, return_index=True)
        self.keep = np.ones(n, bool)
        self.keep[first] = False
        self.Ar = self.A[:, self.keep].tocsc()
        self.ArT = self.Ar.T.tocsr()

    def node_id(self, node) -> int:
        """Node number of (layer, i, j), -1 off copper; a lid's its own after the grid's, -1 when no branch names it
        (peec.py:1864-1869)."""
        if _is_lid(node):
            return self.lid_nid.get(node.lid, -1)
        return int(self.nid[tuple(node)])

    # ---------------------------------------------------------------- what-if stacks (peec.py:1871-1923)
    def with_branches(self, extra: list) -> "PlaneStack":
        """This stack plus a few branches (peec.py:1871-1909): a shallow copy with the branch list and incidence
        extended, the kernels shared; a branch joining two islands makes a stack of its own. Its operators factorise
        the extended S afresh (the reference reuses the base factor through Woodbury: the same S^-1)."""
        extra = [b for b in extra if self.node_id(b.a) >= 0 and self.node_id(b.b) >= 0
                 and self.node_id(b.a) != self.node_id(b.b)]
        if not extra:
            return self
        a = np.array([self.node_id(b.a) for b in extra])
        b = np.array([self.node_id(b.b) for b in extra])
        if np.any(self.full_label[a] != self.full_label[b]):
            return PlaneStack(self.layers, self.branches + extra, self.sigma)
        new = copy.copy(self)
        new.branches = self.branches + extra
        new.nbr = self.nbr + len(extra)
        new.nb = self.nb + len(extra)
        new.a_nodes = np.concatenate([self.a_nodes, a])
        new.b_nodes = np.concatenate([self.b_nodes, b])
        r = np.arange(self.nb, new.nb)
        rows = sp.csr_matrix((np.concatenate([np.ones(len(r)), -np.ones(len(r))]),
                              (np.concatenate([r - self.nb] * 2), np.concatenate([a, b]))),
                             shape=(len(r), self.n_nodes))
        new.A = sp.vstack([self.A, rows]).tocsr()
        new.Ar = new.A[:, self.keep].tocsc()
        new.ArT = new.Ar.T.tocsr()
        new._base = self
        new._ops = {}
        return new

    def map_state(self, x, stack: "PlaneStack"):
        """A solution vector of `stack` in this stack's unknowns: the same stack, or this stack's base, whose added
        branches start at zero current (peec.py:1911-1923)."""
        if x is None:
            return None
        if stack is self:
            return x
        if getattr(self, "_base", None) is stack and len(x) == stack.nb + stack.Ar.shape[1]:
            return np.concatenate([x[:stack.nb], np.zeros(self.nb - stack.nb, complex), x[stack.nb:]])
        return None

    # ---------------------------------------------------------------- inductance (peec.py:1925-1988, 2054-2135)
    def face_gap(self, s: int, t: int) -> float:
        """Distance between the faces two sheets turn to each other, at least 1 um (peec.py:1925-1932)."""
        if s == t:
            return 0.0
        gap = abs(self.sheet_z[s] - self.sheet_z[t]) - 0.5 * (self.sheet_t[s] + self.sheet_t[t])
        return max(gap, 1e-6)

    def _kernels(self):
        nx, ny, d, N = self.nx, self.ny, self.d, self.n_sheets
        self.Px = (next_fast_len(2 * (nx - 1) - 1), next_fast_len(2 * ny - 1))
        self.Py = (next_fast_len(2 * nx - 1), next_fast_len(2 * (ny - 1) - 1))
        Kx = np.empty((self.Px[0] // 2 + 1, self.Px[1] // 2 + 1, N, N))
        Ky = np.empty((self.Py[0] // 2 + 1, self.Py[1] // 2 + 1, N, N))
        self.L_self = kernel_pair(nx, ny, d, 0.0)[2]
        for s in range(N):
            for t in range(s, N):
                kx, ky, _, _, _ = kernel_pair(nx, ny, d, self.face_gap(s, t))
                Kx[:, :, s, t] = Kx[:, :, t, s] = kx
                Ky[:, :, s, t] = Ky[:, :, t, s] = ky
        self._kq = {"x": Kx, "y": Ky}

    def kernel(self, orient: str, t: int, s: int) -> np.ndarray:
        """Full (P0, P1) spectrum unfolded from the quadrant: row i reads quadrant row min(i, P0 - i)
        (peec.py:2083-2090)."""
        Kq, P = (self._kq["x"], self.Px) if orient == "x" else (self._kq["y"], self.Py)
        i = np.arange(P[0])
        j = np.arange(P[1])
        return Kq[np.minimum(i, P[0] - i)[:, None], np.minimum(j, P[1] - j)[None, :], t, s]

    def _conv(self, X: np.ndarray, orient: str) -> np.ndarray:
        """sum_s K[t, s] (*) X[s] on the padded grid: fft2, numpy multiply-add, ifft2 (peec.py:2054-2081)."""
        P = self.Px if orient == "x" else self.Py
        F = fft2(X, s=P, axes=(1, 2), workers=-1)
        out = np.empty_like(F)
        tmp = np.empty_like(F[0])
        for t in range(self.n_sheets):
            np.multiply(self.kernel(orient, t, 0), F[0], out=out[t])
            for s in range(1, self.n_sheets):
                np.multiply(self.kernel(orient, t, s), F[s], out=tmp)
                out[t] += tmp
        return ifft2(out, axes=(1, 2), workers=-1, overwrite_x=True)

    def _strip_index(self, orient: str):
        """Flat indices of the strips in the dense current array and in the padded result (peec.py:2102-2113)."""
        S = self.n_sheets
        a, shape, P = ((self.xa, (S, self.nx - 1, self.ny), self.Px) if orient == "x"
                       else (self.ya, (S, self.nx, self.ny - 1), self.Py))
        return np.ravel_multi_index(tuple(a.T), shape), np.ravel_multi_index(tuple(a.T), (S, *P))

    def Lp_matvec(self, I_s: np.ndarray) -> np.ndarray:
        """L_p I for the strip currents: scatter per orientation, convolve, gather (peec.py:2115-2135)."""
        out = np.empty(self.ns, complex)
        S = self.n_sheets
        for orient, lo, hi, shape in (("x", 0, self.nxb, (S, self.nx - 1, self.ny)),
                                      ("y", self.nxb, self.ns, (S, self.nx, self.ny - 1))):
            if hi > lo:
                into, outof = self._strip_index(orient)
                X = np.zeros(shape[0] * shape[1] * shape[2], complex)
                X[into] = I_s[lo:hi]
                out[lo:hi] = self._conv(X.reshape(shape), orient).reshape(-1)[outof]
        if self.dL is not None:             # strips narrower than the pitch (peec.py:2129-2132)
            out += self.dL * I_s
            if self.M_near is not None:
                out += self.M_near @ I_s
        return out

    # ---------------------------------------------------------------- operators (peec.py:2144-2178, 2345-2511)
    def _impedances(self, f: float):
        """(w, strip sheet impedance per strip, branch impedances) (peec.py:2150-2178, no capacitance)."""
        w = 2 * math.pi * f
        zs = np.array([slab_impedance(l.thickness, f, self.sigma) for l in self.layers])
        R_strip = (np.concatenate([zs[self.xa[:, 0]], zs[self.ya[:, 0]]]) / self.r_width if self.ns
                   else np.zeros(0, complex))
        z_br = np.array([complex(b.z(f)) for b in self.branches], complex)
        return w, R_strip, z_br

    def admittance(self, w: float, R_strip: np.ndarray, z_tail: np.ndarray) -> np.ndarray:
        """diag(Z)^-1 of the diag preconditioner: strips with their own self inductance only
        (peec.py:2465-2478)."""
        L_own = self.L_self if self.dL is None else self.L_self + self.dL
        return np.concatenate([1.0 / (R_strip + 1j * w * L_own), 1.0 / z_tail])

    def operator(self, f: float) -> Operator:
        """The operator at f with its own SuperLU factorisation of S = Ar^T D^-1 Ar (peec.py:2227-2445: a fresh
        factorisation per frequency key; the reference borrows only across rounds, peec.py:2303-2314)."""
        key = cache_value(f)
        op = self._ops.get(key)
        if op is None:
            w, R_strip, z_tail = self._impedances(key)
            Dinv = self.admittance(w, R_strip, z_tail)
            S = (self.ArT @ sp.diags(Dinv) @ self.Ar).tocsc()
            lu = splu(S, permc_spec="MMD_AT_PLUS_A", diag_pivot_thresh=LU_PIVOT_THRESHOLD,
                      options={"SymmetricMode": True})
            op = Operator(w, R_strip, z_tail, Dinv, lu.solve)
            self._ops[key] = op
        return op

    def solve(self, f: float, rhs_top: np.ndarray, J: np.ndarray, tol: float = 1e-5, max_cycles: int = 8,
              guess: Optional[list] = None, restart: Optional[int] = None):
        """K [I; V] = [rhs_top; J[keep]] by right-preconditioned GMRES, with the least-squares warm start over
        `guess` (peec.py:2513-2607, own factorisation: no borrow branch). `restart` sets the GMRES restart length
        (None: the reference's gmres_restart(n))."""
        op = self.operator(f)
        m = HostMath(self)
        nb = self.nb
        rhs = np.concatenate([rhs_top.astype(complex), J[self.keep].astype(complex)])
        bnorm = float(np.linalg.norm(rhs))
        x0, r0 = None, rhs
        G = [g for g in (guess or []) if g is not None and len(g) == len(rhs) and np.any(g)]
        if G:
            KG = np.stack([m.matvec(op, g) for g in G], axis=1)
            c, *_ = np.linalg.lstsq(KG, rhs, rcond=None)
            r = rhs - KG @ c
            if np.linalg.norm(r) < 0.9 * bnorm:
                x0, r0 = np.stack(G, axis=1) @ c, r
        y, its, ok = gmres(lambda v: m.matvec(op, m.precond(op, v)), r0, tol, max_cycles=max_cycles, bnorm=bnorm,
                           restart=restart, extra_cycles=GMRES_EXTRA_CYCLES)
        if not ok:
            raise SolverError(f"GMRES did not converge at {f:.4g} Hz ({its} iterations)")
        x = m.precond(op, y) if y.any() else np.zeros(len(rhs), complex)
        if x0 is not None:
            x = x + x0
        V = np.zeros(self.n_nodes, complex)
        V[self.keep] = x[nb:]
        return x[:nb], V, its, x

    # ---------------------------------------------------------------- coupling to the signal (peec.py:2609-2700)
    def pos(self, node) -> np.ndarray:
        k, i, j = node
        return np.array([i * self.d, j * self.d, self.zs[k]])

    def strip_coupling(self, trace: list) -> np.ndarray:
        """psi = int A.dl along every strip per amp of signal: the segment's vector potential at the strip's centre
        on its face toward the segment, averaged over 3x3 Gauss points of the strip's square within 3 pitches
        (+ 2 widths) of the segment (peec.py:2609-2656)."""
        d = self.d
        x, w = leggauss(3)
        q = 0.5 * x
        W = np.outer(w, w).ravel() / 4.0
        qa, ql = (g.ravel() for g in np.meshgrid(q, q, indexing="ij"))
        horiz = [s for s in trace if np.linalg.norm((s.b - s.a)[:2]) > 0]
        psi = np.zeros(self.ns)
        for axis, idx, off in ((0, self.xa, 0), (1, self.ya, self.nxb)):
            if len(idx) == 0:
                continue
            k, i, j = idx[:, 0], idx[:, 1], idx[:, 2]
            centres = np.stack([i * d + (d / 2 if axis == 0 else 0.0), j * d + (d / 2 if axis == 1 else 0.0),
                                self.sheet_z[k]], axis=1)
            acc = np.zeros(len(idx))
            half = 0.5 * self.sheet_t[k]
            for s in horiz:
                ab = s.b - s.a
                L = np.linalg.norm(ab)
                if abs(ab[axis]) < 1e-15 * L:
                    continue
                zseg = 0.5 * (s.a[2] + s.b[2])
                centres[:, 2] = self.sheet_z[k] + np.where(zseg > self.sheet_z[k], half, -half)
                rho_min = max(0.5 * s.width, 1e-7)
                a_c = segment_vector_potential(s.a, s.b, centres, rho_min)
                near = dist_to_segment(centres, s.a, s.b) < 3 * d + 2 * s.width
                if near.any():
                    c = centres[near]
                    da, dl = (qa, ql) if axis == 0 else (ql, qa)
                    pts = np.stack([c[:, 0, None] + da * d, c[:, 1, None] + dl * d,
                                    np.broadcast_to(c[:, 2, None], (len(c), len(W)))], axis=2)
                    a_q = segment_vector_potential(s.a, s.b, pts.reshape(-1, 3), rho_min).reshape(len(c), len(W))
                    a_c[near] = (a_q * W).sum(1)
                acc += (s.i * ab[axis] / L) * a_c
            psi[off: off + len(idx)] = d * acc
        return psi

    def branch_coupling(self, trace: list, start: int = 0) -> np.ndarray:
        """psi of via branches from the vertical trace segments (peec.py:2658-2666)."""
        vert = [s for s in trace if abs((s.b - s.a)[2]) > 0 and np.linalg.norm((s.b - s.a)[:2]) < 1e-12]
        out = np.zeros(self.nbr - start)
        for n, b in enumerate(self.branches[start:]):
            if b.kind == "via" and vert:
                pa, pb = self.pos(b.a), self.pos(b.b)
                out[n] = sum(s.i * segment_mutual(s.a, s.b, pa, pb, b.radius) for s in vert)
        return out

    def moment(self, trace: list, Ib: np.ndarray) -> np.ndarray:
        """m = 1/2 sum r x I dl over the closed loop (peec.py:2668-2700, no capacitance)."""
        m = np.zeros(3, complex)
        for s in trace:
            m += 0.5 * s.i * np.cross(s.a, s.b)
        d = self.d
        Ix, Iy = Ib[: self.nxb], Ib[self.nxb: self.ns]
        if self.nxb:
            k, _, j = self.xa.T
            z = self.sheet_z[k]
            m[1] += 0.5 * np.sum(z * d * Ix)
            m[2] += 0.5 * np.sum(-(j * d) * d * Ix)
        if self.nyb:
            k, i, _ = self.ya.T
            z = self.sheet_z[k]
            m[0] += 0.5 * np.sum(-z * d * Iy)
            m[2] += 0.5 * np.sum((i * d) * d * Iy)
        for b, cur in zip(self.branches, Ib[self.ns:]):
            m += 0.5 * np.cross(self.pos(b.a), self.pos(b.b)) * cur
        return m


# --------------------------------------------------------------------------- the CPU math (peec.py:2841-3083)
class HostVectors:
    """GMRES vector work on complex128 numpy arrays (peec.py:2841-2897)."""

    @staticmethod
    def zeros(n):
        return np.zeros(n, complex)

    @staticmethod
    def basis(rows, n):
        return np.empty((rows, n), complex)

    @staticmethod
    def put(V, j, v):
        V[j] = v

    @staticmethod
    def row(V, j):
        return V[j]

    def grow(self, V, rows, used):
        out = self.basis(rows, V.shape[1])
        out[:used] = V[:used]
        return out

    @staticmethod
    def copy(v):
        return v.astype(complex)

    @staticmethod
    def norm(v) -> float:
        return float(np.linalg.norm(v))

    @staticmethod
    def any(v) -> bool:
        return bool(v.any())

    @staticmethod
    def project(V, j, w):
        """(V^H w, w - V V^H w) over rows 0..j by zgemv (peec.py:2889-2894)."""
        Vt = V[: j + 1].T
        h = zgemv(1.0, Vt, w, trans=2)
        return h, zgemv(-1.0, Vt, h, beta=1.0, y=w, overwrite_y=1)

    @staticmethod
    def combine(y, V, k):
        return y @ V[:k]


class HostMath:
    """The system product and the preconditioner in numpy / scipy (peec.py:3023-3083, no face coupling)."""

    def __init__(self, stack: PlaneStack):
        self.st = stack
        self.vec = HostVectors()

    def matvec(self, op: Operator, x: np.ndarray) -> np.ndarray:
        """[Z I - Ar V; Ar^T I] with Z I = R I + jw L_p I on the strips, z I on the branches (peec.py:3043-3062)."""
        st = self.st
        nb, ns = st.nb, st.ns
        Ib = x[:nb]
        top = np.empty(nb, complex)
        top[:ns] = op.R_strip * Ib[:ns] + 1j * op.w * st.Lp_matvec(Ib[:ns])
        top[ns:] = op.z_tail * Ib[ns:]
        top -= st.Ar @ x[nb:]
        return np.concatenate([top, st.ArT @ Ib])

    def precond(self, op: Operator, r: np.ndarray) -> np.ndarray:
        """Block elimination with Z -> D: y = S^-1 (r2 - Ar^T D^-1 r1), x1 = D^-1 (r1 + Ar y) (peec.py:3064-3083)."""
        st = self.st
        r1, r2 = r[:st.nb], r[st.nb:]
        y = op.solve_S(r2 - st.ArT @ (op.Dinv * r1))
        return np.concatenate([op.Dinv * (r1 + st.Ar @ y), y])


def gmres_restart(n: int) -> int:
    """min(200, max(100, 4e8 / (16 n))) (peec.py:3087-3090)."""
    return int(min(200, max(GMRES_RESTART_MIN, GMRES_MEMORY // (16 * n))))


def gmres(op: Callable, b, tol: float, max_cycles: int = 8, restart: Optional[int] = None,
          bnorm: Optional[float] = None, max_its: Optional[int] = None, vec=None, extra_cycles: int = 0,
          check_misses: int = GMRES_CHECK_MISSES) -> tuple:
    """Restarted GMRES with CGS + DGKS re-pass (below 0.7 of the product's norm), Givens-rotated Hessenberg,
    true-residual check when the recurrence meets the aim (halved after a miss; after `check_misses` misses the
    cycle restarts from the check), a basis growing from 16 rows by doubling, breakdown at hn <= 1e-14 |h|, and
    extra restarts while each still falls to GMRES_PROGRESS (peec.py:3093-3224). Returns (x, its, converged)."""
    if vec is None:
        vec = HostVectors()
    n = len(b)
    if bnorm is None:
        bnorm = vec.norm(b)
    x = vec.zeros(n)
    if bnorm == 0 or vec.norm(b) <= tol * bnorm:
        return x, 0, True
    if restart is None:
        restart = gmres_restart(n)
    its = 0
    r = vec.copy(b)
    last = math.inf
    V = None
    for cycle in range(max_cycles + extra_cycles):
        beta = vec.norm(r)
        if beta <= tol * bnorm:
            return x, its, True
        if max_its is not None and its >= max_its:
            return x, its, False
        if cycle >= max_cycles and beta > GMRES_PROGRESS * last:
            return x, its, False
        last = beta
        if V is None:
            V = vec.basis(min(restart + 1, GMRES_FIRST_ROWS), n)
        H = np.zeros((restart + 1, restart), complex)
        cs = np.zeros(restart)
        sn = np.zeros(restart, complex)
        g = np.zeros(restart + 1, complex)
        g[0] = beta
        vec.put(V, 0, r / beta)
        k = 0
        aim = tol * bnorm
        misses = 0
        restarted = False
        for j in range(restart):
            w = op(vec.row(V, j))
            w_norm = vec.norm(w)
            h, w = vec.project(V, j, w)
            hn = vec.norm(w)
            if hn < 0.7 * w_norm:
                h2, w = vec.project(V, j, w)
                h = h + h2
                hn = vec.norm(w)
            col = np.empty(j + 2, complex)
            col[: j + 1] = h
            col[j + 1] = hn
            for i in range(j):
                a, c = col[i], col[i + 1]
                col[i] = cs[i] * a + sn[i] * c
                col[i + 1] = -np.conj(sn[i]) * a + cs[i] * c
            a = col[j]
            rho = math.hypot(abs(a), hn)
            if rho == 0.0:
                cs[j], sn[j] = 1.0, 0.0
            elif abs(a) == 0.0:
                cs[j], sn[j] = 0.0, 1.0
                col[j] = hn
            else:
                phase = a / abs(a)
                cs[j], sn[j] = abs(a) / rho, phase * hn / rho
                col[j] = phase * rho
            col[j + 1] = 0.0
            H[: j + 2, j] = col
            g[j + 1] = -np.conj(sn[j]) * g[j]
            g[j] = cs[j] * g[j]
            its += 1
            k = j + 1
            if hn <= 1e-14 * np.linalg.norm(h) or (max_its is not None and its >= max_its):
                break
            if abs(g[j + 1]) <= aim:
                y = solve_triangular(H[:k, :k], g[:k])
                xt = x + vec.combine(y, V, k)
                rt = b - op(xt)
                if vec.norm(rt) <= tol * bnorm:
                    return xt, its, True
                misses += 1
                if misses >= check_misses:
                    x, r = xt, rt
                    restarted = True
                    break
                aim = 0.5 * abs(g[j + 1])
            if j + 1 >= len(V):
                V = vec.grow(V, min(restart + 1, 2 * len(V)), j + 1)
            vec.put(V, j + 1, w / hn)
        if restarted:
            continue
        while k and H[k - 1, k - 1] == 0:
            k -= 1
        y = solve_triangular(H[:k, :k], g[:k]) if k else np.zeros(0, complex)
        x = x + vec.combine(y, V, k)
        r = b - op(x)
    return x, its, vec.norm(r) <= tol * bnorm


# --------------------------------------------------------------------------- the route (peec.py:2703-2817)
@dataclass
class Solution:
    """One solve per amp of loop current (peec.py:1092-1110, the fields the device solver fills)."""
    f: float
    I_x: np.ndarray
    I_y: np.ndarray
    V: np.ndarray
    I_branch: np.ndarray
    Z_loop: complex
    m: np.ndarray
    iterations: int
    x: Optional[np.ndarray] = field(default=None, repr=False)
    V_lid: Optional[np.ndarray] = None      # the lids' potentials, in their numbers' order (peec.py:1110-1112)

    @property
    def L_loop(self) -> float:
        return self.Z_loop.imag / (2 * math.pi * self.f)


def _is_lid(node) -> bool:
    """A part's lid (peec.py:1009-1038): a (layer, i, j) tuple that carries its own `lid` number."""
    return getattr(node, "lid", None) is not None


def dist_to_segment(P: np.ndarray, a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """Distance of each point to the segment a-b (peec.py:3374-3382)."""
    ab = b - a
    L2 = float(ab @ ab)
    if L2 == 0:
        return np.linalg.norm(P - a, axis=1)
    t = np.clip((P - a) @ ab / L2, 0, 1)
    return np.linalg.norm(P - (a + t[:, None] * ab), axis=1)


def trace_self(trace: list, trace_t: float) -> float:
    """Partial self inductance of the signal path: sum_ij I_i I_j M_ij (peec.py:3227-3237)."""
    segs = [s for s in trace if np.linalg.norm(s.b - s.a) > 0]
    L = 0.0
    for i, si in enumerate(segs):
        L += si.i * si.i * strip_self(np.linalg.norm(si.b - si.a), si.width, trace_t)
        for sj in segs[i + 1:]:
            L += 2 * si.i * sj.i * segment_mutual(si.a, si.b, sj.a, sj.b, GMD_FLAT * (si.width + trace_t))
    return L


def trace_resistance(trace: list, f: float, sigma: float, trace_t: float) -> float:
    """I^2-weighted skin-effect resistance of the horizontal segments (peec.py:3240-3250)."""
    R = 0.0
    for s in trace:
        L = np.linalg.norm(s.b - s.a)
        if L == 0 or abs(s.b[2] - s.a[2]) > 0:
            continue
        R += s.i * s.i * L / (sigma * s.width * skin_thickness(trace_t, f, sigma))
    return R


class RouteExcitation:
    """A known trace current plus injections on a stack (peec.py:2703-2817)."""

    def __init__(self, stack: PlaneStack, trace: list, inject_driver=None, inject_receiver=None,
                 trace_thickness: float = 35e-6, injections: Optional[list] = None):
        self.stack = stack
        self.trace = trace
        self.trace_t = trace_thickness
        self.inject = (inject_driver, inject_receiver)
        self.extra = list(injections or [])
        terms = []
        if inject_driver is not None:
            terms.append((inject_driver, -1.0, "driver return node"))
        if inject_receiver is not None:
            terms.append((inject_receiver, 1.0, "receiver return node"))
        terms += [(node, float(amps), f"current injection at {node}") for node, amps in (injections or [])]
        if not terms:
            raise SolverError("no current is injected")
        ids = []
        for node, amps, what in terms:
            n = stack.node_id(node)
            if n < 0:
                raise SolverError(f"{what} is not on plane copper")
            ids.append((int(n), amps))
        balance: dict = {}
        for n, amps in ids:
            c = int(stack.full_label[n])
            balance[c] = balance.get(c, 0.0) + amps
        if any(abs(v) > 1e-9 for v in balance.values()):
            if inject_driver is not None and inject_receiver is not None and not injections:
                raise SolverError("receiver return node is not connected to the driver's return node")
            raise SolverError("the injected currents do not balance on every connected piece of copper")
        self.injections = ids
        self.strip_psi = stack.strip_coupling(trace)
        self.branch_psi = stack.branch_coupling(trace)
        self.psi = np.concatenate([self.strip_psi, self.branch_psi])
        self.L_trace = trace_self(trace, trace_thickness) if trace else 0.0

    def on(self, stack: PlaneStack) -> "RouteExcitation":
        """The same loop on a what-if stack (peec.py:2771-2774); the coupling is recomputed (the same strip values)."""
        return RouteExcitation(stack, self.trace, *self.inject, self.trace_t, self.extra)

    def rhs(self, f: float):
        """(rhs_top = -jw psi, J) of this loop at f (peec.py:2789-2795)."""
        w = 2 * math.pi * f
        J = np.zeros(self.stack.n_nodes)
        for n, amps in self.injections:
            J[n] += amps
        return -1j * w * self.psi, J

    def solve(self, f: float, guess: Optional[list] = None, tol: float = 1e-5, max_cycles: int = 8,
              restart: Optional[int] = None) -> Solution:
        """Solve and unpack I_x, I_y, V, branch currents, Z_loop and the moment (peec.py:2787-2817). `tol`,
        `max_cycles` and `restart` are passed to PlaneStack.solve (the reference uses its defaults there)."""
        st = self.stack
        rhs_top, J = self.rhs(f)
        Ib, V, its, x = st.solve(f, rhs_top, J, tol=tol, max_cycles=max_cycles, guess=guess, restart=restart)
        return self.unpack(f, Ib, V, its, x)

    def unpack(self, f: float, Ib: np.ndarray, V: np.ndarray, its: int, x=None) -> Solution:
        """Solution from the branch currents and node potentials (peec.py:2796-2817)."""
        st = self.stack
        w = 2 * math.pi * f
        I_x = np.zeros((st.N, st.nx - 1, st.ny), complex)
        I_y = np.zeros((st.N, st.nx, st.ny - 1), complex)
        I_x[tuple(st.xa.T)] = Ib[:st.nxb]
        I_y[tuple(st.ya.T)] = Ib[st.nxb: st.ns]
        Vg = np.zeros((st.N, st.nx, st.ny), complex)
        Vg[np.nonzero(st.node)] = V[:st.n_grid]
        v_inj = sum(amps * V[n] for n, amps in self.injections)
        Z = complex(trace_resistance(self.trace, f, st.sigma, self.trace_t) + v_inj
                    + 1j * w * (self.L_trace + np.dot(self.psi, Ib)))
        V_lid = V[st.n_grid:].copy()
        return Solution(f, I_x, I_y, Vg, Ib[st.ns:].copy(), Z, st.moment(self.trace, Ib), its, x, V_lid)
