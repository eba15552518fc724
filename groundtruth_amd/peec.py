"""The plane-current PEEC solve behind the reference's own call signatures, on an MI355X.

Drop-in for `src/groundtruth/solver/peec.py`:
  * `PlaneStack(layers, branches, capacitance, sigma, precond, screen)` (peec.py:1555-1594) numbers the unknowns as
    the reference does (x strips, y strips, branches; nodes in (layer, i, j) order; one reference node per connected
    piece, peec.py:1596-1656 / 1845-1862), computes the inductance kernels, and makes the board resident on the
    device: kernel spectra, strip layout, incidence and the Schur factor (libgtpeec, include/gtpeec.h).
  * `PlaneStack.solve(f, rhs_top, J, tol=1e-5, max_cycles=8, guess=None) -> (Ib, V, iterations, x)`
    (peec.py:2513-2607): the same system, right-preconditioned GMRES with the same control flow, on the device.
  * `RouteExcitation(stack, trace, inject_driver, inject_receiver, ...)` and `.solve(f, guess) -> Solution`
    (peec.py:2703-2817): psi, injections, loop impedance and moment as the reference.
Errors are the reference's: SolverError for models and non-convergence, MemoryError for allocation failures,
DeviceError for a failed device (the reference then redoes the solve on its CPU path, peec.py:2549-2558).

Scope of this package: one current sheet per layer (screen=False, the Simulate default), the "diag" preconditioner,
full-width strips (one copper thickness per layer; layers may differ), lumped branches (vias). Plane-pair
capacitance, face sheets and the "sheet" preconditioner raise NotImplementedError.
"""
from __future__ import annotations

import ctypes as C
import copy
import math
from collections import OrderedDict
from typing import Optional

import numpy as np
import scipy.sparse as sp
from scipy.sparse.csgraph import connected_components

from . import _lib
from . import greens as G
from .problem import SIGMA_CU, DeviceTooSmall, LidNode, Solution, SolverError

GMRES_EXTRA_CYCLES = 24         # peec.py:98
GMRES_CHECK_MISSES = 4          # peec.py:100
SCREEN_MIN_NODES = 16           # peec.py:112
OPERATORS_KEPT = 32             # per-frequency operators kept per stack (small device arrays each)
DEVICE_MEMORY_SHARE = 0.8       # a model uses at most this share of the device's free memory (gpu.py:580)
BATCH = 16                      # right-hand sides per batched device launch
MIXED_INNER_TOL = 1e-6          # mixed precision: the smallest relative residual an FP32 cycle is asked for
MIXED_MAX_REFINE = 64           # mixed precision: FP64 restarts at most (the reference's cycle rule governs)

_KERNELS: dict = {}


def _kernel_pair(nx: int, ny: int, d: float, dz: float, px: tuple, py: tuple):
    """(x quadrant, y quadrant, zero-offset mutual) for the device FFT sizes, cached by the 12-digit keys of d, dz
    (peec.py:3276-3306)."""
    d, dz = G.key(d), G.key(dz)
    k = (nx, ny, d, dz, px, py)
    hit = _KERNELS.get(k)
    if hit is None:
        kx = G.strip_kernel(nx - 1, ny, d, dz)
        ky = G.strip_kernel(ny - 1, nx, d, dz).T
        hit = (G.quadrant(kx, *px), G.quadrant(ky, *py), float(kx[nx - 2, ny - 1]))
        _KERNELS[k] = hit
    return hit


_NEAR: dict = {}
_LP_PLANS: dict = {}            # (sheets, nx, ny, Px, Py) -> the tuned L_p launch shapes (rows, waves per class)


def _near_table(d: float, dz: float) -> np.ndarray:
    """The near-field mutuals of two square strips dz apart: tape(d, a d, l d, dz) at axial offsets a and lateral
    offsets l of 0..4 cells, as strip_kernel sets them (kernel.py:268-294), cached by the keyed d and dz."""
    k = (d, dz)
    hit = _NEAR.get(k)
    if hit is None:
        hit = _NEAR[k] = np.array([[G.tape(d, a * d, l * d, dz) for l in range(G.NEAR + 1)]
                                   for a in range(G.NEAR + 1)])
    return hit


def kernel_quadrant(stack: "PlaneStack", orient: int) -> np.ndarray:
    """The stack's resident kernel quadrant of orientation 0 (x strips) or 1 (y strips), (P1/2+1, P0/2+1, S, S)."""
    P = stack.Px if orient == 0 else stack.Py
    S = stack.n_sheets
    out = np.zeros((P[1] // 2 + 1, P[0] // 2 + 1, S, S))
    _lib.check(_lib.lib().gt_stack_kernel(stack.handle, int(orient), _lib.ptr(out)), "gt_stack_kernel")
    return out


def _woodbury_q(c: np.ndarray, G: np.ndarray) -> np.ndarray:
    """diag(c) (I + G diag(c))^-1, the Woodbury capacitance of one frequency (k x k, LAPACK on the host: a second for
    a few thousand branches)."""
    k = len(c)
    A = np.eye(k, dtype=complex) + G * np.asarray(c)[None, :]
    return np.asarray(c)[:, None] * np.linalg.inv(A)


class Operator:
    """One frequency's device operator (gt_op) and the host values it was made from (peec.py:1505-1516)."""

    def __init__(self, stack: "PlaneStack", f: float):
        self.stack = stack
        stack.handle                    # resident first: its factor gives G
        self.f = G.key(f)
        self.w = 2 * math.pi * self.f
        zs = np.array([G.slab_impedance(l.thickness, self.f, stack.sigma) for l in stack.layers])
        self.r_sheet = zs
        self.z_tail = np.array([complex(b.z(self.f)) for b in stack.branches], complex)
        self.dinv_sheet = 1.0 / (self.r_sheet + 1j * self.w * stack.L_self)
        self.dinv_tail = 1.0 / self.z_tail if len(self.z_tail) else np.zeros(0, complex)
        ds = self.dinv_sheet[0]
        k = stack.nbr
        s_in = s_out = None
        if stack.per_sheet:
            Q, k, s_in = self._node_correction(stack)
            s_out = s_in
        elif k:
            # S = d_s (B + U C U^T): B holds the factored branches at weight 1 (C = d/d_s - 1); branches a what-if
            # stack added (with_branches) are not in B (C = d/d_s)
            in_B = np.arange(k) < stack.nbr_factored
            c = self.dinv_tail / ds - in_B
            Q = _woodbury_q(c, stack.G)
        else:
            Q = np.zeros((0, 0), complex)
        self.Q = Q
        rs, zt, dsh, dt, q = (_lib.cplx(a) for a in (self.r_sheet, self.z_tail if k else np.zeros(1, complex),
                                                        self.dinv_sheet, self.dinv_tail if k else np.zeros(1, complex),
                                                        Q if k else np.zeros(1, complex)))
        desc = _lib.OpDesc()
        desc.w = self.w
        desc.r_sheet, desc.z_tail, desc.dinv_sheet, desc.dinv_tail = (_lib.ptr(a) for a in (rs, zt, dsh, dt))
        sc = 1.0 / ds
        desc.s_scale[0], desc.s_scale[1] = sc.real, sc.imag
        desc.k = k
        desc.q = _lib.ptr(q)
        if s_in is not None:
            self._s = _lib.cplx(s_in)
            desc.s_in = desc.s_out = _lib.ptr(self._s)
        h = C.c_void_p()
        _lib.check(_lib.lib().gt_op_create(stack.handle, C.byref(desc), C.byref(h)), "gt_op_create")
        self.handle = h

    def _node_correction(self, stack: "PlaneStack"):
        """Sheets of different strip admittance lambda_s (layers of different copper thickness): with Lambda the
        diagonal of lambda over the kept nodes, S = Ar^T D^-1 Ar = Lambda^1/2 (B + E) Lambda^1/2, where B is the
        factored Laplacian (strips and branches at weight 1) and E = sum_m (d_m u~ u~^T - u u^T), u~ = Lambda^-1/2 u,
        lives on the branches' end nodes (the stack's unit Woodbury columns, stack.wb_cols). Returns
        (Q = E_v (I + G E_v)^-1, its size, the per-sheet scale lambda^-1/2)."""
        lam = self.dinv_sheet
        scale = 1.0 / np.sqrt(lam)
        cols = stack.wb_cols
        k = len(cols)
        if k == 0:
            return np.zeros((0, 0), complex), 0, scale
        pos = -np.ones(stack.n_keep, np.int64)
        pos[cols] = np.arange(k)
        keep_idx = np.nonzero(stack.keep)[0]
        E = np.zeros((k, k), complex)
        for m in range(stack.nbr):
            ends = []
            for kn, sign in ((stack.a_keep[stack.ns + m], 1.0), (stack.b_keep[stack.ns + m], -1.0)):
                if kn >= 0:
                    ends.append((pos[kn], sign, scale[stack.node_layer[keep_idx[kn]]]))
            for pi, si, ci in ends:
                for pj, sj, cj in ends:
                    E[pi, pj] += self.dinv_tail[m] * si * sj * ci * cj - si * sj
        Q = E @ np.linalg.inv(np.eye(k) + stack.G @ E)
        return Q, k, scale

    def __del__(self):
        try:
            if getattr(self, "handle", None) and _lib._LIB is not None:
                _lib._LIB.gt_op_destroy(self.handle)
        except Exception:       # noqa: BLE001 - interpreter shutdown
            pass


_WOODBURY_FORMS = {"auto": 0, "stored": 1, "implicit": 2}
_ND_ORDERINGS = {"joint": 0, "sheets": 1}


class PlaneStack:
    """Plane layers and lumped branches resident on one device (peec.py:1555). Unknown order:
    [x strips | y strips | branches]."""

    def __init__(self, layers: list, branches: Optional[list] = None, capacitance: Optional[list] = None,
                 sigma: float = SIGMA_CU, precond: str = "diag", screen: bool = False, device: int = 0,
                 setup: str = "device", device_fronts: int = 256, schur_leaf: int = 32, woodbury: str = "auto",
                 schur_ordering: str = "sheets"):
        if not layers:
            raise SolverError("no plane layers")
        if precond not in ("diag", "sheet"):
            raise ValueError("precond must be 'diag' or 'sheet'")
        if precond == "sheet":
            raise NotImplementedError("the 'sheet' preconditioner is not part of this library")
        if capacitance:
            raise NotImplementedError("plane-pair capacitance is not part of this library")
        g0 = layers[0].grid
        for l in layers:
            if (l.grid.nx, l.grid.ny, l.grid.d) != (g0.nx, g0.ny, g0.d):
                raise SolverError("all layers must share one grid")

        # layers of different copper thickness: each sheet's strips have their own admittance (gt_schur_factor_nodes)
        self.per_sheet = len({float(l.thickness) for l in layers}) != 1
        self.precond_mode = precond
        self.layers = list(layers)
        self.N = self.n_sheets = len(layers)
        self.nx, self.ny = g0.nx, g0.ny
        self.d = g0.d * 1e-3
        self.zs = np.array([l.z for l in layers], float)
        self.sigma = sigma
        self.sheet_z = self.zs.copy()
        self.sheet_t = np.array([l.thickness for l in layers], float)
        self.device = device
        if setup not in ("device", "host"):
            raise ValueError("setup is 'device' or 'host'")
        self.setup_mode = setup             # where the kernel spectra and the strips' psi are made
        self.device_fronts = int(device_fronts)     # Schur fronts of at least this many rows factorised on the device
        self.schur_leaf = int(schur_leaf)           # nodes of a nested-dissection leaf box at most
        if woodbury not in _WOODBURY_FORMS:
            raise ValueError("woodbury is 'auto', 'stored' or 'implicit'")
        self.woodbury = woodbury        # the branch correction's form (gt_ctx_woodbury); woodbury_form: the one taken
        if schur_ordering not in _ND_ORDERINGS:
            raise ValueError("schur_ordering is 'sheets' (each sheet dissected apart) or 'joint'")
        self.schur_ordering = schur_ordering
        self._topology(list(branches or []))
        if screen:
            self._check_no_split()
        self._kernels()
        self._widths()
        self._handle = None
        self._ops: OrderedDict = OrderedDict()
        self._base: Optional["PlaneStack"] = None
        self._parent: Optional["PlaneStack"] = None
        self.nbr_factored = self.nbr       # leading branches in the factored Laplacian B (all, for a base stack)

    @property
    def handle(self):
        """The resident device stack (gt_stack), made on first use: spectra, layout, incidence, Schur factor."""
        if self._handle is None:
            self._upload()
        return self._handle

    # ---------------------------------------------------------------- topology (peec.py:1596-1656, 1845-1862)
    def _topology(self, branches: list):
        node = np.stack([np.asarray(l.grid.node, bool) for l in self.layers])
        self.xb = np.stack([np.asarray(l.grid.xb, bool) for l in self.layers]) & node[:, :-1, :] & node[:, 1:, :]
        self.yb = np.stack([np.asarray(l.grid.yb, bool) for l in self.layers]) & node[:, :, :-1] & node[:, :, 1:]
        self.node = node
        at = np.nonzero(node)
        self.nid = -np.ones(node.shape, np.int64)
        self.nid[at] = np.arange(len(at[0]))
        self.n_grid = len(at[0])
        # the parts' lids (LidNode) numbered after the grid's nodes, in the order of their numbers (peec.py:1613-1616)
        lids = {}
        for b in branches:
            for e in (b.a, b.b):
                if isinstance(e, LidNode):
                    lids.setdefault(e.lid, e)
        order = sorted(lids)
        self.lid_nid = {k: self.n_grid + n for n, k in enumerate(order)}
        self.n_nodes = self.n_grid + len(order)
        lid_cells = np.array([lids[k][:3] for k in order], np.int64).reshape(-1, 3)
        self.node_layer, self.node_i, self.node_j = (np.concatenate([a, c]).astype(np.int32)
                                                     for a, c in zip(at, lid_cells.T))
        self.node_is_lid = np.concatenate([np.zeros(self.n_grid, bool), np.ones(len(order), bool)])
        self.xa, self.ya = np.argwhere(self.xb), np.argwhere(self.yb)
        self.nxb, self.nyb = len(self.xa), len(self.ya)
        self.ns = self.nxb + self.nyb
        self.branches = [b for b in branches if self.node_id(b.a) >= 0 and self.node_id(b.b) >= 0
                         and self.node_id(b.a) != self.node_id(b.b)]
        self.nbr = len(self.branches)
        self.ncap = 0
        self.nb = self.ns + self.nbr
        xa, ya = self.xa, self.ya
        ba = np.array([self.node_id(b.a) for b in self.branches], np.int64)
        bb = np.array([self.node_id(b.b) for b in self.branches], np.int64)
        self.a_nodes = np.concatenate([self.nid[tuple(xa.T)], self.nid[tuple(ya.T)], ba]).astype(np.int64)
        self.b_nodes = np.concatenate([self.nid[xa[:, 0], xa[:, 1] + 1, xa[:, 2]],
                                       self.nid[ya[:, 0], ya[:, 1], ya[:, 2] + 1], bb]).astype(np.int64)
        n = self.n_nodes
        g = sp.coo_matrix((np.ones(self.nb), (self.a_nodes, self.b_nodes)), shape=(n, n))
        _, self.full_label = connected_components(g, directed=False)
        self.cond_label = self.full_label
        _, first = np.unique(self.full_label, return_index=True)
        self.keep = np.ones(n, bool)
        self.keep[first] = False
        self.n_keep = int(self.keep.sum())
        kidx = np.cumsum(self.keep) - 1
        self.a_keep = np.where(self.keep[self.a_nodes], kidx[self.a_nodes], -1).astype(np.int32)
        self.b_keep = np.where(self.keep[self.b_nodes], kidx[self.b_nodes], -1).astype(np.int32)

    def _check_no_split(self):
        """screen=True splits a layer with copper above and below it at >= 16 of its nodes into two face sheets
        (peec.py:1716-1765); this library has one sheet per layer, so such a stack is refused."""
        if self.N < 3:
            return
        for k in range(self.N):
            above = np.zeros(self.node.shape[1:], bool)
            below = np.zeros(self.node.shape[1:], bool)
            for m in range(self.N):
                if self.zs[m] > self.zs[k]:
                    above |= self.node[m]
                elif self.zs[m] < self.zs[k]:
                    below |= self.node[m]
            if np.count_nonzero(self.node[k] & above & below) >= SCREEN_MIN_NODES:
                raise NotImplementedError("face sheets (screen=True with an inner plane) are not part of this library")

    def node_id(self, node) -> int:
        """Node number of (layer, i, j), -1 off copper; of a LidNode its own number after the grid's, -1 when no
        branch of this stack names it (peec.py:1864-1869)."""
        if isinstance(node, LidNode):
            return self.lid_nid.get(node.lid, -1)
        return int(self.nid[tuple(node)])

    def face_gap(self, s: int, t: int) -> float:
        """Distance between the faces sheets s and t turn to each other, at least 1 um (peec.py:1925-1932)."""
        if s == t:
            return 0.0
        gap = abs(self.sheet_z[s] - self.sheet_z[t]) - 0.5 * (self.sheet_t[s] + self.sheet_t[t])
        return max(gap, 1e-6)

    # ---------------------------------------------------------------- narrow strips (peec.py:1732-1843)
    def _widths(self):
        """The strips' widths from the grids (Grid.wx / wy, rx / ry, ox / oy; full where a grid has none):
        `strip_width` (the inductance's), `r_width` (the resistance's) and, where any strip is narrower than the
        pitch, `dL` (per strip, H) and `M_near` (CSR over the strips, H): what a narrower strip adds to the full
        strip's self term, and the change of the side-by-side neighbours' mutual up to NEAR_AXIAL strips along.
        None of them while every strip is full."""
        ns = self.ns

        def per_strip(attr, default):
            vals = []
            for idx, name in ((self.xa, attr[0]), (self.ya, attr[1])):
                arrs = [getattr(l.grid, name, None) for l in self.layers]
                if all(a is None for a in arrs):
                    vals.append(np.full(len(idx), default))
                    continue
                full = np.stack([np.asarray(a, float) if a is not None else
                                 np.full(self.xb.shape[1:] if name.endswith("x") else self.yb.shape[1:], default)
                                 for a in arrs])
                vals.append(full[tuple(idx.T)] if len(idx) else np.zeros(0))
            return np.concatenate(vals)
        w = per_strip(("wx", "wy"), 1.0)
        ok = w > 0
        self.strip_width = np.where(ok, w, 1.0)
        rw = per_strip(("rx", "ry"), 0.0)
        self.r_width = np.where(ok & (rw > 0), rw, self.strip_width)
        off = np.where(ok, per_strip(("ox", "oy"), 0.0), 0.0)
        self.dL = self.M_near = None
        self.narrow = bool(np.any(self.strip_width != 1.0) or np.any(self.r_width != 1.0))
        narrow = self.strip_width < 1.0
        if not self.narrow or not np.any(narrow):
            return
        d = self.d
        t = self.sheet_t[np.concatenate([self.xa[:, 0], self.ya[:, 0]])]
        keys, inv = np.unique(np.round(np.column_stack([self.strip_width[narrow], t[narrow]]), 12), axis=0,
                              return_inverse=True)
        dL = np.zeros(ns)
        dL[narrow] = G.narrow_self(d, keys[:, 0], keys[:, 1])[np.ravel(inv)]
        self.dL = dL
        # side-by-side pairs: x strips in rows j and j + 1, y strips in columns i and i + 1, |k| <= NEAR_AXIAL along
        a_all, b_all, k_all = [], [], []
        for idx, base, across, shape in ((self.xa, 0, 2, (self.n_sheets, self.nx - 1, self.ny)),
                                         (self.ya, self.nxb, 1, (self.n_sheets, self.nx, self.ny - 1))):
            if len(idx) == 0:
                continue
            at = -np.ones(shape, np.int64)
            at[tuple(idx.T)] = base + np.arange(len(idx))
            along = 1 if across == 2 else 2
            n_ax = shape[along]
            for k in range(-G.NEAR_AXIAL, G.NEAR_AXIAL + 1):
                if abs(k) >= n_ax:
                    continue
                sa = [slice(None)] * 3
                sb = [slice(None)] * 3
                sa[along] = slice(max(-k, 0), n_ax - max(k, 0))
                sb[along] = slice(max(k, 0), n_ax - max(-k, 0))
                sa[across] = slice(0, -1)
                sb[across] = slice(1, None)
                A, B = at[tuple(sa)], at[tuple(sb)]
                both = (A >= 0) & (B >= 0)
                A, B = A[both], B[both]
                m = narrow[A] | narrow[B]
                a_all.append(A[m])
                b_all.append(B[m])
                k_all.append(np.full(int(m.sum()), abs(k)))
        a = np.concatenate(a_all) if a_all else np.zeros(0, np.int64)
        b = np.concatenate(b_all) if b_all else np.zeros(0, np.int64)
        if len(a) == 0:
            return
        kk = np.concatenate(k_all)
        gap = 1.0 + off[b] - off[a]
        keys, inv = np.unique(np.round(np.column_stack([self.strip_width[a], self.strip_width[b], gap, kk]), 9),
                              axis=0, return_inverse=True)
        ones = np.ones(len(keys))
        dM = (G.side_mutual(d, keys[:, 0], keys[:, 1], keys[:, 2], keys[:, 3])
              - G.side_mutual(d, ones, ones, ones, keys[:, 3]))[np.ravel(inv)]
        M = sp.coo_matrix((dM, (a, b)), shape=(ns, ns))
        self.M_near = (M + M.T).tocsr()

    def _set_narrow(self, h):
        """The narrow strips' terms on the device stack (gt_stack_set_narrow)."""
        if not self.narrow:
            return
        rf = np.ascontiguousarray(1.0 / self.r_width)
        dl = np.ascontiguousarray(self.dL if self.dL is not None else np.zeros(self.ns))
        if self.M_near is not None:
            M = self.M_near.tocsr()
            M.sort_indices()
            ptr = np.ascontiguousarray(M.indptr, np.int32)
            col = np.ascontiguousarray(M.indices, np.int32)
            val = np.ascontiguousarray(M.data, np.float64)
        else:
            ptr = np.zeros(self.ns + 1, np.int32)
            col, val = np.zeros(1, np.int32), np.zeros(1)
        _lib.check(_lib.lib().gt_stack_set_narrow(h, _lib.ptr(rf), _lib.ptr(dl), int(ptr[-1]), _lib.ptr(ptr, _lib.i32p),
                                                 _lib.ptr(col, _lib.i32p), _lib.ptr(val)), "gt_stack_set_narrow")

    # ---------------------------------------------------------------- kernels (peec.py:1941-1988)
    def _kernels(self):
        nx, ny, S = self.nx, self.ny, self.n_sheets
        self.Px = (G.fft_size(nx - 1), G.fft_size(ny))
        self.Py = (G.fft_size(nx), G.fft_size(ny - 1))
        if self.setup_mode == "device":
            # the spectra are made on the device at upload; the host keeps the near field (the reference's graded
            # quadrature, 5 x 5 offsets per sheet pair) and the self term
            d = G.key(self.d)
            self.L_self = _near_table(d, 0.0)[0, 0]
            self._dz = np.array([[G.key(self.face_gap(s, t)) for t in range(S)] for s in range(S)])
            self._near = np.ascontiguousarray(np.stack([np.stack([_near_table(d, self._dz[s, t]) for t in range(S)])
                                                        for s in range(S)]))
            return
        hx = (self.Px[1] // 2 + 1, self.Px[0] // 2 + 1)
        hy = (self.Py[1] // 2 + 1, self.Py[0] // 2 + 1)
        Kx = np.empty((*hx, S, S))
        Ky = np.empty((*hy, S, S))
        self.L_self = _kernel_pair(nx, ny, self.d, 0.0, self.Px, self.Py)[2]
        for s in range(S):
            for t in range(s, S):
                qx, qy, _ = _kernel_pair(nx, ny, self.d, self.face_gap(s, t), self.Px, self.Py)
                Kx[:, :, s, t] = Kx[:, :, t, s] = qx.T / (self.Px[0] * self.Px[1])
                Ky[:, :, s, t] = Ky[:, :, t, s] = qy.T / (self.Py[0] * self.Py[1])
        self._kx, self._ky = np.ascontiguousarray(Kx), np.ascontiguousarray(Ky)

    # ---------------------------------------------------------------- what-if stacks (peec.py:1871-1923)
    def with_branches(self, extra: list) -> "PlaneStack":
        """This stack plus a few branches (candidate vias): shares the grids, kernels, numbering and the device's
        spectra and Schur factor; the added branches enter the preconditioner as Woodbury columns of the base's
        factor (peec.py:1871-1909). A branch that joins two islands changes the reference nodes: then a stack of its
        own."""
        # (a branch to a lid this stack does not have joins nothing here: new lids need a stack of their own)
        extra = [b for b in extra if self.node_id(b.a) >= 0 and self.node_id(b.b) >= 0
                 and self.node_id(b.a) != self.node_id(b.b)]
        if not extra:
            return self
        a = np.array([self.node_id(b.a) for b in extra], np.int64)
        b = np.array([self.node_id(b.b) for b in extra], np.int64)
        if np.any(self.full_label[a] != self.full_label[b]) or self.per_sheet:
            return PlaneStack(self.layers, self.branches + extra, sigma=self.sigma, precond=self.precond_mode,
                              device=self.device, setup=self.setup_mode, device_fronts=self.device_fronts,
                              schur_leaf=self.schur_leaf, woodbury=self.woodbury,
                              schur_ordering=self.schur_ordering)
        base = self if self._base is None else self._base
        new = copy.copy(self)
        new.branches = self.branches + extra
        new.nbr = self.nbr + len(extra)
        new.nb = self.nb + len(extra)
        new.a_nodes = np.concatenate([self.a_nodes, a])
        new.b_nodes = np.concatenate([self.b_nodes, b])
        kidx = np.cumsum(self.keep) - 1
        new.a_keep = np.concatenate([self.a_keep, np.where(self.keep[a], kidx[a], -1)]).astype(np.int32)
        new.b_keep = np.concatenate([self.b_keep, np.where(self.keep[b], kidx[b], -1)]).astype(np.int32)
        new._base = base                    # the stack whose factor serves (the first of a chain of what-ifs)
        new._parent = self                  # the stack map_state maps from
        new.nbr_factored = base.nbr_factored
        new._handle = None
        new._ops = OrderedDict()
        return new

    def with_layers(self, layers: list, branches: Optional[list] = None) -> "PlaneStack":
        """A stack of another copper topology on this stack's grid (a moved slot, a changed plane): the same sheets,
        heights and thicknesses with new masks, and this stack's branches (or `branches`) where both ends still have
        copper. Its numbering and factor are its own (the reference builds a new model for such a what-if); the
        inductance kernels, which depend only on the grid and the sheet gaps (peec.py:3288-3290's cache key), are
        this stack's resident spectra, shared on the device."""
        if len(layers) != self.N or any((l.grid.nx, l.grid.ny, l.grid.d) != (self.layers[0].grid.nx,
                                                                         self.layers[0].grid.ny,
                                                                         self.layers[0].grid.d) for l in layers):
            raise ValueError("with_layers keeps the stack's grid and sheet count")
        if any(a.z != b.z or a.thickness != b.thickness for a, b in zip(layers, self.layers)):
            raise ValueError("with_layers keeps each layer's height and thickness (the kernels' gaps)")
        new = PlaneStack(layers, self.branches if branches is None else branches, sigma=self.sigma,
                         precond=self.precond_mode, device=self.device, setup=self.setup_mode,
                         device_fronts=self.device_fronts, schur_leaf=self.schur_leaf, woodbury=self.woodbury,
                         schur_ordering=self.schur_ordering)
        if self.setup_mode == "device":
            donor = self if self._base is None else self._base
            new._kernel_donor = donor.__dict__.get("_kernel_donor") or donor
        return new

    def plan_bytes(self) -> tuple:
        """(bytes held, bytes at the peak of its set-up) this stack's device copy will take, before it is made: a
        what-if derived from a resident base, exactly (gt_stack_derive_bytes); any other stack, its kernel donor's
        or its own resident bytes if it has one, else an estimate from its unknowns (0 when already resident)."""
        if self._handle is not None:
            return 0, 0
        if self._base is not None and self._base._handle is not None:
            nb0 = self._base.nb
            ea = np.ascontiguousarray(self.a_keep[nb0:], np.int32)
            eb = np.ascontiguousarray(self.b_keep[nb0:], np.int32)
            n = len(ea)
            held, peak = C.c_int64(), C.c_int64()
            _lib.check(_lib.lib().gt_stack_derive_bytes(
                self._base.handle, n, _lib.ptr(ea if n else np.zeros(1, np.int32), _lib.i32p),
                _lib.ptr(eb if n else np.zeros(1, np.int32), _lib.i32p), C.byref(held), C.byref(peak)),
                "gt_stack_derive_bytes")
            return int(held.value), int(peak.value)
        donor = self.__dict__.get("_kernel_donor")
        if donor is not None and donor._handle is not None:
            b = donor.device_bytes()            # the same grid: a like topology's layout, factor and scratch
            return b, b
        b = int(2000 * (self.nb + self.n_keep))
        return b, b

    def map_state(self, x: Optional[np.ndarray], stack: "PlaneStack") -> Optional[np.ndarray]:
        """A solution vector of `stack` in this stack's unknowns (a warm start): the same stack, or this stack's
        base, whose added branches start at zero current (peec.py:1911-1923)."""
        if x is None:
            return None
        if stack is self:
            return x
        if self._parent is stack and len(x) == stack.nb + stack.n_keep:
            return np.concatenate([x[:stack.nb], np.zeros(self.nb - stack.nb, complex), x[stack.nb:]])
        return None

    def device_bytes_estimate(self) -> int:
        """About what this stack and its solves put on the device (bytes), before any of it is made: the kernel
        quadrants (FP64 and FP32, one per distinct face gap at most S of them for equally spaced sheets, S^2 at
        worst), the strip layout and incidence, the L_p intermediates (2 complex per padded row point), the Schur
        factor's panels (nested dissection of a grid: about n (3.8 log2(n / S) - 12.5) entries, fitted to the
        factors of 2 x 256^2, 4 x 512^2 and 6 x 2048^2 boards), the Schur solve's vectors, and the GMRES basis at its
        longest with a few more vectors (the reference's count, gpu.py:593-605)."""
        S = self.n_sheets
        n = self.nb + self.n_keep
        G_k = len({float(v) for v in np.ravel(getattr(self, "_dz", np.zeros((S, S))))}) or S * S
        kern = sum((P[0] // 2 + 1) * (P[1] // 2 + 1) for P, m in ((self.Px, self.nxb), (self.Py, self.nyb)) if m)
        kernels = kern * G_k * 12
        rows = sum(S * P[1] * (nrow) for P, nrow, m in ((self.Px, self.nx - 1, self.nxb), (self.Py, self.nx, self.nyb))
                   if m)
        lp = 2 * rows * 16
        layout = 4 * (self.ns + 2 * S * self.nx + 6 * self.nb + self.n_keep)
        nk = max(self.n_keep, 2)
        nnz = nk * max(8.0, 3.8 * math.log2(max(nk / S, 2.0)) - 12.5)
        factor = int(nnz * 8 + 3 * nk * 16 * BATCH)
        restart = min(200, max(100, int(4e8 / (16 * n))))
        basis = (restart + 1) * n * 16
        return int(kernels + lp + layout + factor + basis + 20 * n * 16)

    def _check_fits(self):
        """DeviceTooSmall when the stack would take more than DEVICE_MEMORY_SHARE of the device's free memory
        (gpu.py:483-488), before anything is uploaded."""
        import torch
        free = int(torch.cuda.mem_get_info(self.device)[0])
        need = self.device_bytes_estimate()
        if need > DEVICE_MEMORY_SHARE * free:
            # memory this process keeps for reuse counts as free: torch's cached blocks and the context's GMRES
            # workspace (made again by the next solve)
            torch.cuda.empty_cache()
            _lib.check(_lib.lib().gt_ctx_trim(_lib.Context.get(self.device).handle), "gt_ctx_trim")
            free = int(torch.cuda.mem_get_info(self.device)[0])
        if need > DEVICE_MEMORY_SHARE * free:
            raise DeviceTooSmall(f"the model needs ~{need / 1e9:.3g} GB on device {self.device}, {free / 1e9:.3g} GB "
                                 "free")

    # ---------------------------------------------------------------- the device copy
    def _upload(self):
        if self._base is not None:
            self._derive()
            return
        self._check_fits()
        h = self._create()
        self._set_narrow(h)
        self._tune(h)
        self._factor(h)

    def _create(self):
        """The device stack without its Schur factor: layout, incidence, kernel spectra (gt_stack_create[_gen])."""
        L = _lib.lib()
        self.ctx = _lib.Context.get(self.device)
        xs = np.ascontiguousarray(self.xa, np.int32)
        ys = np.ascontiguousarray(self.ya, np.int32)
        keep_idx = np.nonzero(self.keep)[0]
        node_sheet = np.ascontiguousarray(self.node_layer[keep_idx], np.int32)
        d = _lib.StackDesc()
        d.n_sheets, d.nx, d.ny = self.n_sheets, self.nx, self.ny
        d.nxb, d.nyb, d.nbr, d.n_keep = self.nxb, self.nyb, self.nbr, self.n_keep
        d.x_sij, d.y_sij = _lib.ptr(xs.reshape(-1) if xs.size else np.zeros(3, np.int32), _lib.i32p), \
            _lib.ptr(ys.reshape(-1) if ys.size else np.zeros(3, np.int32), _lib.i32p)
        a_keep = self.a_keep if len(self.a_keep) else np.zeros(1, np.int32)
        b_keep = self.b_keep if len(self.b_keep) else np.zeros(1, np.int32)
        d.a_keep, d.b_keep = _lib.ptr(a_keep, _lib.i32p), _lib.ptr(b_keep, _lib.i32p)
        d.node_sheet = _lib.ptr(node_sheet if len(node_sheet) else np.zeros(1, np.int32), _lib.i32p)
        d.px[0], d.px[1] = self.Px
        d.py[0], d.py[1] = self.Py
        h = C.c_void_p()
        donor = self.__dict__.get("_kernel_donor")
        if donor is not None and np.array_equal(donor._dz, self._dz) and donor.d == self.d:
            _lib.check(L.gt_stack_create_like(self.ctx.handle, C.byref(d), donor.handle, C.byref(h)),
                       "gt_stack_create_like")
            del self._kernel_donor              # the spectra are shared on the device now
        elif self.setup_mode == "device":
            kd = _lib.KernelDesc(G.key(self.d), G.MU0 / (4 * math.pi), _lib.ptr(self._dz), _lib.ptr(self._near))
            _lib.check(L.gt_stack_create_gen(self.ctx.handle, C.byref(d), C.byref(kd), C.byref(h)),
                       "gt_stack_create_gen")
        else:
            d.kx, d.ky = _lib.ptr(self._kx), _lib.ptr(self._ky)
            _lib.check(L.gt_stack_create(self.ctx.handle, C.byref(d), C.byref(h)), "gt_stack_create")
        return h

    def _tune(self, h):
        """The L_p product's launch shapes for this board's shape: tuned on the first stack of the shape in this
        process (gt_stack_lp_autotune), the same plan applied to every later one."""
        L = _lib.lib()
        key = (self.n_sheets, self.nx, self.ny, tuple(self.Px), tuple(self.Py))
        plan = _LP_PLANS.get(key)
        if plan is None:
            rw, cw = np.zeros(3, np.int32), np.zeros(3, np.int32)
            _lib.check(L.gt_stack_lp_autotune(h, 5, _lib.ptr(rw, _lib.i32p), _lib.ptr(cw, _lib.i32p)),
                       "gt_stack_lp_autotune")
            plan = _LP_PLANS[key] = (rw, cw)
        else:
            _lib.check(L.gt_stack_set_lp_plan(h, _lib.ptr(plan[0], _lib.i32p), _lib.ptr(plan[1], _lib.i32p)),
                       "gt_stack_set_lp_plan")
        self.lp_plan = {"rows_per_workgroup": plan[0].tolist(), "waves_per_workgroup": plan[1].tolist()}

    def _factor(self, h):
        """The Schur factor of the stack `h` (gt_schur_factor), then the stack is this PlaneStack's."""
        L = _lib.lib()
        _lib.check(L.gt_ctx_schur_device_fronts(self.ctx.handle, min(self.device_fronts, 2 ** 31 - 1)),
                   "gt_ctx_schur_device_fronts")
        _lib.check(L.gt_ctx_schur_leaf(self.ctx.handle, self.schur_leaf), "gt_ctx_schur_leaf")
        _lib.check(L.gt_ctx_woodbury(self.ctx.handle, _WOODBURY_FORMS[self.woodbury], 16 << 30), "gt_ctx_woodbury")
        _lib.check(L.gt_ctx_schur_ordering(self.ctx.handle, _ND_ORDERINGS[self.schur_ordering]),
                   "gt_ctx_schur_ordering")
        keep_idx = np.nonzero(self.keep)[0]
        # the nested dissection's sheet of a node: its layer; a lid's a sheet of its own (its bonds then join sheets,
        # so it and the nodes under its pads sit in the top separator with the vias' ends)
        sheet = np.where(self.node_is_lid, self.n_sheets, self.node_layer)
        xyz = np.ascontiguousarray(np.stack([sheet[keep_idx], self.node_i[keep_idx],
                                             self.node_j[keep_idx]], axis=1).astype(np.int32))
        xyz_p = _lib.ptr(xyz.reshape(-1) if xyz.size else np.zeros(3, np.int32), _lib.i32p)
        if self.per_sheet:
            kmax = max(2 * self.nbr, 1)
            Gm = np.zeros((kmax, kmax))
            cols = np.zeros(kmax, np.int32)
            nc = C.c_int32()
            code = L.gt_schur_factor_nodes(h, xyz_p, C.byref(nc), _lib.ptr(cols, _lib.i32p), _lib.ptr(Gm))
            if code != _lib.GT_OK:
                L.gt_stack_destroy(h)
                _lib.check(code, "gt_schur_factor_nodes")
            k = nc.value
            self.wb_cols = cols[:k].astype(np.int64)
            self.G = Gm.reshape(-1)[: k * k].reshape(k, k).copy()
        else:
            Gm = np.zeros((max(self.nbr, 1), max(self.nbr, 1)))
            code = L.gt_schur_factor(h, xyz_p, _lib.ptr(Gm))
            if code != _lib.GT_OK:
                L.gt_stack_destroy(h)
                _lib.check(code, "gt_schur_factor")
            self.G = Gm[: self.nbr, : self.nbr].copy()
        self._handle = h

    def _derive(self):
        """The device stack of a what-if stack: the base's (made first) plus the added branches' incidence and
        Woodbury columns (gt_stack_derive)."""
        base = self._base
        base_h = base.handle
        self.ctx = base.ctx
        nx = self.nbr - base.nbr
        ea = np.ascontiguousarray(self.a_keep[base.nb:], np.int32)
        eb = np.ascontiguousarray(self.b_keep[base.nb:], np.int32)
        Gm = np.zeros((max(self.nbr, 1), max(self.nbr, 1)))
        h = C.c_void_p()
        _lib.check(_lib.lib().gt_stack_derive(base_h, nx, _lib.ptr(ea if nx else np.zeros(1, np.int32), _lib.i32p),
                                              _lib.ptr(eb if nx else np.zeros(1, np.int32), _lib.i32p),
                                              _lib.ptr(Gm), C.byref(h)), "gt_stack_derive")
        self.G = Gm[: self.nbr, : self.nbr].copy()
        self._handle = h

    def release(self):
        """Give the device copy back (its operators first), keeping the host topology: a finished what-if's
        solutions still unpack. The next device use makes it again."""
        if self._handle is None:
            return
        L = _lib.lib()
        for op in self._ops.values():
            if op.handle:
                L.gt_op_destroy(op.handle)
                op.handle = None
        self._ops.clear()
        L.gt_stack_destroy(self._handle)
        self._handle = None

    def __del__(self):
        try:
            if getattr(self, "_handle", None) and _lib._LIB is not None:
                self._ops.clear()
                _lib._LIB.gt_stack_destroy(self._handle)
        except Exception:       # noqa: BLE001 - interpreter shutdown
            pass

    def device_bytes(self) -> int:
        """Device memory this stack holds (its resident spectra, layout, factor and scratch)."""
        return int(_lib.lib().gt_stack_device_bytes(self.handle))

    @property
    def woodbury_form(self) -> str:
        """The branch correction's form this stack took: 'stored' (W = B^-1 U resident), 'implicit' (only
        U^T B^-1 U kept; each preconditioner product makes a second Schur solve) or 'root' (every branch end in the
        factor's last supernode: the correction runs on that supernode's rows between the solve's passes)."""
        form, k = C.c_int32(), C.c_int32()
        _lib.check(_lib.lib().gt_stack_woodbury(self.handle, C.byref(form), C.byref(k)), "gt_stack_woodbury")
        return {1: "stored", 2: "implicit", 3: "root"}[form.value]

    def schur_stats(self) -> dict:
        nnz, nsup, lev, mc = C.c_int64(), C.c_int32(), C.c_int32(), C.c_int32()
        _lib.check(_lib.lib().gt_schur_stats(self.handle, C.byref(nnz), C.byref(nsup), C.byref(lev),
                                             C.byref(mc)), "gt_schur_stats")
        T, cols, ent = C.c_int32(), C.c_int32(), C.c_int64()
        _lib.check(_lib.lib().gt_schur_top_stats(self.handle, C.byref(T), C.byref(cols), C.byref(ent)),
                   "gt_schur_top_stats")
        return {"nnz_L": nnz.value, "supernodes": nsup.value, "levels": lev.value, "max_cols": mc.value,
                "dense_top": {"levels": T.value, "cols": cols.value, "entries": ent.value}}

    # ---------------------------------------------------------------- operators (peec.py:2227-2345)
    @staticmethod
    def freq_key(f: float) -> float:
        return G.key(f)

    def operator(self, f: float) -> Operator:
        """The device operator at f's 12-digit key (kept for later solves at the same key)."""
        k = G.key(f)
        op = self._ops.get(k)
        if op is None:
            op = self._ops[k] = Operator(self, k)
            while len(self._ops) > OPERATORS_KEPT:
                self._ops.popitem(last=False)
        else:
            self._ops.move_to_end(k)
        return op

    def solve(self, f: float, rhs_top: np.ndarray, J: np.ndarray, tol: float = 1e-5, max_cycles: int = 8,
              guess: Optional[list] = None, restart: Optional[int] = None):
        """Solve K [I; V] = [rhs_top; J] (peec.py:2513-2607). Returns (Ib, V, iterations, x). `restart` sets the
        GMRES restart length (None: the reference's min(200, max(100, 4e8 / 16n)))."""
        op = self.operator(f)
        nb = self.nb
        rhs = np.concatenate([np.asarray(rhs_top, complex), np.asarray(J, float)[self.keep].astype(complex)])
        bnorm = float(np.linalg.norm(rhs))
        Gs = [g for g in (guess or []) if g is not None and len(g) == len(rhs)]
        if Gs:
            # the warm start and the solve on device vectors
            import torch
            dev = f"cuda:{self.device}"
            rhs_d = torch.from_numpy(rhs).to(dev)
            G_d = [torch.from_numpy(np.ascontiguousarray(g, np.complex128)).to(dev) for g in Gs]
            torch.cuda.synchronize(self.device)
            x0, r0 = self.warm_start_dev(op, rhs_d, G_d)
            xd = torch.empty_like(rhs_d)
            its, ok, _ = self.gmres_dev(op, r0, xd, tol, bnorm=bnorm, max_cycles=max_cycles, restart=restart)
            with torch.cuda.stream(self.torch_stream()):
                x = (xd + x0 if x0 is not None else xd).cpu().numpy()
        else:
            x, its, ok, _ = self.gmres(op, rhs, tol, bnorm=bnorm, max_cycles=max_cycles, restart=restart)
        if not ok:
            raise SolverError(f"GMRES did not converge at {f:.4g} Hz ({its} iterations)")
        V = np.zeros(self.n_nodes, complex)
        V[self.keep] = x[nb:]
        return x[:nb], V, its, x

    def node_field_dev(self, x_dev, d_mm: float):
        """(Kx, Ky) at every grid node of every layer, torch complex128 tensors (N, nx, ny) on the device, from a solution
        vector on the device (gt_node_field): Solution.node_field (copper=False) without copying x to the host."""
        import torch
        self.handle
        shape = (self.n_sheets, self.nx, self.ny)
        kx = torch.empty(shape, dtype=torch.complex128, device=x_dev.device)
        ky = torch.empty(shape, dtype=torch.complex128, device=x_dev.device)
        torch.cuda.synchronize(self.device)
        _lib.check(_lib.lib().gt_node_field(self.handle, C.c_void_p(x_dev.data_ptr()), float(d_mm) * 1e-3,
                                            C.c_void_p(kx.data_ptr()), C.c_void_p(ky.data_ptr())), "gt_node_field")
        return kx, ky

    def torch_stream(self):
        """The library's stream as a torch stream: torch work in its context is ordered with the library's."""
        import torch
        self.handle
        ts = getattr(self.ctx, "torch_stream", None)
        if ts is None:
            ts = self.ctx.torch_stream = torch.cuda.ExternalStream(_lib.lib().gt_ctx_stream(self.ctx.handle),
                                                                   device=f"cuda:{self.device}")
        return ts

    def warm_start_dev(self, op: Operator, rhs, guess: list, host: Optional["PlaneStack"] = None):
        """(x0, r0) on device vectors (torch complex128, their work finished; the results ordered on the library's
        stream): x0 = G c with c minimising |K G c - rhs| over the guesses, r0 = rhs - K x0 the residual left to
        GMRES, taken when it removes at least a tenth of |rhs|; else (None, rhs) (peec.py:2568-2576). As the
        reference, c is the least-squares solution of K G c = rhs itself (a QR factorisation of the n x m KG on the
        device, not its normal equations), and the acceptance is on the norm of the residual formed on the device.
        host: the stack whose scratch runs the products K g (this stack's base: then a what-if stack needs none of
        its own)."""
        import torch
        n = self.nb + self.n_keep
        Gs = [g for g in guess if g is not None and g.numel() == n]
        if not Gs:
            return None, rhs
        L = _lib.lib()
        with torch.cuda.stream(self.torch_stream()):
            keep = [g for g in Gs if bool(torch.any(g != 0))]     # all-zero guesses take no part
            if not keep:
                return None, rhs
            KG = torch.empty((n, len(keep)), dtype=torch.complex128, device=rhs.device)
            col = torch.empty(n, dtype=torch.complex128, device=rhs.device)
            for i, g in enumerate(keep):
                if host is not None:
                    _lib.check(L.gt_matvec_on(host.handle, op.handle, C.c_void_p(g.data_ptr()),
                                              C.c_void_p(col.data_ptr())), "gt_matvec_on")
                else:
                    _lib.check(L.gt_matvec(op.handle, C.c_void_p(g.data_ptr()), C.c_void_p(col.data_ptr())),
                               "gt_matvec")
                KG[:, i] = col
            Q, R = torch.linalg.qr(KG)
            c = torch.linalg.solve_triangular(R, (Q.conj().T @ rhs).unsqueeze(1), upper=True).squeeze(1)
            r = rhs - KG @ c
            ok = bool(torch.linalg.vector_norm(r) < 0.9 * torch.linalg.vector_norm(rhs))
            if not ok:
                return None, rhs
            x0 = torch.zeros_like(rhs)
            for i, g in enumerate(keep):
                x0.add_(g, alpha=complex(c[i].item()))
        return x0, r

    def map_state_dev(self, x, stack: "PlaneStack"):
        """map_state on a device vector (torch)."""
        import torch
        if x is None:
            return None
        if stack is self:
            return x
        if self._parent is stack and x.numel() == stack.nb + stack.n_keep:
            z = torch.zeros(self.nb - stack.nb, dtype=x.dtype, device=x.device)
            return torch.cat([x[:stack.nb], z, x[stack.nb:]])
        return None

    def gmres(self, op: Operator, b: np.ndarray, tol: float, bnorm: Optional[float] = None, max_cycles: int = 8,
              restart: Optional[int] = None, extra_cycles: int = GMRES_EXTRA_CYCLES, max_its: Optional[int] = None):
        """x = M^-1 y with (K M^-1) y = b by the device GMRES (peec.py:3093-3224). (x, its, converged, result)."""
        n = self.nb + self.n_keep
        bb = _lib.cplx(b)
        assert bb.size == 2 * n
        out = np.zeros(2 * n)
        p = _lib.GmresParams(float(tol), float(bnorm or 0.0), int(restart or 0), int(max_cycles),
                             int(extra_cycles), int(max_its or 0), GMRES_CHECK_MISSES)
        res = _lib.GmresResult()
        _lib.check(_lib.lib().gt_gmres(op.handle, _lib.ptr(bb), _lib.ptr(out), C.byref(p), C.byref(res)), "gt_gmres")
        return out.view(np.complex128).copy(), int(res.iterations), bool(res.converged), res

    def gmres_batch_dev(self, ops: list, b_devs: list, x_devs: list, tol: float, max_cycles: int = 8,
                        restart: Optional[int] = None, extra_cycles: int = GMRES_EXTRA_CYCLES,
                        bnorm: Optional[float] = None, precision: str = "fp64",
                        inner_tol: float = MIXED_INNER_TOL, max_refine: int = MIXED_MAX_REFINE) -> list:
        """Batched device GMRES: x_i = M_i^-1 y_i with (K_i M_i^-1) y_i = b_i for every operator i of this stack, on
        device-resident vectors (torch complex128 tensors, their work finished), at most 16 per launch group. Each
        follows the reference's GMRES; the batch shares every launch. precision "mixed": FP64 iterative refinement
        to tol around FP32 inner solves to inner_tol (gt_gmres_batch_mixed). Returns one GmresResult per solve."""
        if precision not in ("fp64", "mixed", "fp32-basis"):
            raise ValueError("precision is 'fp64', 'mixed' or 'fp32-basis'")
        out = []
        for g0 in range(0, len(ops), BATCH):
            grp = slice(g0, g0 + BATCH)
            o, b, x = ops[grp], b_devs[grp], x_devs[grp]
            nr = len(o)
            H = (C.c_void_p * nr)(*[op.handle for op in o])
            B = (C.c_void_p * nr)(*[t.data_ptr() for t in b])
            X = (C.c_void_p * nr)(*[t.data_ptr() for t in x])
            p = _lib.GmresParams(float(tol), float(bnorm or 0.0), int(restart or 0), int(max_cycles),
                                 int(extra_cycles), 0, GMRES_CHECK_MISSES)
            res = (_lib.GmresResult * nr)()
            if precision == "fp64":
                _lib.check(_lib.lib().gt_gmres_batch(self.handle, nr, H, B, X, C.byref(p), res), "gt_gmres_batch")
            elif precision == "fp32-basis":
                _lib.check(_lib.lib().gt_gmres_batch_fp32_basis(self.handle, nr, H, B, X, C.byref(p), res),
                           "gt_gmres_batch_fp32_basis")
            else:
                _lib.check(_lib.lib().gt_gmres_batch_mixed(self.handle, nr, H, B, X, C.byref(p), float(inner_tol),
                                                           int(max_refine), res), "gt_gmres_batch_mixed")
            out += list(res)
        return out

    def solve_many(self, fs: list, rhs_tops: list, Js: list, tol: float = 1e-5, max_cycles: int = 8,
                   restart: Optional[int] = None, precision: str = "fp64") -> list:
        """PlaneStack.solve for many (f, rhs_top, J) on this stack at once (batched on the device): the list of
        (Ib, V, iterations, x), each the same as a solve of its own. SolverError if any does not converge."""
        import torch
        dev = f"cuda:{self.device}"
        ops = [self.operator(f) for f in fs]
        rhs = [np.concatenate([np.asarray(r, complex), np.asarray(J, float)[self.keep].astype(complex)])
               for r, J in zip(rhs_tops, Js)]
        b = [torch.from_numpy(r).to(dev) for r in rhs]
        x = [torch.empty_like(t) for t in b]
        torch.cuda.synchronize(self.device)
        res = self.gmres_batch_dev(ops, b, x, tol, max_cycles=max_cycles, restart=restart, precision=precision)
        self.ctx.sync()
        out = []
        for f, r, xd in zip(fs, res, x):
            if not r.converged:
                raise SolverError(f"GMRES did not converge at {f:.4g} Hz ({r.iterations} iterations)")
            xv = xd.cpu().numpy()
            V = np.zeros(self.n_nodes, complex)
            V[self.keep] = xv[self.nb:]
            out.append((xv[:self.nb], V, int(r.iterations), xv))
        return out

    def gmres_dev(self, op: Operator, b_dev, x_dev, tol: float, bnorm: Optional[float] = None, max_cycles: int = 8,
                  restart: Optional[int] = None, extra_cycles: int = GMRES_EXTRA_CYCLES):
        """The device GMRES on device-resident vectors (torch complex128 tensors of nb + n_keep values, their work
        finished): x_dev = M^-1 y. Returns (iterations, converged, result)."""
        p = _lib.GmresParams(float(tol), float(bnorm or 0.0), int(restart or 0), int(max_cycles),
                             int(extra_cycles), 0, GMRES_CHECK_MISSES)
        res = _lib.GmresResult()
        _lib.check(_lib.lib().gt_gmres_dev(op.handle, C.c_void_p(b_dev.data_ptr()), C.c_void_p(x_dev.data_ptr()),
                                           C.byref(p), C.byref(res)), "gt_gmres_dev")
        return int(res.iterations), bool(res.converged), res

    # ---------------------------------------------------------------- device products on host vectors
    def _dev(self, v: np.ndarray):
        import torch
        self.handle
        return torch.from_numpy(np.ascontiguousarray(v, np.complex128)).to(f"cuda:{self.device}")

    def _run(self, fn, x: np.ndarray, n_out: int) -> np.ndarray:
        import torch
        xd = self._dev(x)
        yd = torch.zeros(n_out, dtype=torch.complex128, device=f"cuda:{self.device}")
        torch.cuda.synchronize(self.device)
        _lib.check(fn(C.c_void_p(xd.data_ptr()), C.c_void_p(yd.data_ptr())))
        self.ctx.sync()
        return yd.cpu().numpy()

    def Lp_matvec(self, I_s: np.ndarray) -> np.ndarray:
        """L_p I on the device (peec.py:2115-2135)."""
        return self._run(lambda a, b: _lib.lib().gt_lp_matvec(self.handle, a, b), I_s, self.ns)

    def matvec(self, op: Operator, x: np.ndarray) -> np.ndarray:
        """K x on the device (peec.py:3043-3062)."""
        return self._run(lambda a, b: _lib.lib().gt_matvec(op.handle, a, b), x, self.nb + self.n_keep)

    def precond(self, op: Operator, r: np.ndarray) -> np.ndarray:
        """M^-1 r on the device (peec.py:3064-3083)."""
        return self._run(lambda a, b: _lib.lib().gt_precond(op.handle, a, b), r, self.nb + self.n_keep)

    def schur_solve(self, b: np.ndarray) -> np.ndarray:
        """B^-1 b on the device (B = Ar^T Ar, the factorised Schur base)."""
        return self._run(lambda a, c: _lib.lib().gt_schur_solve(self.handle, a, c), b, self.n_keep)

    def schur_solve_host(self, b: np.ndarray) -> np.ndarray:
        """B^-1 b with the same factor on the host (a device-free check of the factorisation)."""
        bb = _lib.cplx(b)
        out = np.zeros_like(bb)
        _lib.check(_lib.lib().gt_schur_solve_host(self.handle, _lib.ptr(bb), _lib.ptr(out)), "gt_schur_solve_host")
        return out.view(np.complex128).copy()

    # ---------------------------------------------------------------- coupling to the signal (peec.py:2609-2700)
    def pos(self, node) -> np.ndarray:
        k, i, j = node
        return np.array([i * self.d, j * self.d, self.zs[k]])

    def strip_coupling(self, trace: list) -> np.ndarray:
        """psi = int A.dl along every strip per amp of signal (peec.py:2609-2656): the segment's vector potential
        at the strip's centre on its face toward the segment, or its mean over 3 x 3 Gauss points of the strip's
        square within 3 pitches (+ 2 widths) of the segment. On the device (gt_strip_coupling) unless the stack's
        set-up is the host's."""
        if self.setup_mode == "host" or self.ns == 0:
            return self.strip_coupling_host(trace)
        return self.strip_coupling_dev(trace).cpu().numpy()

    def strip_coupling_dev(self, trace: list):
        """strip_coupling as a device tensor (torch float64, ns values; its work finished)."""
        import torch
        segs = _segments(trace)
        psi = torch.zeros(self.ns, dtype=torch.float64, device=f"cuda:{self.device}")
        torch.cuda.synchronize(self.device)
        z = np.ascontiguousarray(self.sheet_z, float)
        t = np.ascontiguousarray(self.sheet_t, float)
        _lib.check(_lib.lib().gt_strip_coupling(self.handle, float(self.d), G.MU0 / (4 * math.pi), len(trace), segs,
                                                _lib.ptr(z), _lib.ptr(t), C.c_void_p(psi.data_ptr())),
                   "gt_strip_coupling")
        self.ctx.sync()
        return psi

    def strip_coupling_host(self, trace: list) -> np.ndarray:
        """strip_coupling in numpy."""
        d = self.d
        x, w = np.polynomial.legendre.leggauss(3)
        q = 0.5 * x
        W = (np.outer(w, w) / 4.0).ravel()
        qa, ql = (g.ravel() for g in np.meshgrid(q, q, indexing="ij"))
        psi = np.zeros(self.ns)
        for axis, idx, off in ((0, self.xa, 0), (1, self.ya, self.nxb)):
            if len(idx) == 0:
                continue
            k, i, j = idx[:, 0], idx[:, 1], idx[:, 2]
            cx = i * d + (d / 2 if axis == 0 else 0.0)
            cy = j * d + (d / 2 if axis == 1 else 0.0)
            half = 0.5 * self.sheet_t[k]
            acc = np.zeros(len(idx))
            for s in trace:
                ab = s.b - s.a
                if np.linalg.norm(ab[:2]) == 0:
                    continue
                L = float(np.linalg.norm(ab))
                if abs(ab[axis]) < 1e-15 * L:
                    continue
                zseg = 0.5 * (s.a[2] + s.b[2])
                cz = self.sheet_z[k] + np.where(zseg > self.sheet_z[k], half, -half)
                centres = np.stack([cx, cy, cz], axis=1)
                rho_min = max(0.5 * s.width, 1e-7)
                val = G.segment_potential(s.a, s.b, centres, rho_min)
                near = _dist_to_segment(centres, s.a, s.b) < 3 * d + 2 * s.width
                if near.any():
                    c = centres[near]
                    da, dl = (qa, ql) if axis == 0 else (ql, qa)
                    pts = np.stack([c[:, 0, None] + da * d, c[:, 1, None] + dl * d,
                                    np.broadcast_to(c[:, 2, None], (len(c), len(W)))], axis=2)
                    vq = G.segment_potential(s.a, s.b, pts.reshape(-1, 3), rho_min).reshape(len(c), len(W))
                    val[near] = (vq * W).sum(1)
                acc += (s.i * ab[axis] / L) * val
            psi[off: off + len(idx)] = d * acc
        return psi

    def branch_coupling(self, trace: list, start: int = 0) -> np.ndarray:
        """psi of via branches from the vertical trace segments (peec.py:2658-2666): each via's mutual with every
        vertical segment, a 16-point Gauss rule along the via. On the device (gt_via_coupling) unless the stack's
        set-up is the host's; branches that are not vias have none."""
        if self.setup_mode == "host":
            return self.branch_coupling_host(trace, start)
        out = np.zeros(self.nbr - start)
        vias = [(n, b) for n, b in enumerate(self.branches[start:]) if b.kind == "via"]
        if not vias or not trace:
            return out
        ends = np.array([[*self.pos(b.a), *self.pos(b.b), float(b.radius)] for _, b in vias], float)
        val = np.zeros(len(vias))
        segs = _segments(trace)
        self.handle
        _lib.check(_lib.lib().gt_via_coupling(self.ctx.handle, len(vias), _lib.ptr(ends), len(trace), segs,
                                              G.MU0 / (4 * math.pi), _lib.ptr(val)), "gt_via_coupling")
        out[[n for n, _ in vias]] = val
        return out

    def branch_coupling_host(self, trace: list, start: int = 0) -> np.ndarray:
        """branch_coupling in numpy."""
        vert = [s for s in trace if abs((s.b - s.a)[2]) > 0 and np.linalg.norm((s.b - s.a)[:2]) < 1e-12]
        out = np.zeros(self.nbr - start)
        for n, b in enumerate(self.branches[start:]):
            if b.kind == "via" and vert:
                pa, pb = self.pos(b.a), self.pos(b.b)
                out[n] = sum(s.i * G.segment_mutual(s.a, s.b, pa, pb, b.radius) for s in vert)
        return out

    def moment(self, trace: list, Ib: np.ndarray) -> np.ndarray:
        """m = 1/2 sum r x I dl over the closed loop (peec.py:2668-2700)."""
        m = np.zeros(3, complex)
        for s in trace:
            m += 0.5 * s.i * np.cross(s.a, s.b)
        d = self.d
        Ix, Iy = Ib[: self.nxb], Ib[self.nxb: self.ns]
        if self.nxb:
            k, _, j = self.xa.T
            m[1] += 0.5 * np.sum(self.sheet_z[k] * d * Ix)
            m[2] += 0.5 * np.sum(-(j * d) * d * Ix)
        if self.nyb:
            k, i, _ = self.ya.T
            m[0] += 0.5 * np.sum(-self.sheet_z[k] * d * Iy)
            m[2] += 0.5 * np.sum((i * d) * d * Iy)
        for b, cur in zip(self.branches, Ib[self.ns:]):
            m += 0.5 * np.cross(self.pos(b.a), self.pos(b.b)) * cur
        return m


    def loop_terms_dev(self, x_dev, psi_dev, injections: list):
        """The loop's sums over a device solution (gt_loop_terms): (psi . Ib, the strips' and branches' dipole moment
        (3 complex values, the trace's own part left out), sum a V over the injections), with one small read-back."""
        self.handle
        if getattr(self, "_bm_dev", None) is None:
            import torch
            bm = np.array([0.5 * np.cross(self.pos(b.a), self.pos(b.b)) for b in self.branches], float).reshape(-1, 3)
            with torch.cuda.stream(self.torch_stream()):
                self._bm_dev = torch.from_numpy(np.ascontiguousarray(bm)).to(f"cuda:{self.device}")
        if getattr(self, "_keep_pos", None) is None:
            self._keep_pos = np.where(self.keep, np.cumsum(self.keep) - 1, -1)
        kp = self._keep_pos
        terms = [(self.nb + int(kp[n]), float(a)) for n, a in injections if kp[n] >= 0]   # V = 0 off the kept nodes
        idx = np.array([t[0] for t in terms] or [0], np.int32)
        amps = np.array([t[1] for t in terms] or [0.0], float)
        out = np.zeros(10)
        z = np.ascontiguousarray(self.sheet_z, float)
        bm = C.c_void_p(self._bm_dev.data_ptr()) if self.nbr else None
        _lib.check(_lib.lib().gt_loop_terms(self.handle, C.c_void_p(x_dev.data_ptr()), C.c_void_p(psi_dev.data_ptr()),
                                            bm, len(terms), _lib.ptr(idx, _lib.i32p), _lib.ptr(amps),
                                            float(self.d), _lib.ptr(z), _lib.ptr(out)), "gt_loop_terms")
        c = out.view(np.complex128)
        return complex(c[0]), c[1:4].copy(), complex(c[4])


def _segments(trace: list):
    """The trace as the library's segment records."""
    segs = (_lib.Segment * max(1, len(trace)))()
    for k, s in enumerate(trace):
        segs[k].a[:] = [float(v) for v in s.a]
        segs[k].b[:] = [float(v) for v in s.b]
        segs[k].width, segs[k].current = float(s.width), float(s.i)
    return segs


class _DeviceSolution(Solution):
    """A Solution whose solver vector stays on the device: Z_loop and the moment come from the loop's sums on the
    device (gt_loop_terms) and the branch currents from a slice; the vector itself is copied to pinned host memory
    behind the solve on the library's stream, and the strip-current and potential grids are unpacked from it when
    first read."""

    def __init__(self, ex: "RouteExcitation", f: float, x_dev, its: int):
        import torch
        st = ex.stack
        self.f = f
        self.iterations = its
        self.residual = float("nan")
        P, m_grid, v_inj = st.loop_terms_dev(x_dev, ex.psi_dev, ex.injections)
        self.Z_loop = ex.z_loop_terms(f, P, v_inj)
        self.m = ex.trace_moment() + m_grid
        with torch.cuda.stream(st.torch_stream()):
            self.I_branch = x_dev[st.ns:st.nb].cpu().numpy()
            host = torch.empty(x_dev.shape, dtype=x_dev.dtype, pin_memory=True)
            host.copy_(x_dev, non_blocking=True)
            self._done = torch.cuda.Event()
            self._done.record()
        self._host = host
        self._ex = ex
        self._full = None

    @property
    def x(self) -> np.ndarray:
        if self._done is not None:
            self._done.synchronize()
            self._done = None
        return self._host.numpy()

    def _unpacked(self) -> Solution:
        if self._full is None:
            st = self._ex.stack
            x = self.x
            V = np.zeros(st.n_nodes, complex)
            V[st.keep] = x[st.nb:]
            self._full = self._ex.unpack(self.f, x[:st.nb], V, self.iterations, x, moment=False)
        return self._full

    I_x = property(lambda self: self._unpacked().I_x)
    I_y = property(lambda self: self._unpacked().I_y)
    V = property(lambda self: self._unpacked().V)
    V_lid = property(lambda self: self._unpacked().V_lid)
    lid_at = property(lambda self: self._unpacked().lid_at)


class _LazySolution(Solution):
    """A Solution whose grids (I_x, I_y, V) and moment are unpacked from the solver vector on first read."""

    def __init__(self, ex: "RouteExcitation", f: float, Ib: np.ndarray, Vk: np.ndarray, its: int, x: np.ndarray):
        st = ex.stack
        self.f = f
        self.I_branch = Ib[st.ns:].copy()
        self.Z_loop = ex.z_loop(f, Ib, Vk)
        self.iterations = its
        self.x = x
        self.residual = float("nan")
        self._src = (ex, Ib, Vk)
        self._full = None

    def _unpacked(self) -> Solution:
        if self._full is None:
            ex, Ib, Vk = self._src
            self._full = ex.unpack(self.f, Ib, Vk, self.iterations, self.x)
        return self._full

    I_x = property(lambda self: self._unpacked().I_x)
    I_y = property(lambda self: self._unpacked().I_y)
    V = property(lambda self: self._unpacked().V)
    m = property(lambda self: self._unpacked().m)
    V_lid = property(lambda self: self._unpacked().V_lid)
    lid_at = property(lambda self: self._unpacked().lid_at)


def _dist_to_segment(P: np.ndarray, a: np.ndarray, b: np.ndarray) -> np.ndarray:
    ab = b - a
    L2 = float(ab @ ab)
    if L2 == 0:
        return np.linalg.norm(P - a, axis=1)
    t = np.clip((P - a) @ ab / L2, 0, 1)
    return np.linalg.norm(P - (a + t[:, None] * ab), axis=1)


def trace_self(trace: list, trace_t: float) -> float:
    """Partial self inductance of the signal path, sum_ij I_i I_j M_ij (peec.py:3227-3237)."""
    segs = [s for s in trace if np.linalg.norm(s.b - s.a) > 0]
    L = 0.0
    for a, sa in enumerate(segs):
        L += sa.i * sa.i * G.strip_self_gmd(float(np.linalg.norm(sa.b - sa.a)), sa.width, trace_t)
        for sb in segs[a + 1:]:
            L += 2 * sa.i * sb.i * G.segment_mutual(sa.a, sa.b, sb.a, sb.b, G.GMD_FLAT * (sa.width + trace_t))
    return L


def trace_resistance(trace: list, f: float, sigma: float, trace_t: float) -> float:
    """I^2-weighted skin-effect resistance of the horizontal segments (peec.py:3240-3250)."""
    R = 0.0
    for s in trace:
        L = float(np.linalg.norm(s.b - s.a))
        if L == 0 or abs(s.b[2] - s.a[2]) > 0:
            continue
        R += s.i * s.i * L / (sigma * s.width * G.skin_thickness(trace_t, f, sigma))
    return R


class RouteExcitation:
    """A loop on a PlaneStack: the known trace current plus return injections (peec.py:2703-2817)."""

    def __init__(self, stack: PlaneStack, trace: list, inject_driver=None, inject_receiver=None,
                 trace_thickness: float = 35e-6, strip_psi: Optional[np.ndarray] = None,
                 branch_psi: Optional[np.ndarray] = None, injections: Optional[list] = None):
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
        terms += [(node, float(a), f"current injection at {node}") for node, a in self.extra]
        if not terms:
            raise SolverError("no current is injected")
        ids = []
        for node, amps, what in terms:
            n = stack.node_id(node)
            if n < 0:
                raise SolverError(f"{what} is not on plane copper")
            ids.append((int(n), amps))
        bal: dict = {}
        for n, amps in ids:
            c = int(stack.full_label[n])
            bal[c] = bal.get(c, 0.0) + amps
        if any(abs(v) > 1e-9 for v in bal.values()):
            if inject_driver is not None and inject_receiver is not None and not self.extra:
                raise SolverError("receiver return node is not connected to the driver's return node")
            raise SolverError("the injected currents do not balance on every connected piece of copper")
        self.injections = ids
        self._strip_psi_dev = None
        if strip_psi is None and stack.setup_mode == "device" and stack.ns:
            self._strip_psi_dev = stack.strip_coupling_dev(trace)
            strip_psi = self._strip_psi_dev.cpu().numpy()
        self.strip_psi = stack.strip_coupling(trace) if strip_psi is None else strip_psi
        bp = branch_psi if branch_psi is not None else np.zeros(0)
        if len(bp) < stack.nbr:
            bp = np.concatenate([bp, stack.branch_coupling(trace, start=len(bp))])
        self.branch_psi = bp
        self.psi = np.concatenate([self.strip_psi, self.branch_psi])
        self.L_trace = trace_self(trace, trace_thickness) if trace else 0.0

    def on(self, stack: PlaneStack) -> "RouteExcitation":
        """The same loop on a what-if stack (with_branches): the strip coupling is reused, the added branches'
        coupling computed (peec.py:2771-2774)."""
        drv, rcv = self.inject
        ex = RouteExcitation(stack, self.trace, drv, rcv, self.trace_t, self.strip_psi, self.branch_psi, self.extra)
        if len(ex.strip_psi) == len(self.strip_psi):
            ex._strip_psi_dev = self._strip_psi_dev       # the same strips: their psi stays shared on the device
        return ex

    def rhs(self, f: float):
        """(rhs_top = -jw psi, J) at f (peec.py:2789-2795)."""
        w = 2 * math.pi * f
        J = np.zeros(self.stack.n_nodes)
        for n, amps in self.injections:
            J[n] += amps
        return -1j * w * self.psi, J

    def solve(self, f: float, guess: Optional[list] = None, tol: float = 1e-5, max_cycles: int = 8,
              restart: Optional[int] = None) -> Solution:
        """The loop's solution at f (peec.py:2787-2817). tol / max_cycles / restart go to PlaneStack.solve
        (the reference uses its defaults)."""
        rhs_top, J = self.rhs(f)
        Ib, V, its, x = self.stack.solve(f, rhs_top, J, tol=tol, max_cycles=max_cycles, guess=guess,
                                         restart=restart)
        return self.unpack(f, Ib, V, its, x)

    def solve_many(self, fs: list, tol: float = 1e-5, max_cycles: int = 8, restart: Optional[int] = None,
                   precision: str = "fp64") -> list:
        """The loop's solutions at every frequency of `fs`, solved as one batch on the device (each the same as
        solve(f) alone); precision "mixed": FP32 inner solves under FP64 refinement to tol."""
        rhs = [self.rhs(f) for f in fs]
        sols = self.stack.solve_many(fs, [r[0] for r in rhs], [r[1] for r in rhs], tol=tol, max_cycles=max_cycles,
                                     restart=restart, precision=precision)
        return [self.unpack(f, *s) for f, s in zip(fs, sols)]

    def z_loop(self, f: float, Ib: np.ndarray, V: np.ndarray) -> complex:
        """Z_loop of a solution (peec.py:2787-2800): the trace's resistance and the injection nodes' potentials, plus
        j w (L_trace + psi . Ib)."""
        v_inj = sum(amps * V[n] for n, amps in self.injections)
        return self.z_loop_terms(f, complex(np.dot(self.psi, Ib)), v_inj)

    def z_loop_terms(self, f: float, psi_dot_Ib: complex, v_inj: complex) -> complex:
        """Z_loop from its two solution sums: R_trace + sum a V + j w (L_trace + psi . Ib)."""
        w = 2 * math.pi * f
        return complex(trace_resistance(self.trace, f, self.stack.sigma, self.trace_t) + v_inj
                       + 1j * w * (self.L_trace + psi_dot_Ib))

    def trace_moment(self) -> np.ndarray:
        """The trace's own part of the loop's dipole moment, 1/2 sum i a x b over its segments (peec.py:2668-2700)."""
        m = np.zeros(3, complex)
        for s in self.trace:
            m += 0.5 * s.i * np.cross(s.a, s.b)
        return m

    @property
    def psi_dev(self):
        """psi on the device (torch float64, nb values), made once per excitation: the strips' part as the device
        computed it (shared by the what-ifs of one board), the branches' uploaded."""
        if getattr(self, "_psi_dev", None) is None:
            import torch
            dev = f"cuda:{self.stack.device}"
            with torch.cuda.stream(self.stack.torch_stream()):     # ordered before the library's use of it
                if self._strip_psi_dev is None:
                    self._strip_psi_dev = torch.from_numpy(np.ascontiguousarray(self.strip_psi, float)).to(dev)
                br = torch.from_numpy(np.ascontiguousarray(self.branch_psi, float)).to(dev)
                self._psi_dev = torch.cat([self._strip_psi_dev, br])
        return self._psi_dev

    def unpack_dev(self, f: float, x_dev, its: int) -> "Solution":
        """The Solution of a solver vector on the device (torch complex128, its work ordered on the library's stream):
        Z_loop, the moment and the branch currents made now on the device, the rest when first read."""
        return _DeviceSolution(self, f, x_dev, its)

    def unpack_lazy(self, f: float, x: np.ndarray, its: int) -> "Solution":
        """The Solution of the solver vector x with Z_loop and the branch currents made now and the strip-current
        and potential grids and the moment made when first read."""
        st = self.stack
        Ib = x[:st.nb]
        V = np.zeros(st.n_nodes, complex)
        V[st.keep] = x[st.nb:]
        return _LazySolution(self, f, Ib, V, its, x)

    def unpack(self, f: float, Ib: np.ndarray, V: np.ndarray, its: int, x=None, moment: bool = True) -> Solution:
        st = self.stack
        I_x = np.zeros((st.N, st.nx - 1, st.ny), complex)
        I_y = np.zeros((st.N, st.nx, st.ny - 1), complex)
        I_x[tuple(st.xa.T)] = Ib[:st.nxb]
        I_y[tuple(st.ya.T)] = Ib[st.nxb: st.ns]
        Vg = np.zeros((st.N, st.nx, st.ny), complex)
        Vg[np.nonzero(st.node)] = V[:st.n_grid]
        Z = self.z_loop(f, Ib, V)
        m = st.moment(self.trace, Ib) if moment else None
        # the parts' lids after the grid's nodes (peec.py:2807-2816)
        lid_at = {k: v - st.n_grid for k, v in st.lid_nid.items()} if st.lid_nid else None
        V_lid = V[st.n_grid:].copy() if st.lid_nid else None
        return Solution(f, I_x, I_y, Vg, Ib[st.ns:].copy(), Z, m, its, x, V_lid=V_lid, lid_at=lid_at)
