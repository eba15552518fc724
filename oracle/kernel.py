"""Oracle restatement of `src/groundtruth/solver/kernel.py`: partial inductances on a square grid, slab
impedances and the vector potential of straight current filaments (all lengths in metres).

TEST INFRASTRUCTURE (see oracle/__init__.py): only tests, smoke() and bench.py's cpu_baseline use it.

Every function cites the reference line it follows. Pins (tests/test_oracle.py): golden vectors written from the
reference module itself (tests/golden/make_kernel_vectors.py), scipy adaptive quadrature of Neumann's integral,
closed forms (coth slab form, filament far field, plane-pair sheet inductance mu0*dz).
"""
from __future__ import annotations

from functools import lru_cache

import numpy as np
from numpy.polynomial import legendre

MU0 = 4e-7 * np.pi              # kernel.py:19
GMD_FLAT = 0.2235               # kernel.py:20, Grover's GMD of a thin flat strip over (width + thickness)
TAPE_PANELS = 30                # kernel.py:218
TAPE_ORDER = 8                  # kernel.py:219
NEAR_CELLS = 4                  # kernel.py:286, offsets up to this many cells use the exact width integral


def _skin_x(t, f, sigma):
    """x = t / delta with delta = 1/sqrt(pi f mu0 sigma); 0 at f <= 0 (kernel.py:37-38)."""
    t = np.asarray(t, float)
    f = np.asarray(f, float)
    with np.errstate(divide="ignore", over="ignore", invalid="ignore"):
        delta = 1.0 / np.sqrt(np.pi * np.maximum(f, 1e-300) * MU0 * sigma)
        return np.where(f > 0, t / delta, 0.0)


def skin_thickness(t, f, sigma):
    """Effective conducting thickness t / F with Dowell's one-sided factor F (kernel.py:23-44)."""
    t = np.asarray(t, float)
    x = _skin_x(t, f, sigma)
    with np.errstate(divide="ignore", over="ignore", invalid="ignore"):
        y = np.minimum(2.0 * x, 60.0)
        full = x * (np.sinh(y) + np.sin(y)) / (np.cosh(y) - np.cos(y))
        F = np.where(x < 0.1, 1.0 + 4.0 * x ** 4 / 45.0, full)
    out = t / F
    return float(out) if out.ndim == 0 else out


def slab_impedance(t, f, sigma):
    """One-face slab internal impedance per square, (1+j)x coth((1+j)x) / (sigma t) with the thin-foil series
    below x = 0.1 and coth's argument capped at x = 30 (kernel.py:47-75)."""
    t = np.asarray(t, float)
    x = _skin_x(t, f, sigma)
    with np.errstate(divide="ignore", over="ignore", invalid="ignore"):
        u = (1 + 1j) * np.minimum(x, 30.0)
        closed = (1 + 1j) * x / np.tanh(np.where(x > 0, u, 1.0))
        thin = 1.0 + 4.0 * x ** 4 / 45.0 + 1j * (2.0 * x ** 2 / 3.0)
        z = np.where(x < 0.1, thin, closed)
    out = z / (sigma * t)
    return complex(out) if out.ndim == 0 else out


@lru_cache(maxsize=None)
def leggauss(n: int):
    """Gauss-Legendre nodes and weights on [-1, 1] (kernel.py:155-162)."""
    x, w = legendre.leggauss(n)
    x.setflags(write=False)
    w.setflags(write=False)
    return x, w


def _antideriv(u, rho):
    """G(u) with G'' = 1/sqrt(u^2 + rho^2) (kernel.py:165-167)."""
    return u * np.arcsinh(u / rho) - np.sqrt(u * u + rho * rho)


def filament_mutual(length: float, axial, rho):
    """Neumann mutual of two parallel equal filaments as a second difference of G (kernel.py:170-176)."""
    rho = np.maximum(rho, 1e-12 * length)
    return MU0 / (4 * np.pi) * (_antideriv(axial + length, rho) - 2.0 * _antideriv(axial, rho)
                                + _antideriv(axial - length, rho))


def strip_self(length: float, width: float, thickness: float = 0.0) -> float:
    """Self inductance at Grover's GMD 0.2235 (w + t) (kernel.py:179-184)."""
    return float(filament_mutual(length, np.array(0.0), np.array(GMD_FLAT * (width + thickness))))


def _panel_rule(p: float, q: float, toward, panels: int, order: int):
    """Composite Gauss-Legendre on [p, q]: equal panels, or panels halving toward the end `toward`
    (kernel.py:222-236)."""
    x, w = leggauss(order)
    if toward is None:
        edges = np.linspace(0.0, 1.0, panels + 1)
    else:
        edges = np.concatenate([[0.0], 0.5 ** np.arange(panels - 1, -1, -1)])
        if toward == q:
            edges = 1.0 - edges[::-1]
    lo, hi = edges[:-1], edges[1:]
    nodes = p + (q - p) * (lo[:, None] + (hi - lo)[:, None] * (x[None, :] + 1.0) / 2.0)
    weights = (q - p) * (hi - lo)[:, None] * w[None, :] / 2.0
    return nodes.ravel(), weights.ravel()


@lru_cache(maxsize=1)
def _unit_rule():
    """The (0, 1) rule graded toward 0 for the self integral (kernel.py:209-215)."""
    s, ws = _panel_rule(0.0, 1.0, 0.0, TAPE_PANELS, TAPE_ORDER)
    s.setflags(write=False)
    ws.setflags(write=False)
    return s, ws


def strip_self_numeric(length: float, width):
    """2 int_0^1 (1 - s) M_fil(w s) ds: the flat strip's self inductance (kernel.py:187-206)."""
    s, ws = _unit_rule()
    w = np.asarray(width, float)
    flat = np.atleast_1d(w).ravel()
    out = np.empty(len(flat))
    for a in range(0, len(flat), 2048):
        u = flat[a:a + 2048, None] * s[None, :]
        out[a:a + 2048] = 2.0 * (filament_mutual(length, 0.0, u) * ((1.0 - s) * ws)[None, :]).sum(axis=1)
    return float(out[0]) if w.ndim == 0 else out.reshape(w.shape)


def tape_mutual(length: float, wa: float, wb: float, axial: float, lateral: float, dz: float) -> float:
    """Mutual of two parallel flat tapes: the trapezoid-weighted 1-D integral over the widths' offset v, split at
    the trapezoid's corners and at the filaments' crossing v = -lateral (graded there) (kernel.py:239-265)."""
    h, k, s = 0.5 * (wa + wb), 0.5 * abs(wa - wb), -lateral
    cuts = {-h, -k, k, h}
    if -h < s < h:
        cuts.add(s)
    pts = sorted(cuts)
    total = 0.0
    for p, q in zip(pts[:-1], pts[1:]):
        if q - p <= 0.0:
            continue
        toward = p if p == s else q if q == s else None
        v, wt = _panel_rule(p, q, toward, TAPE_PANELS if toward is not None else 4, TAPE_ORDER)
        overlap = np.clip(np.minimum(0.5 * wa, v + 0.5 * wb) - np.maximum(-0.5 * wa, v - 0.5 * wb), 0.0, None)
        rho = np.sqrt((lateral + v) ** 2 + dz * dz)
        total += float(np.sum(wt * overlap * filament_mutual(length, np.full(len(v), axial), rho)))
    return total / (wa * wb)


def strip_kernel(n_axial: int, n_lateral: int, d: float, dz: float = 0.0) -> np.ndarray:
    """Mutual between square strips at every offset in [-(n-1), n-1]^2: filaments in the far field, tape_mutual for
    |offsets| <= 4 cells (kernel.py:268-294). Centre = zero offset."""
    ia = np.arange(-(n_axial - 1), n_axial) * d
    il = np.arange(-(n_lateral - 1), n_lateral) * d
    A, L = np.meshgrid(ia, il, indexing="ij")
    K = filament_mutual(d, A, np.sqrt(L * L + dz * dz))
    ca, cl = n_axial - 1, n_lateral - 1
    for a in range(0, min(NEAR_CELLS, n_axial - 1) + 1):
        for b in range(0, min(NEAR_CELLS, n_lateral - 1) + 1):
            m = tape_mutual(d, d, d, a * d, b * d, dz)
            for ra in {ca + a, ca - a}:
                for rb in {cl + b, cl - b}:
                    K[ra, rb] = m
    return K


def segment_vector_potential(a: np.ndarray, b: np.ndarray, points: np.ndarray, rho_min: float = 1e-9) -> np.ndarray:
    """|A| per ampere of a straight filament a->b at points (N, 3): mu0/4pi (asinh((L-pu)/rho) + asinh(pu/rho))
    (kernel.py:297-312)."""
    ab = b - a
    L = np.linalg.norm(ab)
    if L == 0:
        return np.zeros(len(points))
    u = ab / L
    p = points - a
    pu = p @ u
    rho2 = np.einsum("ij,ij->i", p, p) - pu * pu
    rho = np.sqrt(np.maximum(rho2, rho_min * rho_min))
    return MU0 / (4 * np.pi) * (np.arcsinh((L - pu) / rho) + np.arcsinh(pu / rho))


def segment_mutual(a1, b1, a2, b2, rho_min: float, n: int = 16) -> float:
    """Neumann mutual of two straight filaments: the first's vector potential integrated along the second with an
    n-point Gauss rule (kernel.py:315-337)."""
    a2 = np.asarray(a2, float)
    b2 = np.asarray(b2, float)
    d2 = b2 - a2
    L2 = np.linalg.norm(d2)
    if L2 == 0:
        return 0.0
    x, w = leggauss(n)
    pts = a2 + (0.5 * (x + 1))[:, None] * d2
    a1 = np.asarray(a1, float)
    b1 = np.asarray(b1, float)
    L1 = np.linalg.norm(b1 - a1)
    if L1 == 0:
        return 0.0
    psi = segment_vector_potential(a1, b1, pts, rho_min)
    return float(np.dot((b1 - a1) / L1, d2 / L2) * 0.5 * L2 * np.sum(w * psi))
