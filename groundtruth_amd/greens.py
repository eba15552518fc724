"""Set-up of the partial-inductance operator: the Green's-function kernels of square strips on the grid, their
spectra for the device FFT sizes, and the impedances the per-frequency operator needs.

The mutual partial inductance between two parallel square strips depends only on their grid offset and the gap dz
between their sheets (src/groundtruth/solver/kernel.py:268-294 `strip_kernel`): far away, two filaments along the
strips' centre lines (Neumann's integral in closed form, kernel.py:165-176); within 4 cells, the exact mean over the
two strips' widths, a one-dimensional integral over the widths' offset weighted by their overlap, on Gauss-Legendre
panels graded toward the point where the filaments cross (kernel.py:222-265); the strip itself the flat strip's
exact self inductance. The quadrature rule (30 geometric panels of 8 points toward the crossing, 4 equal panels
elsewhere) is the reference's, because it fixes the discretisation's numbers to ~1e-10.

Differences from the reference's set-up, none of which changes the operator: the FFT sizes are powers of two
(>= 2n - 1, the size the device FFT runs; the circular embedding of the kernel is a linear convolution for any
size >= 2n - 1), the spectrum's quadrant is stored pixel-major over sheet pairs as the device reads it, and the
1 / (P0 P1) of the inverse FFT is folded into it.
"""
from __future__ import annotations

import math
from functools import lru_cache

import numpy as np

MU0 = 4e-7 * math.pi
GMD_FLAT = 0.2235
PANELS, ORDER, NEAR = 30, 8, 4
DIGITS = 12


def key(v: float) -> float:
    """A length or frequency as the kernels and operators are computed for it: 12 significant digits
    (peec.py:3264-3273), so a value is a function of its key alone."""
    return float(f"{float(v):.{DIGITS}g}")


def fft_size(n: int) -> int:
    """The device FFT length for n points of a linear convolution: the power of two >= 2n - 1."""
    return 1 << max(1, int(math.ceil(math.log2(max(2 * n - 1, 2)))))


def filament(length: float, axial, rho):
    """Mutual of two parallel filaments `length` long, axial offset, distance rho: the second difference of
    u asinh(u/rho) - sqrt(u^2 + rho^2) over the ends' offsets [Grover 1946]."""
    rho = np.maximum(rho, 1e-12 * length)

    def g(u):
        return u * np.arcsinh(u / rho) - np.sqrt(u * u + rho * rho)
    return MU0 / (4 * math.pi) * (g(axial + length) - 2.0 * g(axial) + g(axial - length))


@lru_cache(maxsize=None)
def _gauss(order: int):
    x, w = np.polynomial.legendre.leggauss(order)
    return x, w


def _rule(p: float, q: float, graded_to, panels: int):
    """Gauss-Legendre nodes/weights on [p, q]: `panels` equal panels, or panels halving toward `graded_to`."""
    x, w = _gauss(ORDER)
    if graded_to is None:
        e = np.linspace(0.0, 1.0, panels + 1)
    else:
        e = np.concatenate([[0.0], 0.5 ** np.arange(panels - 1, -1, -1)])
        if graded_to == q:
            e = 1.0 - e[::-1]
    a, b = e[:-1, None], e[1:, None]
    nodes = p + (q - p) * (a + (b - a) * (x[None, :] + 1.0) / 2.0)
    wts = (q - p) * (b - a) * w[None, :] / 2.0
    return nodes.ravel(), wts.ravel()


def strip_self_numeric(length: float, width):
    """Self inductance of a flat strip (zero thickness) with its current even across it: 2 int_0^1 (1 - s)
    M_fil(w s) ds, on PANELS panels graded toward s = 0 (kernel.py:187-206); `width` may be an array."""
    s, ws = _rule(0.0, 1.0, 0.0, PANELS)
    w = np.atleast_1d(np.asarray(width, float))
    out = np.empty(w.size)
    flat = w.ravel()
    for a in range(0, len(flat), 2048):
        u = flat[a:a + 2048, None] * s[None, :]
        out[a:a + 2048] = 2.0 * (filament(length, 0.0, u) * ((1.0 - s) * ws)[None, :]).sum(axis=1)
    return float(out[0]) if np.ndim(width) == 0 else out.reshape(w.shape)


NEAR_AXIAL = 2     # peec.py:3325: side-by-side partial strips corrected up to this many strips along


def narrow_self(d: float, phi, t):
    """What a strip d long and phi d wide of copper t thick adds to the full square strip's self term (peec.py:3309-
    3322): the flat-strip integral's change, and the thickness's change from the GMD formula [Grover 1946]."""
    phi, t = np.asarray(phi, float), np.asarray(t, float)

    def thick(w):
        return filament(d, 0.0, GMD_FLAT * w) - filament(d, 0.0, GMD_FLAT * (w + t))
    return strip_self_numeric(d, phi * d) - strip_self_numeric(d, d) - thick(phi * d) + thick(d)


def side_mutual(length: float, wa, wb, gap, steps=0.0, n: int = 8) -> np.ndarray:
    """Mutual of two flat coplanar parallel strips `length` long, wa and wb wide, centres gap across and `steps`
    lengths along apart (fractions of length; arrays broadcast): the mean filament mutual of n filaments across each,
    touching filaments a filament's GMD apart (peec.py:3328-3341)."""
    wa, wb, gap, steps = (np.atleast_1d(v) * 1.0 for v in np.broadcast_arrays(*(np.asarray(v, float)
                                                                                 for v in (wa, wb, gap, steps))))
    wa, wb, gap = wa * length, wb * length, gap * length
    u = (np.arange(n) + 0.5) / n - 0.5
    ya = wa[:, None, None] * u[None, :, None]
    yb = gap[:, None, None] + wb[:, None, None] * u[None, None, :]
    rho = np.maximum(np.abs(yb - ya), GMD_FLAT * np.maximum(wa, wb)[:, None, None] / n)
    axial = np.broadcast_to((steps * length)[:, None, None], rho.shape)
    return filament(length, axial, rho).mean(axis=(1, 2))


def tape(d: float, axial: float, lateral: float, dz: float) -> float:
    """Mutual of two parallel square tapes of side d (zero thickness), centres `axial` apart along them and
    `lateral` across, in planes dz apart: (1/d^2) int W(v) M_fil(d, axial, |(lateral + v, dz)|) dv with W the
    overlap of the two widths at offset v, split at the overlap's corners and at v = -lateral."""
    s = -lateral
    cuts = sorted({-d, 0.0, d} | ({s} if -d < s < d else set()))
    total = 0.0
    for p, q in zip(cuts[:-1], cuts[1:]):
        if q <= p:
            continue
        toward = p if p == s else (q if q == s else None)
        v, w = _rule(p, q, toward, PANELS if toward is not None else 4)
        W = np.clip(np.minimum(0.5 * d, v + 0.5 * d) - np.maximum(-0.5 * d, v - 0.5 * d), 0.0, None)
        total += float(np.sum(w * W * filament(d, np.full(v.shape, axial), np.sqrt((lateral + v) ** 2 + dz * dz))))
    return total / (d * d)


def strip_kernel(n_ax: int, n_lat: int, d: float, dz: float) -> np.ndarray:
    """Mutual partial inductance of square strips (axis 0 along them) at every offset in [-(n-1), n-1]; centre =
    zero offset. Filaments in the far field, tapes for |offsets| <= 4 cells."""
    a = np.arange(-(n_ax - 1), n_ax) * d
    l = np.arange(-(n_lat - 1), n_lat) * d
    K = filament(d, a[:, None], np.sqrt(l[None, :] ** 2 + dz * dz))
    ca, cl = n_ax - 1, n_lat - 1
    for ia in range(min(NEAR, n_ax - 1) + 1):
        for il in range(min(NEAR, n_lat - 1) + 1):
            m = tape(d, ia * d, il * d, dz)
            for r in {ca - ia, ca + ia}:
                for c in {cl - il, cl + il}:
                    K[r, c] = m
    return K


def quadrant(K: np.ndarray, P0: int, P1: int) -> np.ndarray:
    """The real quadrant [0, P0/2] x [0, P1/2] of the FFT of the centred kernel embedded circularly in (P0, P1)
    (its centre at (0, 0): negative offsets wrap to the end). The kernel is even, so the spectrum is real and even."""
    Z = np.zeros((P0, P1))
    c0, c1 = K.shape[0] // 2, K.shape[1] // 2
    # offsets 0..c go to 0..c, offsets -c..-1 to P-c..P-1
    Z[: c0 + 1, : c1 + 1] = K[c0:, c1:]
    Z[: c0 + 1, P1 - c1:] = K[c0:, :c1]
    Z[P0 - c0:, : c1 + 1] = K[:c0, c1:]
    Z[P0 - c0:, P1 - c1:] = K[:c0, :c1]
    return np.fft.fft2(Z).real[: P0 // 2 + 1, : P1 // 2 + 1]


def slab_impedance(t: float, f: float, sigma: float) -> complex:
    """Internal impedance per square of a foil t thick with its field on one face: (1+j) x coth((1+j) x) / (sigma t),
    x = t / skin depth, thin-foil series below x = 0.1, coth's argument capped at x = 30 (kernel.py:47-75)."""
    if f <= 0:
        return complex(1.0 / (sigma * t))
    delta = 1.0 / math.sqrt(math.pi * f * MU0 * sigma)
    x = t / delta
    if x < 0.1:
        z = complex(1.0 + 4.0 * x ** 4 / 45.0, 2.0 * x ** 2 / 3.0)
    else:
        u = complex(1.0, 1.0) * min(x, 30.0)
        z = complex(1.0, 1.0) * x / np.tanh(u)
    return complex(z) / (sigma * t)


def skin_thickness(t: float, f: float, sigma: float) -> float:
    """The thickness a foil conducts in for its resistance, t / F with Dowell's one-sided F (kernel.py:23-44)."""
    if f <= 0:
        return t
    delta = 1.0 / math.sqrt(math.pi * f * MU0 * sigma)
    x = t / delta
    if x < 0.1:
        F = 1.0 + 4.0 * x ** 4 / 45.0
    else:
        y = min(2.0 * x, 60.0)
        F = x * (math.sinh(y) + math.sin(y)) / (math.cosh(y) - math.cos(y))
    return t / F


def strip_self_gmd(length: float, width: float, thickness: float = 0.0) -> float:
    """Self inductance of a flat strip as the filament mutual at its GMD 0.2235 (w + t) [Grover 1946]."""
    return float(filament(length, 0.0, GMD_FLAT * (width + thickness)))


def segment_potential(a: np.ndarray, b: np.ndarray, pts: np.ndarray, rho_min: float) -> np.ndarray:
    """|A| per amp of the straight filament a -> b at points (N, 3): mu0/4pi (asinh((L - s)/rho) + asinh(s/rho)),
    s the points' position along the axis, rho their distance from it (at least rho_min) [Jackson 1999, 5.4]."""
    ab = b - a
    L = float(np.linalg.norm(ab))
    if L == 0:
        return np.zeros(len(pts))
    u = ab / L
    p = pts - a
    s = p @ u
    rho = np.sqrt(np.maximum(np.einsum("ij,ij->i", p, p) - s * s, rho_min * rho_min))
    return MU0 / (4 * math.pi) * (np.arcsinh((L - s) / rho) + np.arcsinh(s / rho))


def segment_mutual(a1, b1, a2, b2, rho_min: float, n: int = 16) -> float:
    """Mutual of two straight filaments: the first's vector potential integrated along the second by an n-point
    Gauss rule, times the cosine between them [Paul 2010]."""
    a1, b1, a2, b2 = (np.asarray(v, float) for v in (a1, b1, a2, b2))
    d2 = b2 - a2
    L2 = float(np.linalg.norm(d2))
    L1 = float(np.linalg.norm(b1 - a1))
    if L2 == 0 or L1 == 0:
        return 0.0
    x, w = _gauss(n)
    pts = a2 + (0.5 * (x + 1.0))[:, None] * d2
    psi = segment_potential(a1, b1, pts, rho_min)
    return float(np.dot((b1 - a1) / L1, d2 / L2) * 0.5 * L2 * np.sum(w * psi))
