"""Copper geometries for the rasteriser's tests, built without shapely (groundtruth_amd.geometry), after the
reference's own raster tests (tests/test_raster.py, tests/test_track_width.py, tests/test_validation_audit.py SP-27,
tests/test_vectors.py, tests/test_pad_bond.py).
"""
import math

import numpy as np

from groundtruth_amd.geometry import MultiPolygon, Polygon, box, disc

PLANE = (0.0, 0.0, 20.0, 10.0)


def path_outline(pts, half: float) -> np.ndarray:
    """The outline of a polyline widened by `half` on each side with mitred joins and flat ends (a track's copper),
    counter-clockwise."""
    P = np.asarray(pts, float)
    seg = np.diff(P, axis=0)
    nrm = np.column_stack([-seg[:, 1], seg[:, 0]]) / np.hypot(seg[:, 0], seg[:, 1])[:, None]
    left, right = [], []
    for k in range(len(P)):
        if k == 0:
            n = nrm[0]
            s = 1.0
        elif k == len(P) - 1:
            n = nrm[-1]
            s = 1.0
        else:
            m = nrm[k - 1] + nrm[k]
            n = m / np.hypot(*m)
            s = 1.0 / float(n @ nrm[k])
        left.append(P[k] + half * s * n)
        right.append(P[k] - half * s * n)
    return np.array(right + left[::-1])


def webbed_plane():
    """test_raster.py: 0.36 mm antipads every 0.5 mm along x = 10 (inside the plane) leave 0.14 mm webs."""
    return box(*PLANE, holes=[disc(10.0, 0.5 * k, 0.18, 32)[::-1] for k in range(1, 20)])


def slotted_plane():
    """test_raster.py: a 0.3 mm slot across the plane."""
    return MultiPolygon([box(0, 0, 9.85, 10), box(10.15, 0, 20, 10)])


def perforated_plane():
    """test_raster.py (capacitance test): a 16 x 6 field of 0.6 mm holes on a 1 mm pitch."""
    return box(*PLANE, holes=[disc(2 + i, 2 + j, 0.3, 32)[::-1] for i in range(16) for j in range(6)])


def track_through_pour():
    """test_raster.py: a 0.4 mm track of net A with 45 degree bends through net B's pour, 0.25 mm clearance, pads at
    both ends (the pads' clearance folded into the track's)."""
    pts = [(1.0, 2.0), (6.0, 2.0), (11.0, 7.0), (19.0, 7.0), (19.0, 9.0)]
    track = Polygon(path_outline(pts, 0.2))
    pour = box(*PLANE, holes=[path_outline(pts, 0.45)[::-1]])
    return pts, {"B": pour, "A": track}


def narrow_strip(y: float, w: float = 0.35):
    """SP-27: a strip of width w along x."""
    return {"N": box(2.0, y - w / 2, 18.0, y + w / 2)}


def angled_track(theta_deg: float, w: float = 0.3):
    """A straight track at an angle (TW3)."""
    t = math.radians(theta_deg)
    a = np.array([3.0, 3.0])
    b = a + 12.0 * np.array([math.cos(t), math.sin(t)])
    return [tuple(a), tuple(b)], {"T": Polygon(path_outline([a, b], w / 2))}


def necked_track():
    """TW4: a 0.4 mm track narrowing to 0.1 mm over 4 mm."""
    outline = [(2, 4.8), (8, 4.8), (8, 4.95), (12, 4.95), (12, 4.8), (18, 4.8), (18, 5.2), (12, 5.2), (12, 5.05),
               (8, 5.05), (8, 5.2), (2, 5.2)]
    return {"N": Polygon(outline)}


def two_nets_and_island():
    """test_vectors.py: a pour with a separate island of one net and a strip of another."""
    return {"A": MultiPolygon([box(0, 0, 10, 4), box(14, 0, 15, 0.6)]), "B": box(0, 6, 20, 6.4)}


def touching_fills():
    """Two fills of one net that only touch along an edge, and a third 0.1 mm away (_pieces, TW12): the first two
    are one piece, the third its own."""
    return {"P": MultiPolygon([box(0, 0, 5, 5), box(5, 0, 10, 5), box(10.1, 0, 15, 5)])}


def diagonal_lattice():
    """Edges at 45 degrees through lattice points and boxes on lattice lines: the exact on-edge cases."""
    diamond = Polygon([(5, 1), (9, 5), (5, 9), (1, 5)])
    return {"D": diamond, "R": box(10, 1, 16, 8, holes=[[(12, 3), (12, 6), (14, 6), (14, 3)]])}


def overlapping_nets():
    """Two nets' copper overlapping (the first to claim a sample keeps it)."""
    return {"X": box(1, 1, 9, 6), "Y": box(5, 3, 15, 9)}
