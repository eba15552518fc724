"""Write tests/golden/kernel_vectors.npz: input/output vectors of the reference's
`src/groundtruth/solver/kernel.py` (partial inductances, slab impedances, vector potentials).

Run in a container that has the reference checked out (default /root/reference; override with GT_REFERENCE):

    python tests/golden/make_kernel_vectors.py

kernel.py needs only numpy, so it is loaded by file path on its own (no package import, nothing else of the
reference runs). The output is data only: inputs and the reference's outputs, which tests/test_oracle.py compares
the oracle against. Nothing at test time reads the reference.
"""
from __future__ import annotations

import importlib.util
import os
from pathlib import Path

import numpy as np

REF = Path(os.environ.get("GT_REFERENCE", "/root/reference")) / "src" / "groundtruth" / "solver" / "kernel.py"
OUT = Path(__file__).with_name("kernel_vectors.npz")


def load():
    spec = importlib.util.spec_from_file_location("gt_reference_kernel", REF)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def main():
    k = load()
    rng = np.random.default_rng(7)
    out = {}
    # filament mutuals over a spread of offsets, including coincident filaments (rho -> 1e-12 length)
    L = np.array([1e-3, 0.5e-3, 0.25e-3, 2e-3, 1e-3, 1e-3])
    ax = np.array([0.0, 0.5e-3, 1e-3, -3e-3, 7e-3, 0.0])
    rho = np.array([0.2e-3, 1e-3, 0.0, 0.4e-3, 2.5e-3, 1e-15])
    out["fil_in"] = np.stack([L, ax, rho], 1)
    out["fil_out"] = np.array([float(k.filament_mutual(a, np.array(b), np.array(c))) for a, b, c in zip(L, ax, rho)])
    # flat-strip self (numeric integral) and GMD self
    widths = np.array([1e-3, 0.5e-3, 0.1e-3, 2e-3, 0.25e-3])
    out["ssn_in"] = widths
    out["ssn_out"] = k.strip_self_numeric(1e-3, widths)
    out["ss_in"] = np.array([[1e-3, 0.2e-3, 35e-6], [5e-3, 0.15e-3, 35e-6], [0.5e-3, 0.5e-3, 0.0]])
    out["ss_out"] = np.array([k.strip_self(a, b, c) for a, b, c in out["ss_in"]])
    # tapes: the reference's own audit cases (tests/test_numerics_audit.py:309-310) in units of d = 0.5 mm, plus
    # the kernel's near-field offsets at a few layer gaps
    d = 0.5e-3
    cases = [(1, 1, 0, 0, 0), (1, 1, 1, 0, 0), (1, 1, 0, 1, 0), (1, 1, 0, 0, 0.05), (1, 1, 1, 1, 0.02),
             (0.3, 1, 0, 0.6, 0), (0.1, 0.7, 2, 0.4, 0.001), (1, 1, 4, 3, 0.4), (1, 1, 2, 4, 0.47), (1, 1, 3, 0, 0.1)]
    out["tape_in"] = np.array(cases, float) * d
    out["tape_out"] = np.array([k.tape_mutual(d, *c) for c in out["tape_in"]])
    # whole strip kernels (far and near field) at three shapes and layer gaps
    kin = [(6, 6, 1e-3, 0.0), (9, 7, 0.5e-3, 0.2e-3), (5, 8, 0.25e-3, 0.035e-3), (12, 10, 0.25e-3, 0.2e-3)]
    out["sk_in"] = np.array(kin, float)
    for n, (a, b, dd, dz) in enumerate(kin):
        out[f"sk_out{n}"] = k.strip_kernel(a, b, dd, dz)
    # slab impedance and skin thickness from DC to 20 GHz for four foil thicknesses
    f = np.concatenate([[0.0], np.logspace(2, np.log10(2e10), 41)])
    t = np.array([18e-6, 35e-6, 70e-6, 210e-6])
    out["slab_f"], out["slab_t"] = f, t
    out["slab_out"] = np.array([[k.slab_impedance(tt, ff, 5.8e7) for ff in f] for tt in t])
    out["skin_out"] = np.array([[k.skin_thickness(tt, ff, 5.8e7) for ff in f] for tt in t])
    # vector potential of a segment at random points (some on its axis: rho_min), and segment mutuals
    a, b = np.array([1e-3, 2e-3, 0.2e-3]), np.array([6e-3, 2e-3, 0.2e-3])
    pts = rng.uniform(0, 8e-3, (40, 3)) * np.array([1, 1, 0.1])
    pts[:4] = [[3e-3, 2e-3, 0.2e-3], [0.0, 2e-3, 0.2e-3], [1e-3, 2e-3, 0.2e-3], [3e-3, 2.1e-3, 0.0]]
    out["svp_seg"] = np.stack([a, b])
    out["svp_pts"] = pts
    out["svp_out"] = k.segment_vector_potential(a, b, pts, 1e-7)
    segs = [((0, 0, 0), (1e-3, 0, 0), (0, 0.5e-3, 0), (1e-3, 0.5e-3, 0), 1e-5),
            ((0, 0, 0), (0, 0, 1e-3), (0.3e-3, 0, 0), (0.3e-3, 0, 1.6e-3), 1.5e-4),
            ((0, 0, 0), (2e-3, 0, 0), (2e-3, 0, 0), (2e-3, 3e-3, 0), 1e-4)]
    out["sm_in"] = np.array([np.concatenate([s[0], s[1], s[2], s[3], [s[4]]]) for s in segs], float)
    out["sm_out"] = np.array([k.segment_mutual(*s) for s in segs])
    np.savez_compressed(OUT, **out)
    print(f"wrote {OUT} ({OUT.stat().st_size} bytes)")


if __name__ == "__main__":
    main()
