"""The reference's winding-number examples over the batched GPU calls.

``fix_orientation`` (modules/winding/examples/fix_orientation.cpp): for every facet of an input
mesh, probe the reference mesh's winding number at barycentre +- epsilon x unit normal; a facet
whose criterion (omega(+) - omega(-)) / (4 pi) exceeds the threshold is flipped (its first two
vertices swapped).  Defaults epsilon 1e-2, threshold 0.8 (:40-41).

``sample_points_in_mesh`` (modules/winding/examples/sample_points_in_mesh.cpp): uniform samples in
the mesh's bounding box, kept when inside, written as an .xyz file.  Default 10,000 samples
(:36).

Command line (mirrors the reference programs' arguments):
  python -m lagrange_amd.examples fix_orientation INPUT REFERENCE [OUTPUT] [-e EPS] [-t THRESH]
                                   [--criterion-json PATH]
  python -m lagrange_amd.examples sample_points_in_mesh INPUT [OUTPUT.xyz] [-n N] [--seed S]
"""
from __future__ import annotations

import argparse
import json
import sys

import numpy as np

from . import io as mio
from .winding import FastWindingNumber


def fix_orientation(vertices, facets, reference, epsilon: float = 1e-2, threshold: float = 0.8):
    """(fixed facets, criterion per facet, counts) for an input triangle mesh; ``reference`` is a
    FastWindingNumber or a (V, F) mesh."""
    engine = reference if isinstance(reference, FastWindingNumber) else FastWindingNumber(reference)
    F = np.array(facets, dtype=np.int32, copy=True).reshape(-1, 3)
    crit = engine.orientation_criterion(vertices, F, epsilon)
    flip = crit > threshold
    F[flip, 0], F[flip, 1] = F[flip, 1].copy(), F[flip, 0].copy()
    counts = {"pos": int(flip.sum()), "neg": int((crit < threshold).sum()), "tot": int(crit.size)}
    return F, crit, counts


def sample_points_in_mesh(mesh, num_samples: int = 10000, seed: int = 0):
    """Samples uniform in the mesh's bounding box that are inside the mesh, in sample order."""
    engine = mesh if isinstance(mesh, FastWindingNumber) else FastWindingNumber(mesh)
    return engine.sample_inside(num_samples, seed)


def _load_triangles(path):
    m = mio.load_mesh(path, triangulate=True)
    return m.vertices.astype(np.float32), m.facets


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="python -m lagrange_amd.examples")
    sub = ap.add_subparsers(dest="cmd", required=True)
    fo = sub.add_parser("fix_orientation")
    fo.add_argument("input")
    fo.add_argument("reference")
    fo.add_argument("output", nargs="?")
    fo.add_argument("-e", "--epsilon", type=float, default=1e-2)
    fo.add_argument("-t", "--threshold", type=float, default=0.8)
    fo.add_argument("--criterion-json")
    sp = sub.add_parser("sample_points_in_mesh")
    sp.add_argument("input")
    sp.add_argument("output", nargs="?", default="output.xyz")
    sp.add_argument("-n", "--num-samples", type=int, default=10000)
    sp.add_argument("--seed", type=int, default=0)
    a = ap.parse_args(argv)
    if a.cmd == "fix_orientation":
        V, F = _load_triangles(a.input)
        ref = _load_triangles(a.reference)
        F2, crit, counts = fix_orientation(V, F, ref, a.epsilon, a.threshold)
        print(f"pos {counts['pos']} neg {counts['neg']} tot {counts['tot']}")
        if a.criterion_json:
            with open(a.criterion_json, "w") as f:
                json.dump({"criterions": crit.tolist(), "epsilon": a.epsilon, "threshold": a.threshold}, f)
        if a.output:
            mio.save_mesh(a.output, V, F2)
        return 0
    V, F = _load_triangles(a.input)
    if not a.output.endswith(".xyz"):
        print(f"Output file extension should be .xyz. '{a.output}' was given.", file=sys.stderr)
    pts = sample_points_in_mesh((V, F), a.num_samples, a.seed)
    np.savetxt(a.output, pts, fmt="%.9g")
    print(f"{pts.shape[0]} of {a.num_samples} samples inside")
    return 0


if __name__ == "__main__":
    sys.exit(main())
