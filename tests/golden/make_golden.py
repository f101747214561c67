"""Writes the committed fixtures of tests/golden/ (run in a container that
holds the reference checkout at /root/reference and a built oracle/_ref):

* mesh_small.npz -- the reference's assets/mesh_small.json (DFG channel,
  324 P2 elements) as arrays: float32 nodes, elements, Dirichlet node lists
  with their expressions evaluated.
* kd_golden.npz  -- query points and the outputs of the reference's OWN
  KDTree::findElement (gpu_cpu_shared/kd_tree.cuh, compiled into
  oracle/_ref/libref_kdtree.so) on mesh_small and on a 24x24 cavity:
  element index, xi, eta, nearest node.  The trees themselves are the
  oracle's restatement of cpp/kd_tree_builder.cpp (that file cannot be
  compiled here: its header chain needs cuda.h).

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from navier_stokes_fem_amd import mesh  # noqa: E402
from oracle import oracle as O  # noqa: E402

REF_MESH = "/root/reference/assets/mesh_small.json"


def query_points(m, rng, n_random):
    """Random points over an enlarged bbox, every node, node +/- tiny
    offsets (ties on edges), edge midpoints and far-away points."""
    xy = m.nodes.astype(np.float64)
    lo, hi = xy.min(0), xy.max(0)
    span = hi - lo
    pts = [rng.uniform(lo - 0.1 * span, hi + 0.1 * span, size=(n_random, 2))]
    pts.append(xy)
    eps = 1e-6 * span
    pts.append(xy + eps)
    pts.append(xy - eps)
    e = m.elements
    pts.append(0.5 * (xy[e[:, 0]] + xy[e[:, 1]]))
    pts.append(lo - span * rng.uniform(0.5, 2.0, size=(32, 2)))
    pts.append(hi + span * rng.uniform(0.5, 2.0, size=(32, 2)))
    return np.concatenate(pts).astype(np.float32)


def main():
    O.build(force=True)
    if O.ref_lib() is None:
        raise SystemExit("oracle/_ref/libref_kdtree.so is missing (needs /root/reference)")
    ms = mesh.load_json(REF_MESH)
    np.savez_compressed(os.path.join(HERE, "mesh_small.npz"), nodes=ms.nodes,
                        elements=ms.elements, pressure_count=ms.pressure_count,
                        vel_bc_nodes=ms.vel_bc_nodes, vel_bc_u=ms.vel_bc_u,
                        vel_bc_v=ms.vel_bc_v, p_bc_nodes=ms.p_bc_nodes,
                        p_bc_values=ms.p_bc_values)
    rng = np.random.default_rng(20240531)
    out = {}
    for tag, m in (("small", ms), ("cav24", mesh.cavity(24))):
        t = O.kd_build(m.xy(), m.elems())
        pts = query_points(m, rng, 4000)
        e, xi, eta, nn = O.ref_kd_find(t, m.xy(), m.elems(), pts)
        out[f"{tag}_points"] = pts
        out[f"{tag}_element"] = e
        out[f"{tag}_xi"] = xi
        out[f"{tag}_eta"] = eta
        out[f"{tag}_nearest"] = nn
        out[f"{tag}_tree_nodes"] = t.nodes
        out[f"{tag}_tree_leaf"] = t.leaf_indices
        out[f"{tag}_tree_bbox"] = t.bbox
        print(tag, "points", len(pts), "found", int((e >= 0).sum()), "tree nodes",
              len(t.nodes) // 2)
    np.savez_compressed(os.path.join(HERE, "kd_golden.npz"), **out)


if __name__ == "__main__":
    main()
