"""Pins of the oracle's KD-tree build and point location (CPU only).

* point location is compared bit for bit with golden vectors written by the
  reference's own KDTree::findElement (gpu_cpu_shared/kd_tree.cuh:40-169,
  compiled from the reference sources into oracle/_ref), and, where that
  library is present, with it directly on fresh points;
* the build (cpp/kd_tree_builder.cpp:81-148) is checked through structural
  facts of the reference's algorithm (midpoint splits on alternating axes,
  left child = next node, leaf cut-off, depth limit, bbox with the
  BBox2D::expand defect) and through the property the traversal relies on:
  the leaf reached by a point's descent lists a triangle containing it;
* point-in-triangle and the located elements are checked against exact
  geometry (brute force over all triangles in FP64).
"""
import math
import os

import numpy as np
import pytest

from conftest import GOLDEN


def _mesh(tag, mesh_small):
    from navier_stokes_fem_amd import mesh

    return mesh_small if tag == "small" else mesh.cavity(24)


@pytest.mark.parametrize("tag", ["small", "cav24"])
def test_locate_matches_reference_golden(oracle, mesh_small, tag):
    z = np.load(os.path.join(GOLDEN, "kd_golden.npz"))
    m = _mesh(tag, mesh_small)
    t = oracle.kd_build(m.xy(), m.elems())
    # the golden locate vectors were produced on exactly this tree
    np.testing.assert_array_equal(t.nodes, z[f"{tag}_tree_nodes"])
    np.testing.assert_array_equal(t.leaf_indices, z[f"{tag}_tree_leaf"])
    np.testing.assert_array_equal(t.bbox, z[f"{tag}_tree_bbox"])
    e, xi, eta, nn = oracle.kd_find(t, m.xy(), m.elems(), z[f"{tag}_points"])
    np.testing.assert_array_equal(e, z[f"{tag}_element"])
    np.testing.assert_array_equal(xi.view(np.uint32), z[f"{tag}_xi"].view(np.uint32))
    np.testing.assert_array_equal(eta.view(np.uint32), z[f"{tag}_eta"].view(np.uint32))
    miss = e < 0
    np.testing.assert_array_equal(nn[miss], z[f"{tag}_nearest"][miss])


def test_locate_matches_reference_build_live(oracle):
    """Fresh points against the reference's traversal (oracle/_ref)."""
    from navier_stokes_fem_amd import mesh

    if oracle.ref_lib() is None:
        pytest.skip("oracle/_ref not built (the reference checkout is absent)")
    m = mesh.channel(nx=48, ny=12, seed=7)
    t = oracle.kd_build(m.xy(), m.elems())
    rng = np.random.default_rng(8)
    pts = rng.uniform([-0.1, -0.05], [2.3, 0.46], size=(20000, 2)).astype(np.float32)
    a = oracle.kd_find(t, m.xy(), m.elems(), pts)
    b = oracle.ref_kd_find(t, m.xy(), m.elems(), pts)
    np.testing.assert_array_equal(a[0], b[0])
    np.testing.assert_array_equal(a[1].view(np.uint32), b[1].view(np.uint32))
    np.testing.assert_array_equal(a[2].view(np.uint32), b[2].view(np.uint32))
    miss = a[0] < 0
    np.testing.assert_array_equal(a[3][miss], b[3][miss])


def _walk(nodes):
    """Yield (index, depth, kind, payload) for every node, root first."""
    out = []
    stack = [(0, 0)]
    while stack:
        i, d = stack.pop()
        w0, w1 = int(nodes[2 * i]), int(nodes[2 * i + 1])
        kind = w0 & 3
        out.append((i, d, kind, w0 >> 2, w1))
        if kind != 2:
            stack.append((w0 >> 2, d + 1))
            stack.append((i + 1, d + 1))
    return out


@pytest.mark.parametrize("tag", ["small", "cav24", "channel"])
def test_build_structure(oracle, mesh_small, tag):
    from navier_stokes_fem_amd import mesh

    m = mesh.channel(nx=64, ny=16, seed=3) if tag == "channel" else _mesh(tag, mesh_small)
    t = oracle.kd_build(m.xy(), m.elems())
    E = m.element_count
    # depth limit: min(29, round(8 + 1.3f*log(E))) (cpp/kd_tree_builder.cpp:87)
    want_depth = min(29, int(np.round(np.float32(8) + np.float32(1.3) * np.float32(math.log(E)))))
    assert t.max_depth == want_depth
    # bbox: min over nodes, max = LAST node (BBox2D::expand defect)
    xy = m.nodes
    np.testing.assert_array_equal(t.bbox[:2], xy.min(0))
    np.testing.assert_array_equal(t.bbox[2:], xy[-1])
    nodes = t.nodes.reshape(-1)
    walk = _walk(nodes)
    assert len(walk) == len(nodes) // 2  # every node reached exactly once
    assert sorted(i for i, *_ in walk) == list(range(len(nodes) // 2))
    # boxes along the descent: split = midpoint of the current box on the
    # axis of the level (x at the root)
    boxes = {0: (t.bbox[:2].astype(np.float32).copy(), t.bbox[2:].astype(np.float32).copy())}
    covered = np.zeros(E, bool)
    for i, d, kind, hi30, w1 in sorted(walk, key=lambda w: w[0]):
        lo, hi = boxes[i]
        if kind == 2:
            cnt, off = hi30, w1
            assert cnt <= 16 or d >= t.max_depth
            covered[t.leaf_indices[off:off + cnt]] = True
            continue
        assert kind == d % 2
        split = np.uint32(w1).view(np.float32)
        assert split == np.float32((lo[kind] + hi[kind]) * np.float32(0.5))
        right = hi30
        assert right > i + 1
        lhi = hi.copy(); lhi[kind] = split
        rlo = lo.copy(); rlo[kind] = split
        boxes[i + 1] = (lo, lhi)
        boxes[right] = (rlo, hi)
    assert covered.all()


def _contains(m, e, p, eps=1e-9):
    """Exact-ish FP64 containment of p in triangle e (vertices 0..2)."""
    a, b, c = (m.nodes[m.elements[e, k]].astype(np.float64) for k in range(3))
    d = (b[0] - a[0]) * (c[1] - a[1]) - (c[0] - a[0]) * (b[1] - a[1])
    l1 = ((p[0] - a[0]) * (c[1] - a[1]) - (c[0] - a[0]) * (p[1] - a[1])) / d
    l2 = ((b[0] - a[0]) * (p[1] - a[1]) - (p[0] - a[0]) * (b[1] - a[1])) / d
    return l1 >= -eps and l2 >= -eps and l1 + l2 <= 1 + eps, l1, l2


def _descend(nodes, p):
    i = 0
    while (nodes[2 * i] & 3) != 2:
        ax = nodes[2 * i] & 3
        split = np.uint32(nodes[2 * i + 1]).view(np.float32)
        i = i + 1 if p[ax] <= split else int(nodes[2 * i] >> 2)
    return int(nodes[2 * i] >> 2), int(nodes[2 * i + 1])


def test_descent_leaf_holds_containing_triangle(oracle):
    from navier_stokes_fem_amd import mesh

    m = mesh.channel(nx=40, ny=10, seed=5)
    t = oracle.kd_build(m.xy(), m.elems())
    rng = np.random.default_rng(6)
    # points at triangle centroids: each lies strictly inside its triangle
    es = rng.choice(m.element_count, 400, replace=False)
    cent = m.nodes[m.elements[es, :3]].astype(np.float64).mean(1).astype(np.float32)
    nodes = t.nodes.reshape(-1)
    for e, p in zip(es, cent):
        cnt, off = _descend(nodes, p)
        assert e in set(t.leaf_indices[off:off + cnt].tolist())


def test_located_elements_contain_points(oracle):
    from navier_stokes_fem_amd import mesh

    m = mesh.channel(nx=40, ny=10, seed=5)
    t = oracle.kd_build(m.xy(), m.elems())
    rng = np.random.default_rng(9)
    pts = rng.uniform([-0.2, -0.1], [2.4, 0.5], size=(3000, 2)).astype(np.float32)
    e, xi, eta, nn = oracle.kd_find(t, m.xy(), m.elems(), pts)
    xy = m.nodes.astype(np.float64)
    for k in range(len(pts)):
        p = pts[k].astype(np.float64)
        if e[k] >= 0:
            ok, l1, l2 = _contains(m, e[k], p, eps=1e-5)
            assert ok
            assert abs(l1 - xi[k]) < 1e-4 and abs(l2 - eta[k]) < 1e-4
        else:
            # no triangle holds the point, and the nearest node is nearest
            d = ((xy - p) ** 2).sum(1)
            assert d[nn[k]] <= d.min() * (1 + 1e-5) + 1e-12
    # points strictly inside the domain are always found
    inside = (pts[:, 0] > 0.01) & (pts[:, 0] < 2.19) & (pts[:, 1] > 0.01) & (pts[:, 1] < 0.40)
    hole = ((pts[:, 0] - 0.2) ** 2 + (pts[:, 1] - 0.2) ** 2) < 0.07 ** 2
    assert np.all(e[inside & ~hole] >= 0)


def test_point_in_triangle_ccw_only(oracle):
    ok, xi, eta = oracle.point_in_triangle((0.25, 0.5), (0, 0), (1, 0), (0, 1))
    assert ok and xi == 0.25 and eta == 0.5
    # the reference divides by |D| but compares with the signed D
    # (gpu_cpu_shared/kd_tree_common.cuh:18-61): clockwise triangles reject
    ok, _, _ = oracle.point_in_triangle((0.25, 0.5), (0, 0), (0, 1), (1, 0))
    assert not ok
    # vertices and edges are inside (closed triangle)
    for p in [(0, 0), (1, 0), (0, 1), (0.5, 0.5), (0.5, 0)]:
        assert oracle.point_in_triangle(p, (0, 0), (1, 0), (0, 1))[0]
    assert not oracle.point_in_triangle((0.6, 0.6), (0, 0), (1, 0), (0, 1))[0]


def test_float_offset_is_exact_below_2_24(oracle):
    """KDNode::getTriangleOffset returns float; offsets below 2^24 are exact,
    so both decodings agree on meshes of this size."""
    from navier_stokes_fem_amd import mesh

    m = mesh.cavity(20)
    t = oracle.kd_build(m.xy(), m.elems())
    pts = np.random.default_rng(1).uniform(-0.1, 1.1, size=(5000, 2)).astype(np.float32)
    a = oracle.kd_find(t, m.xy(), m.elems(), pts, float_offset=False)
    b = oracle.kd_find(t, m.xy(), m.elems(), pts, float_offset=True)
    for u, v in zip(a, b):
        np.testing.assert_array_equal(u, v)


def _shift_leaves(t, shift):
    """The same tree with its leaf index array behind `shift` entries of
    padding (every leaf offset grows by `shift`)."""
    w = t.nodes.reshape(-1, 2).copy()
    leaf = (w[:, 0] & 3) == 2
    w[leaf, 1] += np.uint32(shift)
    pad = np.zeros(shift, np.int32)
    return type(t)(w.reshape(-1), np.concatenate([pad, t.leaf_indices]), t.bbox, t.max_depth)


def test_float_offset_decoding_above_2_24(oracle):
    """Above 2^24 (and below 2^25) the float decoding rounds odd offsets to
    an even neighbour.  On a structured cavity every leaf holds an even
    number of triangles, so every offset is even: with the leaves moved by
    2^24 + 2 the reference's decoding stays exact and no point moves (the
    2048^2 cavity's case: leaves of 12 or 16 triangles in an array of
    33.5M < 2^25 entries), while a move by 2^24 + 1 makes every offset odd
    and the windows one entry off place points elsewhere."""
    from navier_stokes_fem_amd import mesh

    m = mesh.cavity(20)
    t = oracle.kd_build(m.xy(), m.elems())
    w = t.nodes.reshape(-1, 2)
    leaf = (w[:, 0] & 3) == 2
    assert ((w[leaf, 0] >> 2) % 2 == 0).all() and (w[leaf, 1] % 2 == 0).all()
    pts = np.random.default_rng(2).uniform(0.0, 1.0, size=(4000, 2)).astype(np.float32)
    base = oracle.kd_find(t, m.xy(), m.elems(), pts)
    moved = {}
    for shift in (2 ** 24 + 2, 2 ** 24 + 1):
        ts = _shift_leaves(t, shift)
        exact = oracle.kd_find(ts, m.xy(), m.elems(), pts)
        for u, v in zip(exact, base):
            np.testing.assert_array_equal(u, v)  # integer decoding: the same tree
        rounded = oracle.kd_find(ts, m.xy(), m.elems(), pts, float_offset=True)
        moved[shift] = int((rounded[0] != base[0]).sum())
    print(f"points placed elsewhere by the float decoding: {moved}")
    assert moved[2 ** 24 + 2] == 0
    assert moved[2 ** 24 + 1] > 0
