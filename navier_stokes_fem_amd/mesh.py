"""P2-P1 triangular meshes in the reference's layout.

The reference reads a JSON mesh (cpp/grid.cpp:9-77) whose element rows are
``[v0, v1, v2, m12, m02, m01]`` -- three CCW vertices followed by the
midpoints opposite each vertex, the order P2ShapeEval expects
(gpu/advection.cu:5-12, include/assembly.h:201-235).  Vertices come first
and are the pressure nodes ``0..P-1`` (include/grid.h:118-121).  Node
coordinates are ``float`` (gpu_cpu_shared/misc_common.cuh:8), so every mesh
here stores ``float32`` coordinates; FP64 arithmetic on the mesh starts from
those values.

Dirichlet data are carried as explicit per-node values (the reference
evaluates expression strings per node, include/assembly.h:503-524); lists are
kept in boundary order, which is the order the reference patches right-hand
sides in.  :func:`load_json` reads the reference's own file format, including
its boundary-condition expressions (cpp/expression.cpp grammar: ``+ - * / ^``,
unary minus, ``sin``, ``cos``, ``pi``, variables ``x`` and ``y``).
"""
from __future__ import annotations

import json
import math
import re
from dataclasses import dataclass, field

import numpy as np

__all__ = ["Mesh", "cavity", "channel", "load_json", "evaluate_expression"]


@dataclass
class Mesh:
    nodes: np.ndarray              # float32 [V, 2]
    elements: np.ndarray           # int32 [E, 6]
    pressure_count: int            # P: nodes 0..P-1 are the P1 vertices
    vel_bc_nodes: np.ndarray       # int32 [B]
    vel_bc_u: np.ndarray           # float64 [B]
    vel_bc_v: np.ndarray           # float64 [B]
    p_bc_nodes: np.ndarray         # int32 [Bp]
    p_bc_values: np.ndarray        # float64 [Bp]
    name: str = "mesh"
    info: dict = field(default_factory=dict)
    # int32 [P, 2]: logical (i, j) of each pressure node on the generator's
    # structured grid (cavity, channel), None for other meshes
    grid: np.ndarray = None

    @property
    def node_count(self) -> int:
        return int(self.nodes.shape[0])

    @property
    def element_count(self) -> int:
        return int(self.elements.shape[0])

    def xy(self) -> np.ndarray:
        return np.ascontiguousarray(self.nodes, dtype=np.float32).reshape(-1)

    def elems(self) -> np.ndarray:
        return np.ascontiguousarray(self.elements, dtype=np.int32).reshape(-1)

    def check(self) -> None:
        """Layout invariants the kernels rely on (raise on violation)."""
        e = self.elements
        V, P = self.node_count, self.pressure_count
        if e.ndim != 2 or e.shape[1] != 6:
            raise ValueError("elements must be [E, 6]")
        if e.min() < 0 or e.max() >= V:
            raise ValueError("element node index out of range")
        if e[:, :3].max() >= P or e[:, 3:].min() < P:
            raise ValueError("vertices must be nodes 0..P-1, midpoints P..V-1")
        xy = self.nodes.astype(np.float64)
        a, b, c = xy[e[:, 0]], xy[e[:, 1]], xy[e[:, 2]]
        det = (b[:, 0] - a[:, 0]) * (c[:, 1] - a[:, 1]) - (b[:, 1] - a[:, 1]) * (c[:, 0] - a[:, 0])
        if not np.all(det > 0):
            raise ValueError("all triangles must be counter-clockwise")
        if len(self.vel_bc_u) != len(self.vel_bc_nodes) or len(self.vel_bc_v) != len(self.vel_bc_nodes):
            raise ValueError("velocity Dirichlet values must match the node list")
        if len(self.p_bc_values) != len(self.p_bc_nodes):
            raise ValueError("pressure Dirichlet values must match the node list")


def _p2_from_p1(xy_v: np.ndarray, tris: np.ndarray):
    """Add one midpoint node per unique edge.  Midpoints are numbered after
    the vertices in ascending (min vertex, max vertex) order, which keeps a
    row-major vertex numbering local for the midpoints too."""
    P = xy_v.shape[0]
    t = tris.astype(np.int64)
    # edges opposite v0, v1, v2: (v1,v2), (v0,v2), (v0,v1)
    pairs = np.stack([t[:, [1, 2]], t[:, [0, 2]], t[:, [0, 1]]], axis=1)  # [E,3,2]
    lo = pairs.min(axis=2)
    hi = pairs.max(axis=2)
    key = lo * P + hi
    uniq, inv, counts = np.unique(key.reshape(-1), return_inverse=True, return_counts=True)
    mid_idx = (P + inv).reshape(-1, 3)
    ulo = uniq // P
    uhi = uniq % P
    mids = ((xy_v[ulo].astype(np.float64) + xy_v[uhi].astype(np.float64)) * 0.5).astype(np.float32)
    nodes = np.concatenate([xy_v.astype(np.float32), mids], axis=0)
    elements = np.concatenate([tris.astype(np.int64), mid_idx], axis=1).astype(np.int32)
    edges = np.stack([ulo, uhi], axis=1)
    return nodes, elements, edges, counts


def cavity(n: int, lid_speed: float = 1.0) -> Mesh:
    """Lid-driven cavity on the unit square: ``n x n`` squares, each split
    along the (i,j)-(i+1,j+1) diagonal into two CCW triangles.  Lid
    ``u = (lid_speed, 0)`` on ``y = 1`` excluding the two top corners, no-slip
    on every other boundary node; the pure-Neumann pressure is pinned with
    ``p = 0`` at node 0 (the reference's pDirichlet mechanism,
    include/assembly.h:583-603)."""
    if n < 1:
        raise ValueError("n must be >= 1")
    ii, jj = np.meshgrid(np.arange(n + 1), np.arange(n + 1), indexing="xy")
    xy_v = np.stack([ii.reshape(-1) / n, jj.reshape(-1) / n], axis=1).astype(np.float32)
    sq_i, sq_j = np.meshgrid(np.arange(n), np.arange(n), indexing="xy")
    sq_i = sq_i.reshape(-1)
    sq_j = sq_j.reshape(-1)
    v00 = sq_j * (n + 1) + sq_i
    v10 = v00 + 1
    v01 = v00 + (n + 1)
    v11 = v01 + 1
    t1 = np.stack([v00, v10, v11], axis=1)
    t2 = np.stack([v00, v11, v01], axis=1)
    tris = np.stack([t1, t2], axis=1).reshape(-1, 3)
    nodes, elements, edges, counts = _p2_from_p1(xy_v, tris)
    P = xy_v.shape[0]
    vi, vj = ii.reshape(-1), jj.reshape(-1)
    on_b = (vi == 0) | (vi == n) | (vj == 0) | (vj == n)
    top_v = (vj == n) & (vi > 0) & (vi < n)
    bedge = counts == 1
    e_top = bedge & (vj[edges[:, 0]] == n) & (vj[edges[:, 1]] == n)
    lid = np.concatenate([np.nonzero(top_v)[0], P + np.nonzero(e_top)[0]])
    walls = np.concatenate([np.nonzero(on_b & ~top_v)[0], P + np.nonzero(bedge & ~e_top)[0]])
    lid = np.sort(lid)
    walls = np.sort(walls)
    vb = np.concatenate([lid, walls]).astype(np.int32)
    gu = np.concatenate([np.full(len(lid), float(lid_speed)), np.zeros(len(walls))])
    gv = np.zeros(len(vb))
    m = Mesh(nodes=nodes, elements=elements, pressure_count=P,
             vel_bc_nodes=vb, vel_bc_u=gu, vel_bc_v=gv,
             p_bc_nodes=np.array([0], dtype=np.int32), p_bc_values=np.zeros(1),
             name=f"cavity{n}", info={"n": n, "lid_speed": float(lid_speed)},
             grid=np.stack([vi, vj], axis=1).astype(np.int32))
    return m


def cavity_prolongation(n: int, ratio: int):
    """P1 interpolation from the pressure nodes of the (n/ratio)^2 cavity to
    those of the n^2 cavity (both numbered row-major, squares split along the
    same diagonal, so every coarse triangle is a union of fine ones): a fine
    node in the coarse square (I, J) at local position (a, b) takes the
    barycentric weights of its coarse triangle, (1-a, a-b, b) on the corners
    (0,0), (1,0), (1,1) when a >= b, else (1-b, a, b-a) on (0,0), (1,1),
    (0,1).  Linear functions are reproduced exactly.  Returns (row_start,
    col, val, coarse_count) for ``Simulation.set_coarse_space``."""
    if ratio < 2 or n % ratio:
        raise ValueError("ratio must be >= 2 and divide n")
    nc = n // ratio
    f = np.arange((n + 1) ** 2)
    j, i = np.divmod(f, n + 1)
    I = np.minimum(i // ratio, nc - 1)
    J = np.minimum(j // ratio, nc - 1)
    a = (i - I * ratio) / ratio
    b = (j - J * ratio) / ratio
    c00 = J * (nc + 1) + I
    c10, c01 = c00 + 1, c00 + nc + 1
    c11 = c01 + 1
    lower = a >= b
    cols = np.stack([c00, np.where(lower, c10, c11), np.where(lower, c11, c01)], axis=1)
    w = np.stack([np.where(lower, 1 - a, 1 - b), np.where(lower, a - b, a),
                  np.where(lower, b, b - a)], axis=1)
    rows = np.repeat(f, 3)
    cols = cols.reshape(-1)
    w = w.reshape(-1)
    keep = w > 0
    rows, cols, w = rows[keep], cols[keep], w[keep]
    order = np.lexsort((cols, rows))
    rows, cols, w = rows[order], cols[order], w[order]
    rs = np.zeros(len(f) + 1, np.int64)
    np.cumsum(np.bincount(rows, minlength=len(f)), out=rs[1:])
    return rs.astype(np.int32), cols.astype(np.int32), w.astype(np.float64), (nc + 1) ** 2


def _coarse_grid_level(grid: np.ndarray, extent, dirichlet: np.ndarray, ratio: int):
    """The coarse level of a structured grid (node positions `grid`, logical
    extents `extent`, Dirichlet flags): every coarse position (I*ratio,
    J*ratio) that carries a node, in raster order, with its coarse logical
    position and Dirichlet flag; plus which of them are coarse unknowns."""
    gi = grid[:, 0].astype(np.int64)
    gj = grid[:, 1].astype(np.int64)
    nx, ny = extent
    lut = -np.ones((ny + 1) * (nx + 1), np.int64)
    lut[gj * (nx + 1) + gi] = np.arange(len(gi))
    mx, my = nx // ratio, ny // ratio
    cI, cJ = np.meshgrid(np.arange(mx + 1), np.arange(my + 1), indexing="xy")
    cf = lut[(cJ.reshape(-1) * ratio) * (nx + 1) + cI.reshape(-1) * ratio]
    exists = cf >= 0
    valid = exists.copy()
    valid[exists] = ~dirichlet[cf[exists]]
    return cI.reshape(-1), cJ.reshape(-1), cf, exists, valid


def _grid_interpolation(grid, extent, dirichlet, ratio):
    gi = grid[:, 0].astype(np.int64)
    gj = grid[:, 1].astype(np.int64)
    P = len(gi)
    nx, ny = extent
    if nx % ratio or ny % ratio:
        raise ValueError("ratio must divide the grid's extents")
    cI, cJ, cf, exists, valid = _coarse_grid_level(grid, extent, dirichlet, ratio)
    mx, my = nx // ratio, ny // ratio
    cid = -np.ones(len(cf), np.int64)
    cid[valid] = np.arange(int(valid.sum()))
    I = np.minimum(gi // ratio, mx - 1)
    J = np.minimum(gj // ratio, my - 1)
    a = (gi - I * ratio) / ratio
    b = (gj - J * ratio) / ratio
    c00 = J * (mx + 1) + I
    c10, c01 = c00 + 1, c00 + mx + 1
    c11 = c01 + 1
    lower = a >= b
    corners = np.stack([c00, np.where(lower, c10, c11), np.where(lower, c11, c01)], axis=1)
    w = np.stack([np.where(lower, 1 - a, 1 - b), np.where(lower, a - b, a),
                  np.where(lower, b, b - a)], axis=1)
    w = w * exists[corners]
    tot = w.sum(axis=1, keepdims=True)
    w = np.divide(w, tot, out=np.zeros_like(w), where=tot > 0)
    w = w * valid[corners]
    w[dirichlet] = 0.0
    rows = np.repeat(np.arange(P), 3)
    cols = cid[corners].reshape(-1)
    w = w.reshape(-1)
    keep = w > 0
    rows, cols, w = rows[keep], cols[keep], w[keep]
    order = np.lexsort((cols, rows))
    rows, cols, w = rows[order], cols[order], w[order]
    rs = np.zeros(P + 1, np.int64)
    np.cumsum(np.bincount(rows, minlength=P), out=rs[1:])
    return rs.astype(np.int32), cols.astype(np.int32), w, int(valid.sum())


def grid_hierarchy(mesh: Mesh, levels: int = -1, min_size: int = 16):
    """Prolongations of a multigrid hierarchy on a mesh built on a
    structured grid (``mesh.grid``: cavity, channel): the first is
    ``grid_prolongation(mesh, 2)`` (for ``Simulation.set_coarse_space``),
    each further one interpolates from the next coarser grid onto the
    previous level's coarse unknowns (for ``Simulation.add_coarse_level``)
    with the same rules: holes rescale, positions of pressure Dirichlet
    nodes stay zero.  ``levels`` further levels (< 0: halve while both
    extents stay even and the smaller one above ``min_size``)."""
    g = mesh.grid
    if g is None:
        raise ValueError("the mesh carries no structured grid")
    grid = np.asarray(g, np.int64)
    extent = (int(grid[:, 0].max()), int(grid[:, 1].max()))
    dirichlet = np.zeros(mesh.pressure_count, bool)
    dirichlet[np.asarray(mesh.p_bc_nodes, np.int64)] = True
    out = [_grid_interpolation(grid, extent, dirichlet, 2)]
    count = 0
    while levels < 0 or count < levels:
        # the level just built, as a grid: its existing coarse positions
        cI, cJ, cf, exists, valid = _coarse_grid_level(grid, extent, dirichlet, 2)
        extent = (extent[0] // 2, extent[1] // 2)
        grid = np.stack([cI[exists], cJ[exists]], axis=1)
        dirichlet = ~valid[exists]
        if extent[0] % 2 or extent[1] % 2:
            if levels >= 0:
                raise ValueError("cannot halve the coarse grid again")
            break
        if levels < 0 and min(extent) // 2 < min_size:
            break
        rs, col, val, nc = _grid_interpolation(grid, extent, dirichlet, 2)
        # rows of the level's coarse unknowns only (raster order, as their ids)
        keep = np.nonzero(~dirichlet)[0]
        cnt = np.diff(rs)[keep]
        idx = np.concatenate([np.arange(rs[k], rs[k + 1]) for k in keep]) if len(keep) else \
            np.zeros(0, np.int64)
        nrs = np.zeros(len(keep) + 1, np.int64)
        np.cumsum(cnt, out=nrs[1:])
        out.append((nrs.astype(np.int32), col[idx], val[idx], nc))
        count += 1
    return out


def _grid_levels(grid, extent, dirichlet, levels, min_size):
    """Prolongations below a structured level (node positions `grid`, logical
    `extent`, Dirichlet flags), each from the next coarser grid onto the
    level's non-Dirichlet nodes in their order (rows), as grid_hierarchy's
    deeper levels."""
    out = []
    count = 0
    while levels < 0 or count < levels:
        if extent[0] % 2 or extent[1] % 2:
            if levels >= 0:
                raise ValueError("cannot halve the coarse grid again")
            break
        if levels < 0 and min(extent) // 2 < min_size:
            break
        rs, col, val, nc = _grid_interpolation(grid, extent, dirichlet, 2)
        keep = np.nonzero(~dirichlet)[0]
        cnt = np.diff(rs)[keep]
        idx = np.concatenate([np.arange(rs[k], rs[k + 1]) for k in keep]) if len(keep) else \
            np.zeros(0, np.int64)
        nrs = np.zeros(len(keep) + 1, np.int64)
        np.cumsum(cnt, out=nrs[1:])
        out.append((nrs.astype(np.int32), col[idx], val[idx], nc))
        cI, cJ, cf, exists, valid = _coarse_grid_level(grid, extent, dirichlet, 2)
        extent = (extent[0] // 2, extent[1] // 2)
        grid = np.stack([cI[exists], cJ[exists]], axis=1)
        dirichlet = ~valid[exists]
        count += 1
    return out


def velocity_hierarchy(mesh: Mesh, levels: int = -1, min_size: int = 32):
    """Prolongations of a multigrid hierarchy below the velocity (P2) space
    of a mesh built on a structured grid, for
    ``Simulation.set_velocity_coarse_space`` (the first) and
    ``add_velocity_coarse_level`` (the rest).  The first maps the P1 space
    of the mesh's vertices without velocity Dirichlet values onto the P2
    nodes (a vertex keeps its value, an edge midpoint takes half of each
    endpoint's; Dirichlet nodes get empty rows and Dirichlet vertices no
    column, so every coarse function vanishes on the velocity boundary);
    the others halve the vertex grid as grid_hierarchy does (holes
    rescale, Dirichlet positions stay zero), ``levels`` of them (< 0: while
    both extents stay even and the smaller one halves to at least
    ``min_size``: a cavity ends on its 32^2 grid, whose 31^2 free vertices
    the library solves by a dense inverse).  Rows are in the mesh's node
    order."""
    g = mesh.grid
    if g is None:
        raise ValueError("the mesh carries no structured grid")
    V, P = mesh.node_count, mesh.pressure_count
    vdir = np.zeros(V, bool)
    vdir[np.asarray(mesh.vel_bc_nodes, np.int64)] = True
    free = ~vdir[:P]
    vid = -np.ones(P, np.int64)
    vid[free] = np.arange(int(free.sum()))
    el = np.asarray(mesh.elements, np.int64)
    mids = el[:, 3:6].reshape(-1)
    ends = np.stack([el[:, [1, 2]], el[:, [0, 2]], el[:, [0, 1]]], axis=1).reshape(-1, 2)
    um, first = np.unique(mids, return_index=True)
    e = ends[first]
    rows = [np.nonzero(free)[0]]
    cols = [vid[free]]
    vals = [np.ones(int(free.sum()))]
    for k in (0, 1):
        ok = ~vdir[um] & free[e[:, k]]
        rows.append(um[ok])
        cols.append(vid[e[ok, k]])
        vals.append(np.full(int(ok.sum()), 0.5))
    rows, cols, vals = np.concatenate(rows), np.concatenate(cols), np.concatenate(vals)
    order = np.lexsort((cols, rows))
    rows, cols, vals = rows[order], cols[order], vals[order]
    rs = np.zeros(V + 1, np.int64)
    np.cumsum(np.bincount(rows, minlength=V), out=rs[1:])
    out = [(rs.astype(np.int32), cols.astype(np.int32), vals, int(free.sum()))]
    grid = np.asarray(g, np.int64)
    extent = (int(grid[:, 0].max()), int(grid[:, 1].max()))
    return out + _grid_levels(grid, extent, vdir[:P], levels, min_size)


def grid_prolongation(mesh: Mesh, ratio: int):
    """Interpolation from a coarse grid onto the pressure nodes of a mesh
    built on a structured grid (``mesh.grid``: cavity, channel), for
    ``Simulation.set_coarse_space``.  Coarse nodes are the grid positions
    (I*ratio, J*ratio) that carry a pressure node without a pressure
    Dirichlet value.  A fine node at logical position (i, j) takes the
    barycentric weights, in logical coordinates, of the coarse triangle
    around it (squares split along the same diagonal as the mesh's):

    * corners whose grid position has no node (a hole: the channel's disk)
      are dropped and the other weights rescaled to sum 1;
    * corners on the pressure Dirichlet boundary are dropped without
      rescaling, so every coarse function vanishes there;
    * pressure Dirichlet nodes get an empty row.

    Returns (row_start, col, val, coarse_count)."""
    g = mesh.grid
    if g is None:
        raise ValueError("the mesh carries no structured grid")
    if ratio < 2:
        raise ValueError("ratio must be >= 2")
    P = mesh.pressure_count
    gi = g[:, 0].astype(np.int64)
    gj = g[:, 1].astype(np.int64)
    nx, ny = int(gi.max()), int(gj.max())
    if nx % ratio or ny % ratio:
        raise ValueError("ratio must divide the grid's extents")
    lut = -np.ones((ny + 1) * (nx + 1), np.int64)
    lut[gj * (nx + 1) + gi] = np.arange(P)
    dirichlet = np.zeros(P, bool)
    dirichlet[np.asarray(mesh.p_bc_nodes, np.int64)] = True
    mx, my = nx // ratio, ny // ratio
    cI, cJ = np.meshgrid(np.arange(mx + 1), np.arange(my + 1), indexing="xy")
    cf = lut[(cJ.reshape(-1) * ratio) * (nx + 1) + cI.reshape(-1) * ratio]
    exists = cf >= 0
    valid = exists.copy()
    valid[exists] = ~dirichlet[cf[exists]]
    cid = -np.ones(len(cf), np.int64)
    cid[valid] = np.arange(int(valid.sum()))
    I = np.minimum(gi // ratio, mx - 1)
    J = np.minimum(gj // ratio, my - 1)
    a = (gi - I * ratio) / ratio
    b = (gj - J * ratio) / ratio
    c00 = J * (mx + 1) + I
    c10, c01 = c00 + 1, c00 + mx + 1
    c11 = c01 + 1
    lower = a >= b
    corners = np.stack([c00, np.where(lower, c10, c11), np.where(lower, c11, c01)], axis=1)
    w = np.stack([np.where(lower, 1 - a, 1 - b), np.where(lower, a - b, a),
                  np.where(lower, b, b - a)], axis=1)
    w = w * exists[corners]
    tot = w.sum(axis=1, keepdims=True)
    w = np.divide(w, tot, out=np.zeros_like(w), where=tot > 0)
    w = w * valid[corners]
    w[dirichlet] = 0.0
    rows = np.repeat(np.arange(P), 3)
    cols = cid[corners].reshape(-1)
    w = w.reshape(-1)
    keep = w > 0
    rows, cols, w = rows[keep], cols[keep], w[keep]
    order = np.lexsort((cols, rows))
    rows, cols, w = rows[order], cols[order], w[order]
    rs = np.zeros(P + 1, np.int64)
    np.cumsum(np.bincount(rows, minlength=P), out=rs[1:])
    return rs.astype(np.int32), cols.astype(np.int32), w, int(valid.sum())


def cavity_batch(count: int = 64, seed: int = 3, n_min: int = 16, n_max: int = 256):
    """A batch of independent lid-driven cavities of uneven size: side
    ``n ~ U{n_min..n_max}``, lid speed ``U ~ U(0.5, 2)``, viscosity
    ``nu ~ logU(1e-3, 1e-2)`` (seeded).  Returns ``[(mesh, nu), ...]``."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        n = int(rng.integers(n_min, n_max + 1))
        lid = float(rng.uniform(0.5, 2.0))
        nu = float(np.exp(rng.uniform(np.log(1e-3), np.log(1e-2))))
        m = cavity(n, lid_speed=lid)
        m.name = f"cavity{n}_lid{lid:.3f}"
        out.append((m, nu))
    return out


def _inflow(y: np.ndarray, umax: float, height: float) -> np.ndarray:
    # 4*Um*y*(H-y)/H^2, the inflow profile of assets/mesh_small.json
    return 4.0 * umax * y * (height - y) / (height * height)


def channel(nx: int = 512, ny: int = 128, seed: int = 2, jitter: float = 0.2,
            length: float = 2.2, height: float = 0.41,
            disk: tuple = (0.2, 0.2, 0.05), umax: float = 1.5) -> Mesh:
    """Channel past a cylinder: ``[0,length] x [0,height]`` on an ``nx x ny``
    grid, interior vertices jittered uniformly by ``+-jitter*h`` per axis
    (seeded), triangles whose centroid lies inside the disk removed, orphan
    vertices dropped.  Parabolic inflow on ``x = 0``, no-slip on the walls and
    the hole, do-nothing outlet with ``p = 0`` on ``x = length``."""
    hx, hy = length / nx, height / ny
    rng = np.random.default_rng(seed)
    ii, jj = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), indexing="xy")
    ii = ii.reshape(-1)
    jj = jj.reshape(-1)
    x = ii * hx
    y = jj * hy
    x[ii == nx] = length
    y[jj == ny] = height
    dx = rng.uniform(-jitter * hx, jitter * hx, size=x.shape)
    dy = rng.uniform(-jitter * hy, jitter * hy, size=y.shape)
    interior = (ii > 0) & (ii < nx) & (jj > 0) & (jj < ny)
    x = np.where(interior, x + dx, x)
    y = np.where(interior, y + dy, y)
    xy_all = np.stack([x, y], axis=1).astype(np.float32)
    sq_i, sq_j = np.meshgrid(np.arange(nx), np.arange(ny), indexing="xy")
    sq_i = sq_i.reshape(-1)
    sq_j = sq_j.reshape(-1)
    v00 = sq_j * (nx + 1) + sq_i
    v10 = v00 + 1
    v01 = v00 + (nx + 1)
    v11 = v01 + 1
    tris = np.stack([np.stack([v00, v10, v11], 1), np.stack([v00, v11, v01], 1)], 1).reshape(-1, 3)
    cxy = xy_all[tris].astype(np.float64).mean(axis=1)
    cx, cy, r = disk
    keep = (cxy[:, 0] - cx) ** 2 + (cxy[:, 1] - cy) ** 2 >= r * r
    tris = tris[keep]
    used = np.zeros(xy_all.shape[0], dtype=bool)
    used[tris.reshape(-1)] = True
    remap = -np.ones(xy_all.shape[0], dtype=np.int64)
    remap[used] = np.arange(int(used.sum()))
    tris = remap[tris]
    xy_v = xy_all[used]
    gi, gj = ii[used], jj[used]
    nodes, elements, edges, counts = _p2_from_p1(xy_v, tris)
    P = xy_v.shape[0]
    left_v, right_v = gi == 0, gi == nx
    bot_v, top_v = gj == 0, gj == ny
    bedge = counts == 1
    e0, e1 = edges[:, 0], edges[:, 1]
    e_left = bedge & left_v[e0] & left_v[e1]
    e_right = bedge & right_v[e0] & right_v[e1]
    e_bot = bedge & bot_v[e0] & bot_v[e1]
    e_top = bedge & top_v[e0] & top_v[e1]
    e_hole = bedge & ~(e_left | e_right | e_bot | e_top)
    hole_v = np.zeros(P, dtype=bool)
    hole_v[e0[e_hole]] = True
    hole_v[e1[e_hole]] = True
    inflow = np.concatenate([np.nonzero(left_v)[0], P + np.nonzero(e_left)[0]])
    walls = np.concatenate([np.nonzero((bot_v | top_v) & ~left_v)[0],
                            P + np.nonzero(e_bot | e_top)[0]])
    hole = np.concatenate([np.nonzero(hole_v)[0], P + np.nonzero(e_hole)[0]])
    inflow, walls, hole = np.sort(inflow), np.sort(walls), np.sort(hole)
    vb = np.concatenate([inflow, walls, hole]).astype(np.int32)
    yin = nodes[inflow, 1].astype(np.float64)
    gu = np.concatenate([_inflow(yin, umax, height), np.zeros(len(walls) + len(hole))])
    gv = np.zeros(len(vb))
    pb = np.nonzero(right_v)[0].astype(np.int32)
    m = Mesh(nodes=nodes, elements=elements, pressure_count=P,
             vel_bc_nodes=vb, vel_bc_u=gu, vel_bc_v=gv,
             p_bc_nodes=pb, p_bc_values=np.zeros(len(pb)),
             name=f"channel{nx}x{ny}s{seed}",
             info={"nx": nx, "ny": ny, "seed": seed, "umax": umax,
                   "removed_triangles": int((~keep).sum())},
             grid=np.stack([gi, gj], axis=1).astype(np.int32))
    return m


# ---------------------------------------------------------------------------
# Reference JSON mesh format (cpp/grid.cpp:9-77) and its BC expressions.
# ---------------------------------------------------------------------------
_TOKEN = re.compile(r"\s*(?:(\d+\.?\d*(?:[eE][-+]?\d+)?|\.\d+(?:[eE][-+]?\d+)?)|(sin|cos|pi)|([xy])|(.))")
_BIN_PREC = {"+": 0, "-": 0, "*": 1, "/": 1, "^": 3}


def _tokenize(s: str):
    out = []
    pos = 0
    s = s.strip()
    while pos < len(s):
        m = _TOKEN.match(s, pos)
        if not m or m.end() == pos:
            raise ValueError(f"cannot parse expression {s!r} at {pos}")
        num, fn, var, op = m.groups()
        if num is not None:
            out.append(("num", float(num)))
        elif fn is not None:
            out.append(("num", math.pi) if fn == "pi" else ("fn", fn))
        elif var is not None:
            out.append(("var", var))
        elif op is not None and op.strip():
            if op not in "+-*/^()":
                raise ValueError(f"unexpected symbol {op!r} in {s!r}")
            out.append(("op", op))
        pos = m.end()
    return out


def evaluate_expression(expr: str, x, y):
    """Evaluate a reference BC expression (operator precedence of
    cpp/expression.cpp:195-213: + - < * / < unary minus < ^ < sin cos; all
    binary operators left-associative) in float64, vectorised over x, y."""
    toks = _tokenize(expr)
    pos = 0

    def peek():
        return toks[pos] if pos < len(toks) else (None, None)

    def parse(min_prec):
        nonlocal pos
        kind, val = peek()
        if kind == "op" and val == "-":
            pos += 1
            lhs = -parse(2)
        elif kind == "op" and val == "(":
            pos += 1
            lhs = parse(0)
            if peek() != ("op", ")"):
                raise ValueError(f"missing ')' in {expr!r}")
            pos += 1
        elif kind == "fn":
            pos += 1
            arg = parse(10)
            lhs = np.sin(arg) if val == "sin" else np.cos(arg)
        elif kind == "num":
            pos += 1
            lhs = np.asarray(val, dtype=np.float64)
        elif kind == "var":
            pos += 1
            lhs = np.asarray(x if val == "x" else y, dtype=np.float64)
        else:
            raise ValueError(f"unexpected token in {expr!r}")
        while True:
            kind, val = peek()
            if kind != "op" or val not in _BIN_PREC or _BIN_PREC[val] < min_prec:
                break
            prec = _BIN_PREC[val]
            pos += 1
            rhs = parse(prec + 1)
            if val == "+":
                lhs = lhs + rhs
            elif val == "-":
                lhs = lhs - rhs
            elif val == "*":
                lhs = lhs * rhs
            elif val == "/":
                lhs = lhs / rhs
            else:
                lhs = np.power(lhs, rhs)
        return lhs

    res = parse(0)
    if pos != len(toks):
        raise ValueError(f"trailing tokens in {expr!r}")
    return np.broadcast_to(res, np.broadcast(np.asarray(x), np.asarray(y)).shape).astype(np.float64)


def load_json(path_or_dict) -> Mesh:
    """Read the reference's mesh JSON (keys ``elements, nodes, elementsCount,
    velocityNodesCount, pressureNodesCount, elementSize, uDirichlet,
    pDirichlet``; cpp/grid.cpp:19-72)."""
    d = path_or_dict
    if not isinstance(d, dict):
        with open(path_or_dict) as f:
            d = json.load(f)
    if int(d["elementSize"]) != 6:
        raise ValueError("only P2-P1 elements (elementSize 6) are supported")
    nodes = np.asarray(d["nodes"], dtype=np.float64).astype(np.float32)
    elements = np.asarray(d["elements"], dtype=np.int32).reshape(-1, 6)
    if nodes.shape[0] != int(d["velocityNodesCount"]) or elements.shape[0] != int(d["elementsCount"]):
        raise ValueError("node/element count mismatch")
    xs = nodes[:, 0].astype(np.float64)
    ys = nodes[:, 1].astype(np.float64)
    vb, gu, gv = [], [], []
    for bc in d["uDirichlet"]:
        idx = np.asarray(bc["nodes"], dtype=np.int32)
        vb.append(idx)
        gu.append(evaluate_expression(str(bc["u"]), xs[idx], ys[idx]))
        gv.append(evaluate_expression(str(bc["v"]), xs[idx], ys[idx]))
    pb, gp = [], []
    for bc in d["pDirichlet"]:
        idx = np.asarray(bc["nodes"], dtype=np.int32)
        pb.append(idx)
        gp.append(evaluate_expression(str(bc["p"]), xs[idx], ys[idx]))
    cat = lambda a, dt: np.concatenate(a).astype(dt) if a else np.zeros(0, dt)
    return Mesh(nodes=nodes, elements=elements, pressure_count=int(d["pressureNodesCount"]),
                vel_bc_nodes=cat(vb, np.int32), vel_bc_u=cat(gu, np.float64),
                vel_bc_v=cat(gv, np.float64), p_bc_nodes=cat(pb, np.int32),
                p_bc_values=cat(gp, np.float64), name="json")
