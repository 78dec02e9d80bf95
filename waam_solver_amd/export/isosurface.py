"""Isosurface of a cell field on the host: cell-to-point averaging and marching tetrahedra.

``cell_to_point`` gives each grid point the mean of the (up to eight) cells that share it, the
averaging VTK's cell-data-to-point-data filter applies. ``marching_tetrahedra`` splits every
point-grid cube into the six tetrahedra around its main diagonal (the Kuhn split, which matches
across neighbouring cubes, so the surface has no cracks) and cuts each tetrahedron that straddles
the level with one triangle (one or three vertices inside) or two (two inside). Every surface
vertex lies on a grid edge or a cube/face diagonal at the linearly interpolated crossing; point
arrays are interpolated with the same weights. A crossing that falls exactly on a grid point (a
point value equal to the level) is that point, shared by every tetrahedron there, and triangles
that collapse through it are dropped. The triangles are consistently oriented with their normals
pointing from the inside (value > level) outward.
"""

from __future__ import annotations

import numpy as np

# cube corner c = dx + 2 dy + 4 dz; six tetrahedra around the 0-7 diagonal
_CORNER = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0],
                    [0, 0, 1], [1, 0, 1], [0, 1, 1], [1, 1, 1]], dtype=np.int64)
_TETS = ((0, 1, 3, 7), (0, 3, 2, 7), (0, 2, 6, 7), (0, 6, 4, 7), (0, 4, 5, 7), (0, 5, 1, 7))


def cell_to_point(a: np.ndarray) -> np.ndarray:
    """(nx, ny, nz) cell values -> (nx+1, ny+1, nz+1) point values, each the mean of its cells."""
    a = np.asarray(a, dtype=np.float64)
    nx, ny, nz = a.shape
    s = np.zeros((nx + 1, ny + 1, nz + 1))
    n = np.zeros((nx + 1, ny + 1, nz + 1))
    for ox in (0, 1):
        for oy in (0, 1):
            for oz in (0, 1):
                s[ox:ox + nx, oy:oy + ny, oz:oz + nz] += a
                n[ox:ox + nx, oy:oy + ny, oz:oz + nz] += 1.0
    return s / n


def _tet_cases():
    """For each of the 16 inside-masks of a tetrahedron (bit v = vertex v inside), its
    triangles as triples of (inside vertex, outside vertex) edges."""
    cases = {}
    for m in range(16):
        ins = [v for v in range(4) if m >> v & 1]
        out = [v for v in range(4) if not m >> v & 1]
        if len(ins) in (0, 4):
            cases[m] = []
        elif len(ins) == 1:
            a = ins[0]
            cases[m] = [((a, out[0]), (a, out[1]), (a, out[2]))]
        elif len(ins) == 3:
            b = out[0]
            cases[m] = [((ins[0], b), (ins[1], b), (ins[2], b))]
        else:
            a, b = ins
            c, d = out
            cases[m] = [((a, c), (a, d), (b, d)), ((a, c), (b, d), (b, c))]
    return cases


_CASES = _tet_cases()


def marching_tetrahedra(P: np.ndarray, level: float, point_arrays: dict[str, np.ndarray]):
    """(vertices in point-index coordinates (m, 3), triangles (t, 3), {name: (m,) values})."""
    P = np.asarray(P, dtype=np.float64)
    X, Y, Z = P.shape
    if min(X, Y, Z) < 2:
        return np.zeros((0, 3)), np.zeros((0, 3), np.int64), {k: np.zeros(0) for k in point_arrays}
    cx, cy, cz = X - 1, Y - 1, Z - 1
    inside = P > level
    # cubes with both inside and outside corners
    any_in = np.zeros((cx, cy, cz), bool)
    all_in = np.ones((cx, cy, cz), bool)
    for ox, oy, oz in _CORNER:
        v = inside[ox:ox + cx, oy:oy + cy, oz:oz + cz]
        any_in |= v
        all_in &= v
    cubes = np.argwhere(any_in & ~all_in)
    flatP = P.reshape(-1)
    strides = np.array([Y * Z, Z, 1], dtype=np.int64)
    edges_a, edges_b = [], []
    for tet in _TETS:
        corner_pts = [(cubes + _CORNER[c]) @ strides for c in tet]   # global point ids (n, 4)
        ids = np.stack(corner_pts, axis=1)
        mask = np.zeros(len(cubes), np.int64)
        for v in range(4):
            mask |= (flatP[ids[:, v]] > level).astype(np.int64) << v
        for m, tris in _CASES.items():
            if not tris:
                continue
            sel = ids[mask == m]
            if len(sel) == 0:
                continue
            for tri in tris:
                edges_a.append(np.stack([sel[:, e[0]] for e in tri], axis=1))
                edges_b.append(np.stack([sel[:, e[1]] for e in tri], axis=1))
    if not edges_a:
        return np.zeros((0, 3)), np.zeros((0, 3), np.int64), {k: np.zeros(0) for k in point_arrays}
    ea = np.concatenate(edges_a)  # (t, 3) inside point of each triangle vertex's edge
    eb = np.concatenate(edges_b)  # (t, 3) outside point
    Np = X * Y * Z
    # a crossing exactly at the outside point (value == level) is that grid point itself: key it
    # by the point so every tetrahedron meeting there shares the vertex
    at_point = flatP[eb] == level
    key = np.where(at_point, -(eb + 1), ea * Np + eb)
    uniq, inv = np.unique(key.reshape(-1), return_inverse=True)
    inv = inv.reshape(-1, 3)
    on_pt = uniq < 0
    pa = np.where(on_pt, -uniq - 1, uniq // Np)
    pb = np.where(on_pt, -uniq - 1, uniq % Np)
    va, vb = flatP[pa], flatP[pb]
    with np.errstate(invalid="ignore", divide="ignore"):
        t = np.where(on_pt, 0.0, (level - va) / np.where(on_pt, 1.0, vb - va))
    ca = np.stack(np.unravel_index(pa, (X, Y, Z)), axis=1).astype(np.float64)
    cb = np.stack(np.unravel_index(pb, (X, Y, Z)), axis=1).astype(np.float64)
    verts = ca + t[:, None] * (cb - ca)
    # triangles that collapsed onto a grid point or an edge through merged vertices are dropped
    keep = (inv[:, 0] != inv[:, 1]) & (inv[:, 1] != inv[:, 2]) & (inv[:, 0] != inv[:, 2])
    tris = inv[keep].astype(np.int64)
    ea0, eb0 = ea[keep, 0], eb[keep, 0]
    # orientation: each normal points along its first vertex's edge, inside -> outside
    p0, p1, p2 = verts[tris[:, 0]], verts[tris[:, 1]], verts[tris[:, 2]]
    nrm = np.cross(p1 - p0, p2 - p0)
    outward = (np.stack(np.unravel_index(eb0, (X, Y, Z)), axis=1)
               - np.stack(np.unravel_index(ea0, (X, Y, Z)), axis=1)).astype(np.float64)
    flip = (nrm * outward).sum(axis=1) < 0
    tris[flip] = tris[flip][:, [0, 2, 1]]
    data = {}
    for name, arr in point_arrays.items():
        f = np.asarray(arr, dtype=np.float64).reshape(-1)
        data[name] = f[pa] + t * (f[pb] - f[pa])
    return verts, tris, data
