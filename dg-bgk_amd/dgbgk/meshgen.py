"""Synthetic meshes in the reference's conventions.

A mesh is the triple the reference reads from its ``.plt`` file
(src/gtinpt.f90:24-50): ``intma`` (E,3) 1-based point ids of each linear
triangle, ``coord`` (P,2), and ``bsido`` (B,6) boundary sides
``p1 p2 element type Tw alpha*100``.  Elements are counter-clockwise and every
boundary side has the domain on its left, the orientation the reference's
side builder and normals assume (src/getsid.f90:85-129, src/normal.f90:27-31,
src/getnor.f90:21-31).  Boundary types: 1 inflow, 2 outflow, 3 wall
without the mass-flux parameter, 4 wall with it.

Elements and points are numbered along a Hilbert curve so neighbouring
triangles are neighbours in memory.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass
class Mesh:
    intma: np.ndarray  # (E,3) int32, 1-based
    coord: np.ndarray  # (P,2) float64
    bsido: np.ndarray  # (B,6) int32
    name: str = "mesh"

    @property
    def nelem(self) -> int:
        return int(self.intma.shape[0])

    @property
    def npoin(self) -> int:
        return int(self.coord.shape[0])

    @property
    def nboun(self) -> int:
        return int(self.bsido.shape[0])


def hilbert_index(x: np.ndarray, y: np.ndarray, order: int = 16) -> np.ndarray:
    """Hilbert-curve index of points scaled into [0, 2**order)^2."""
    n = 1 << order
    lo = np.array([x.min(), y.min()])
    span = max(float(x.max() - lo[0]), float(y.max() - lo[1]), 1e-300)
    xi = np.minimum(((x - lo[0]) / span * (n - 1)).astype(np.int64), n - 1)
    yi = np.minimum(((y - lo[1]) / span * (n - 1)).astype(np.int64), n - 1)
    d = np.zeros_like(xi)
    s = n >> 1
    while s > 0:
        rx = ((xi & s) > 0).astype(np.int64)
        ry = ((yi & s) > 0).astype(np.int64)
        d += s * s * ((3 * rx) ^ ry)
        # rotate the quadrant
        flip = ry == 0
        swap_x = np.where(flip & (rx == 1), n - 1 - xi, xi)
        swap_y = np.where(flip & (rx == 1), n - 1 - yi, yi)
        xi = np.where(flip, swap_y, swap_x)
        yi = np.where(flip, swap_x, swap_y)
        s >>= 1
    return d


def _orient_ccw(coord: np.ndarray, tri: np.ndarray) -> np.ndarray:
    p = coord[tri]
    area2 = (p[:, 1, 0] - p[:, 0, 0]) * (p[:, 2, 1] - p[:, 0, 1]) - (
        p[:, 2, 0] - p[:, 0, 0]) * (p[:, 1, 1] - p[:, 0, 1])
    tri = tri.copy()
    neg = area2 < 0
    tri[neg] = tri[neg][:, [0, 2, 1]]
    return tri


def _boundary_edges(tri: np.ndarray):
    """Edges used by exactly one triangle, oriented as in that (CCW)
    triangle, so the triangle lies on the edge's left."""
    e = np.concatenate([tri[:, [0, 1]], tri[:, [1, 2]], tri[:, [2, 0]]])
    owner = np.concatenate([np.arange(len(tri))] * 3)
    key = np.sort(e, axis=1)
    k = key[:, 0].astype(np.int64) * (int(tri.max()) + 1) + key[:, 1]
    order = np.argsort(k, kind="stable")
    ks = k[order]
    uniq, first, counts = np.unique(ks, return_index=True, return_counts=True)
    single = order[first[counts == 1]]
    if np.any(counts > 2):
        raise ValueError("non-manifold triangulation")
    return e[single], owner[single]


def _finish(coord, tri, classify, name, tw=1, alpha100=100):
    """Orient, renumber along a Hilbert curve, drop unused points and build
    BSIDO with types from ``classify(midpoints) -> int array``."""
    tri = _orient_ccw(coord, tri)
    used = np.unique(tri)
    remap = -np.ones(len(coord), dtype=np.int64)
    # points along the Hilbert curve
    hp = hilbert_index(coord[used, 0], coord[used, 1])
    used = used[np.argsort(hp, kind="stable")]
    remap[used] = np.arange(len(used))
    coord = coord[used]
    tri = remap[tri]
    # elements along the Hilbert curve of their centroids
    cen = coord[tri].mean(axis=1)
    he = hilbert_index(cen[:, 0], cen[:, 1])
    tri = tri[np.argsort(he, kind="stable")]
    edges, owner = _boundary_edges(tri)
    mid = coord[edges].mean(axis=1)
    btype = np.asarray(classify(mid, coord[edges[:, 0]], coord[edges[:, 1]]),
                       dtype=np.int64)
    # boundary sides in element order, then first point
    order = np.lexsort((edges[:, 0], owner))
    edges, owner, btype = edges[order], owner[order], btype[order]
    bsido = np.zeros((len(edges), 6), dtype=np.int32)
    bsido[:, 0] = edges[:, 0] + 1
    bsido[:, 1] = edges[:, 1] + 1
    bsido[:, 2] = owner + 1
    bsido[:, 3] = btype
    bsido[:, 4] = np.where(btype >= 3, tw, 0)
    bsido[:, 5] = np.where(btype >= 3, alpha100, 0)
    return Mesh(intma=(tri + 1).astype(np.int32), coord=coord.astype(np.float64),
                bsido=bsido, name=name)


def unit_square(n: int = 32, left: int = 1, right: int = 2, bottom: int = 3,
                top: int = 3, tw: int = 1, alpha100: int = 0) -> Mesh:
    """n x n cells on the unit square, each split along its rising diagonal
    into two CCW triangles: 2 n^2 elements, (n+1)^2 points, 4 n boundary
    sides.  Elements are numbered cell by cell, row-major (already local)."""
    m = n + 1
    xs = np.linspace(0.0, 1.0, m)
    X, Y = np.meshgrid(xs, xs, indexing="xy")
    coord = np.stack([X.ravel(), Y.ravel()], axis=1)

    def pid(i, j):
        return j * m + i

    tris = []
    for j in range(n):
        for i in range(n):
            p00, p10, p11, p01 = pid(i, j), pid(i + 1, j), pid(i + 1, j + 1), pid(i, j + 1)
            tris.append((p00, p10, p11))
            tris.append((p00, p11, p01))
    tri = np.asarray(tris, dtype=np.int64)
    b = []
    for i in range(n):  # bottom, left to right: cell (i,0), lower triangle
        b.append((pid(i, 0), pid(i + 1, 0), 2 * (0 * n + i), bottom))
    for j in range(n):  # right, upwards: cell (n-1,j), lower triangle
        b.append((pid(n, j), pid(n, j + 1), 2 * (j * n + n - 1), right))
    for i in range(n):  # top, right to left: cell (i,n-1), upper triangle
        b.append((pid(i + 1, n), pid(i, n), 2 * ((n - 1) * n + i) + 1, top))
    for j in range(n):  # left, downwards: cell (0,j), upper triangle
        b.append((pid(0, j + 1), pid(0, j), 2 * (j * n) + 1, left))
    bsido = np.zeros((len(b), 6), dtype=np.int32)
    for k, (p1, p2, e, t) in enumerate(b):
        bsido[k] = (p1 + 1, p2 + 1, e + 1, t, tw if t >= 3 else 0,
                    alpha100 if t >= 3 else 0)
    return Mesh(intma=(tri + 1).astype(np.int32), coord=coord, bsido=bsido,
                name=f"unit_square_{n}")


def _ring_points(cx, cy, r, h):
    n = max(8, int(round(2 * np.pi * r / h)))
    return n


def body_in_box(lx: float, ly: float, cx: float, cy: float, diameter: float,
                h0: float, growth: float, hmax: float, seed: int = 0,
                wall_type: int = 4, tw: int = 1, alpha100: int = 100,
                far_types=(1, 2, 2, 2), name: str = "body") -> Mesh:
    """Delaunay mesh of the box [0,lx]x[0,ly] minus a circle of the given
    diameter at (cx,cy), graded from spacing h0 on the circle by
    h(d) = min(hmax, h0 + growth*d) with d the distance from the surface.

    Points: concentric rings around the body (alternately rotated by half a
    spacing so the rings triangulate into near-equilateral triangles), then
    a jittered hexagonal lattice of spacing hmax outside, plus points on the
    box edges.  far_types = types of the (left, right, bottom, top) box
    sides; the circle's sides get wall_type.
    """
    rng = np.random.default_rng(seed)
    r0 = 0.5 * diameter
    pts = []
    r = r0
    k = 0
    while True:
        h = min(hmax, h0 + growth * (r - r0))
        n = _ring_points(cx, cy, r, h)
        off = (0.5 * (k % 2) + 0.05 * rng.standard_normal()) * 2 * np.pi / n
        th = off + 2 * np.pi * np.arange(n) / n
        ring = np.stack([cx + r * np.cos(th), cy + r * np.sin(th)], axis=1)
        inside = (ring[:, 0] > 0.5 * hmax) & (ring[:, 0] < lx - 0.5 * hmax) & (
            ring[:, 1] > 0.5 * hmax) & (ring[:, 1] < ly - 0.5 * hmax)
        pts.append(ring[inside])
        if h >= hmax or not inside.any():
            break
        r = r + h * np.sqrt(3.0) / 2.0
        k += 1
    r_rings = r
    # hexagonal lattice outside the rings
    dy = hmax * np.sqrt(3.0) / 2.0
    ys = np.arange(0.5 * dy, ly - 0.25 * dy, dy)
    lat = []
    for j, y in enumerate(ys):
        x0 = 0.5 * hmax * (1 + (j % 2))
        xs = np.arange(x0, lx - 0.25 * hmax, hmax)
        lat.append(np.stack([xs, np.full_like(xs, y)], axis=1))
    lat = np.concatenate(lat)
    lat += 0.08 * hmax * rng.standard_normal(lat.shape)
    d = np.hypot(lat[:, 0] - cx, lat[:, 1] - cy)
    lat = lat[d > r_rings + 0.6 * hmax]
    lat = lat[(lat[:, 0] > 0.3 * hmax) & (lat[:, 0] < lx - 0.3 * hmax) &
              (lat[:, 1] > 0.3 * hmax) & (lat[:, 1] < ly - 0.3 * hmax)]
    pts.append(lat)
    # box edges
    nx = max(2, int(round(lx / hmax)))
    ny = max(2, int(round(ly / hmax)))
    bx = np.linspace(0.0, lx, nx + 1)
    by = np.linspace(0.0, ly, ny + 1)[1:-1]
    edge = np.concatenate([
        np.stack([bx, np.zeros_like(bx)], 1), np.stack([bx, np.full_like(bx, ly)], 1),
        np.stack([np.zeros_like(by), by], 1), np.stack([np.full_like(by, lx), by], 1)])
    pts.append(edge)
    coord = np.concatenate(pts)
    from scipy.spatial import Delaunay

    tri = Delaunay(coord).simplices.astype(np.int64)
    cen = coord[tri].mean(axis=1)
    keep = np.hypot(cen[:, 0] - cx, cen[:, 1] - cy) > r0
    tri = tri[keep]
    p = coord[tri]
    area2 = np.abs((p[:, 1, 0] - p[:, 0, 0]) * (p[:, 2, 1] - p[:, 0, 1]) -
                   (p[:, 2, 0] - p[:, 0, 0]) * (p[:, 1, 1] - p[:, 0, 1]))
    tri = tri[area2 > 1e-10 * h0 * h0]
    tol = 1e-6 * max(lx, ly)

    def classify(mid, a, b):
        t = np.full(len(mid), -1, dtype=np.int64)
        on_body = (np.hypot(a[:, 0] - cx, a[:, 1] - cy) < r0 * (1 + 1e-9) + tol) & (
            np.hypot(b[:, 0] - cx, b[:, 1] - cy) < r0 * (1 + 1e-9) + tol)
        t[on_body] = wall_type
        t[(t < 0) & (np.abs(mid[:, 0]) < tol)] = far_types[0]
        t[(t < 0) & (np.abs(mid[:, 0] - lx) < tol)] = far_types[1]
        t[(t < 0) & (np.abs(mid[:, 1]) < tol)] = far_types[2]
        t[(t < 0) & (np.abs(mid[:, 1] - ly) < tol)] = far_types[3]
        if np.any(t < 0):
            raise ValueError("unclassified boundary side (mesh has a hole)")
        return t

    return _finish(coord, tri, classify, name, tw=tw, alpha100=alpha100)


def renumber_elements(m: Mesh, perm: np.ndarray, name: str | None = None) -> Mesh:
    """The same mesh with element i being element perm[i] of m: INTMA rows
    moved, BSIDO's element column (BSIDO(3)) renumbered."""
    perm = np.asarray(perm, dtype=np.int64)
    inv = np.empty_like(perm)
    inv[perm] = np.arange(len(perm))
    bsido = m.bsido.copy()
    bsido[:, 2] = (inv[m.bsido[:, 2] - 1] + 1).astype(np.int32)
    return Mesh(intma=np.ascontiguousarray(m.intma[perm]), coord=m.coord, bsido=bsido,
                name=name or m.name)


def shuffle_elements(m: Mesh, seed: int = 0) -> tuple[Mesh, np.ndarray]:
    """m with its elements in a random order (an input mesh in arbitrary
    element numbering); returns the mesh and perm (new element i is element
    perm[i] of m)."""
    perm = np.random.default_rng(seed).permutation(m.nelem)
    return renumber_elements(m, perm, f"{m.name}-shuffled{seed}"), perm


def mesh_quality(m: Mesh) -> dict:
    p = m.coord[m.intma - 1]
    a = np.linalg.norm(p[:, 1] - p[:, 0], axis=1)
    b = np.linalg.norm(p[:, 2] - p[:, 1], axis=1)
    c = np.linalg.norm(p[:, 0] - p[:, 2], axis=1)
    area2 = (p[:, 1, 0] - p[:, 0, 0]) * (p[:, 2, 1] - p[:, 0, 1]) - (
        p[:, 2, 0] - p[:, 0, 0]) * (p[:, 1, 1] - p[:, 0, 1])
    hmin = area2 / np.maximum(np.maximum(a, b), c)
    return {"min_area2": float(area2.min()), "min_altitude": float(hmin.min()),
            "max_edge": float(np.maximum(np.maximum(a, b), c).max())}


def wall_distance(m: Mesh, wall_types=(3, 4)) -> np.ndarray:
    """Distance from every point to the nearest wall side (segment)."""
    sel = np.isin(m.bsido[:, 3], wall_types)
    if not sel.any():
        return np.full(m.npoin, np.inf)
    a = m.coord[m.bsido[sel, 0] - 1]
    b = m.coord[m.bsido[sel, 1] - 1]
    out = np.empty(m.npoin)
    ab = b - a
    L2 = np.maximum((ab * ab).sum(1), 1e-300)
    chunk = max(1, 4_000_000 // max(1, len(a)))
    for s in range(0, m.npoin, chunk):
        q = m.coord[s:s + chunk, None, :]
        t = np.clip(((q - a[None]) * ab[None]).sum(-1) / L2[None], 0.0, 1.0)
        proj = a[None] + t[..., None] * ab[None]
        out[s:s + chunk] = np.sqrt(((q - proj) ** 2).sum(-1)).min(axis=1)
    return out
