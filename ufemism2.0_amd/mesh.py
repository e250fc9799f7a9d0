"""Triangular mesh container and synthetic mesh generation.

The container mirrors the parts of UFEMISM's ``type_mesh`` that the velocity
solve and the heat columns read (src/UPSY/types/mesh_types.f90:57-94):

* ``V``      vertex coordinates                     (nV, 2)  float64
* ``Tri``    triangle vertices, counter-clockwise     (nT, 3)  int32, 0-based
* ``TriC``   neighbour across from vertex n, -1 if none (nT, 3) int32
* ``VBI``    vertex border index 0..8 (1 = N, 2 = NE, ..., 8 = NW)
* ``TriBI``  triangle border index 0..8 (mesh_secondary.f90:72-135)
* ``TriGC``  triangle geometric centre                (nT, 2)
* ``iTri``   triangles around each vertex, counter-clockwise (CSR)
* ``C``      vertex neighbours (CSR)

Indices are 0-based here; the C-ABI converts nothing (it takes 0-based
arrays) and the Fortran shim subtracts one before calling it.

Mesh *generation* in the reference is a Delaunay refinement code
(src/UPSY/mesh/mesh_refinement_*.f90) that is out of scope for this engine.
For synthetic workloads this module places points with a size-field-driven
quadtree, triangulates them with Qhull (scipy.spatial.Delaunay) and smooths
the interior with a few Laplacian sweeps.  The topology conventions
(counter-clockwise ordering, TriC across from vertex n, border indices) follow
the reference so that the operator construction sees the same kind of mesh.
"""

from __future__ import annotations

import dataclasses
import hashlib

import numpy as np


@dataclasses.dataclass
class Mesh:
    xmin: float
    xmax: float
    ymin: float
    ymax: float
    V: np.ndarray
    Tri: np.ndarray
    TriC: np.ndarray
    VBI: np.ndarray
    TriBI: np.ndarray
    TriGC: np.ndarray
    iTri_ptr: np.ndarray
    iTri_ind: np.ndarray
    C_ptr: np.ndarray
    C_ind: np.ndarray

    @property
    def nV(self) -> int:
        return int(self.V.shape[0])

    @property
    def nT(self) -> int:
        return int(self.Tri.shape[0])

    def digest(self) -> str:
        h = hashlib.sha1()
        h.update(np.ascontiguousarray(self.V).tobytes())
        h.update(np.ascontiguousarray(self.Tri).tobytes())
        return h.hexdigest()[:12]


# --------------------------------------------------------------------------
# point placement
# --------------------------------------------------------------------------

def _quadtree_points(xmin, xmax, ymin, ymax, size_fn, rng, jitter=0.15):
    """Cell centres of a quadtree refined until every cell is <= size_fn."""
    Lx, Ly = xmax - xmin, ymax - ymin
    # Level-0 cells are no larger than the largest target size sampled on a
    # coarse probe grid; cells are then split while larger than sqrt(2) times
    # the local target, so the final sizes straddle the target.
    gx, gy = np.meshgrid(np.linspace(xmin, xmax, 65), np.linspace(ymin, ymax, 65), indexing="ij")
    hmax = float(np.max(size_fn(gx.ravel(), gy.ravel())))
    nx = max(1, int(np.ceil(Lx / hmax - 1e-9)))
    ny = max(1, int(np.ceil(Ly / hmax - 1e-9)))
    sx, sy = Lx / nx, Ly / ny
    ix, iy = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    x0 = (xmin + ix * sx).ravel()
    y0 = (ymin + iy * sy).ravel()
    w = np.full(x0.shape, sx)
    h = np.full(x0.shape, sy)
    done_x, done_y, done_w, done_h = [], [], [], []
    for _ in range(40):
        cx, cy = x0 + 0.5 * w, y0 + 0.5 * h
        # a cell is fine enough when its diagonal-equivalent size is below
        # the smallest target size sampled at its centre and corners
        hs = np.minimum.reduce([size_fn(cx, cy), size_fn(x0, y0), size_fn(x0 + w, y0),
                                size_fn(x0, y0 + h), size_fn(x0 + w, y0 + h)])
        split = np.maximum(w, h) > np.sqrt(2.0) * hs
        done_x.append(x0[~split]); done_y.append(y0[~split])
        done_w.append(w[~split]); done_h.append(h[~split])
        if not split.any():
            break
        x0, y0, w, h = x0[split], y0[split], 0.5 * w[split], 0.5 * h[split]
        x0 = np.concatenate([x0, x0 + w, x0, x0 + w])
        y0 = np.concatenate([y0, y0, y0 + h, y0 + h])
        w = np.concatenate([w] * 4)
        h = np.concatenate([h] * 4)
    x0 = np.concatenate(done_x); y0 = np.concatenate(done_y)
    w = np.concatenate(done_w); h = np.concatenate(done_h)

    # interior points: jittered cell centres
    px = x0 + 0.5 * w + jitter * w * rng.uniform(-1, 1, x0.shape)
    py = y0 + 0.5 * h + jitter * h * rng.uniform(-1, 1, y0.shape)

    # border points: the corners of cells that touch the domain edges
    tol = 1e-9 * max(Lx, Ly)
    bx, by = [], []
    s = np.abs(y0 - ymin) < tol
    bx += [x0[s], x0[s] + w[s]]; by += [np.full(s.sum(), ymin)] * 2
    s = np.abs(y0 + h - ymax) < tol
    bx += [x0[s], x0[s] + w[s]]; by += [np.full(s.sum(), ymax)] * 2
    s = np.abs(x0 - xmin) < tol
    by += [y0[s], y0[s] + h[s]]; bx += [np.full(s.sum(), xmin)] * 2
    s = np.abs(x0 + w - xmax) < tol
    by += [y0[s], y0[s] + h[s]]; bx += [np.full(s.sum(), xmax)] * 2
    B = np.stack([np.concatenate(bx), np.concatenate(by)], axis=1)
    # snap to the exact domain edges and deduplicate
    B[:, 0] = np.clip(B[:, 0], xmin, xmax)
    B[:, 1] = np.clip(B[:, 1], ymin, ymax)
    key = np.round((B - [xmin, ymin]) / tol).astype(np.int64)
    _, first = np.unique(key, axis=0, return_index=True)
    B = B[np.sort(first)]
    P = np.stack([px, py], axis=1)
    return B, P


def _triangulate(P):
    from scipy.spatial import Delaunay  # setup-time only
    d = Delaunay(P, qhull_options="Qbb Qc Qz Q12")
    tri = d.simplices.astype(np.int64)
    nb = d.neighbors.astype(np.int64)
    # orient counter-clockwise; neighbours stay "across from vertex n"
    a, b, c = P[tri[:, 0]], P[tri[:, 1]], P[tri[:, 2]]
    cross = (b[:, 0] - a[:, 0]) * (c[:, 1] - a[:, 1]) - (b[:, 1] - a[:, 1]) * (c[:, 0] - a[:, 0])
    flip = cross < 0
    tri[flip, 1], tri[flip, 2] = tri[flip, 2].copy(), tri[flip, 1].copy()
    nb[flip, 1], nb[flip, 2] = nb[flip, 2].copy(), nb[flip, 1].copy()
    return tri, nb


def _laplacian_smooth(V, tri, is_border, sweeps):
    nV = V.shape[0]
    e = np.concatenate([tri[:, [0, 1]], tri[:, [1, 2]], tri[:, [2, 0]]])
    e = np.concatenate([e, e[:, ::-1]])
    for _ in range(sweeps):
        acc = np.zeros_like(V)
        np.add.at(acc, e[:, 0], V[e[:, 1]])
        cnt = np.bincount(e[:, 0], minlength=nV).astype(np.float64)
        Vn = acc / np.maximum(cnt, 1)[:, None]
        V = np.where(is_border[:, None], V, 0.5 * V + 0.5 * Vn)
    return V


def build_topology(V, tri, nb, xmin, xmax, ymin, ymax) -> Mesh:
    """Derive VBI, TriBI, TriGC, iTri and C from a CCW triangulation."""
    nV = V.shape[0]
    nT = tri.shape[0]
    tol = 1e-9 * max(xmax - xmin, ymax - ymin)
    x, y = V[:, 0], V[:, 1]
    on_n, on_s = np.abs(y - ymax) < tol, np.abs(y - ymin) < tol
    on_e, on_w = np.abs(x - xmax) < tol, np.abs(x - xmin) < tol
    VBI = np.zeros(nV, np.int32)
    VBI[on_n] = 1
    VBI[on_e] = 3
    VBI[on_s] = 5
    VBI[on_w] = 7
    VBI[on_n & on_e] = 2
    VBI[on_s & on_e] = 4
    VBI[on_s & on_w] = 6
    VBI[on_n & on_w] = 8

    # Triangle border index: the reference traces the border and stamps each
    # triangle touching border vertex vi with VBI(vi) (mesh_secondary.f90:97-124);
    # a triangle touching several border vertices keeps the last stamp.  Here
    # the stamp is taken from the vertex with the largest border index, and a
    # corner vertex's stamp wins, as in the corner fix-up (:127-130).
    vbi_t = VBI[tri]
    is_corner = (vbi_t > 0) & (vbi_t % 2 == 0)
    TriBI = np.where(is_corner.any(axis=1),
                     np.max(np.where(is_corner, vbi_t, 0), axis=1),
                     vbi_t.max(axis=1)).astype(np.int32)

    TriGC = V[tri].mean(axis=1)

    # iTri: triangles around each vertex, counter-clockwise by angle
    vt = tri.ravel()
    tt = np.repeat(np.arange(nT), 3)
    ang = np.arctan2(TriGC[tt, 1] - V[vt, 1], TriGC[tt, 0] - V[vt, 0])
    order = np.lexsort((ang, vt))
    iTri_ind = tt[order].astype(np.int32)
    iTri_ptr = np.concatenate([[0], np.cumsum(np.bincount(vt, minlength=nV))]).astype(np.int32)

    # C: vertex neighbours, counter-clockwise by angle
    e = np.concatenate([tri[:, [0, 1]], tri[:, [1, 2]], tri[:, [2, 0]]])
    e = np.concatenate([e, e[:, ::-1]])
    e = np.unique(e, axis=0)
    ang = np.arctan2(V[e[:, 1], 1] - V[e[:, 0], 1], V[e[:, 1], 0] - V[e[:, 0], 0])
    order = np.lexsort((ang, e[:, 0]))
    C_ind = e[order, 1].astype(np.int32)
    C_ptr = np.concatenate([[0], np.cumsum(np.bincount(e[:, 0], minlength=nV))]).astype(np.int32)

    return Mesh(float(xmin), float(xmax), float(ymin), float(ymax),
                np.ascontiguousarray(V, np.float64), np.ascontiguousarray(tri, np.int32),
                np.ascontiguousarray(nb, np.int32), VBI, TriBI,
                np.ascontiguousarray(TriGC), iTri_ptr, iTri_ind, C_ptr, C_ind)


def generate_mesh(xmin, xmax, ymin, ymax, size_fn, seed=0, smooth_sweeps=4) -> Mesh:
    """Graded Delaunay mesh of the rectangle with target edge length size_fn(x, y)."""
    rng = np.random.default_rng(seed)
    B, P = _quadtree_points(xmin, xmax, ymin, ymax, size_fn, rng)
    V = np.concatenate([B, P])
    is_border = np.zeros(V.shape[0], bool)
    is_border[: B.shape[0]] = True
    tri, nb = _triangulate(V)
    if smooth_sweeps > 0:
        V = _laplacian_smooth(V, tri, is_border, smooth_sweeps)
        tri, nb = _triangulate(V)
    return build_topology(V, tri, nb, xmin, xmax, ymin, ymax)


def uniform_mesh(xmin, xmax, ymin, ymax, h, seed=0) -> Mesh:
    return generate_mesh(xmin, xmax, ymin, ymax, lambda x, y: np.full(np.shape(x), float(h)), seed)


# --------------------------------------------------------------------------
# Voronoi geometry of the vertices (the mass-conservation operators)
# --------------------------------------------------------------------------

@dataclasses.dataclass
class VertexGeometry:
    """Per-vertex Voronoi data of the thickness solver, in the C (neighbour)
    order of the mesh: A (cell areas), Cw (shared boundary length per
    connection) and ETri (the triangles either side of each connection's
    edge, -1 where the edge is on the border)."""
    A: np.ndarray       # (nV,)
    Cw: np.ndarray      # (nC,)
    ETri: np.ndarray    # (nC, 2) int32


def circumcentres(mesh: Mesh) -> np.ndarray:
    """Triangle circumcentres (the reference's mesh%Tricc)."""
    a, b, c = (mesh.V[mesh.Tri[:, k]] for k in range(3))
    d = 2.0 * (a[:, 0] * (b[:, 1] - c[:, 1]) + b[:, 0] * (c[:, 1] - a[:, 1]) + c[:, 0] * (a[:, 1] - b[:, 1]))
    a2, b2, c2 = (np.sum(p * p, axis=1) for p in (a, b, c))
    ux = (a2 * (b[:, 1] - c[:, 1]) + b2 * (c[:, 1] - a[:, 1]) + c2 * (a[:, 1] - b[:, 1])) / d
    uy = (a2 * (c[:, 0] - b[:, 0]) + b2 * (a[:, 0] - c[:, 0]) + c2 * (b[:, 0] - a[:, 0])) / d
    return np.stack([ux, uy], axis=1)


def vertex_geometry(mesh: Mesh) -> VertexGeometry:
    """Voronoi cell areas and connection widths of the vertices.

    A: the circumcentric dual cell of each vertex, i.e. for every triangle
    around the vertex the two triangles (vertex, edge midpoint, circumcentre),
    signed, which is the Voronoi cell clipped to the domain the reference
    builds from the circumcentres of the surrounding triangles
    (calc_Voronoi_cell, mesh_utilities.f90:72-303; calc_Voronoi_cell_areas,
    mesh_secondary.f90:137-186).  The cells tile the domain.
    Cw: the shared Voronoi boundary of the connection's edge
    (find_shared_Voronoi_boundary, mesh_utilities.f90:305-372): between the
    circumcentres of the two triangles either side, cropped to the domain;
    for an edge on the border, from the triangle's circumcentre (clamped into
    the domain) to a point dx = (width + height) / 100 outside the border,
    perpendicular to it, as the reference draws it
    (calc_connection_widths, mesh_secondary.f90:246-298)."""
    V, Tri = mesh.V, mesh.Tri
    nV, nT = mesh.nV, mesh.nT
    cc = circumcentres(mesh)
    # edges -> the triangles either side: a CCW triangle has each of its
    # directed edges (va, vb) on its left
    nC = int(mesh.C_ptr[-1])
    ea = np.concatenate([Tri[:, n] for n in range(3)]).astype(np.int64)
    eb = np.concatenate([Tri[:, (n + 1) % 3] for n in range(3)]).astype(np.int64)
    et = np.tile(np.arange(nT, dtype=np.int64), 3)
    keys = ea * nV + eb
    order = np.argsort(keys)
    keys, et = keys[order], et[order]
    src = np.repeat(np.arange(nV, dtype=np.int64), np.diff(mesh.C_ptr))
    dst = mesh.C_ind.astype(np.int64)

    def lookup(k):
        pos = np.minimum(np.searchsorted(keys, k), len(keys) - 1)
        return np.where(keys[pos] == k, et[pos], -1).astype(np.int32)

    ETri = np.stack([lookup(src * nV + dst), lookup(dst * nV + src)], axis=1)
    # cell areas: signed sub-triangles (vertex, edge midpoint, circumcentre)
    def area2(p, q, r):
        return (q[:, 0] - p[:, 0]) * (r[:, 1] - p[:, 1]) - (q[:, 1] - p[:, 1]) * (r[:, 0] - p[:, 0])
    A = np.zeros(nV)
    for n in range(3):
        vi, vb, vc = Tri[:, n], Tri[:, (n + 1) % 3], Tri[:, (n + 2) % 3]
        p = V[vi]
        mb, mc = 0.5 * (p + V[vb]), 0.5 * (p + V[vc])
        s = 0.5 * (area2(p, mb, cc) + area2(p, cc, mc))
        np.add.at(A, vi, s)
    # connection widths
    xmin, xmax, ymin, ymax = mesh.xmin, mesh.xmax, mesh.ymin, mesh.ymax
    dx = ((xmax - xmin) + (ymax - ymin)) / 100.0
    tol = 1e-9 * max(xmax - xmin, ymax - ymin)
    lo, hi = np.array([xmin, ymin]), np.array([xmax, ymax])
    tl, tr = ETri[:, 0], ETri[:, 1]
    interior = (tl >= 0) & (tr >= 0)
    p = np.clip(cc[np.where(tl >= 0, tl, tr)], lo, hi)
    q = np.clip(cc[np.where(tr >= 0, tr, tl)], lo, hi)
    a, b = V[src], V[dst]
    north = (np.abs(a[:, 1] - ymax) < tol) & (np.abs(b[:, 1] - ymax) < tol)
    east = (np.abs(a[:, 0] - xmax) < tol) & (np.abs(b[:, 0] - xmax) < tol)
    south = (np.abs(a[:, 1] - ymin) < tol) & (np.abs(b[:, 1] - ymin) < tol)
    qb = np.where(north[:, None], np.stack([p[:, 0], np.full(nC, ymax + dx)], 1),
                  np.where(east[:, None], np.stack([np.full(nC, xmax + dx), p[:, 1]], 1),
                           np.where(south[:, None], np.stack([p[:, 0], np.full(nC, ymin - dx)], 1),
                                    np.stack([np.full(nC, xmin - dx), p[:, 1]], 1))))
    q = np.where(interior[:, None], q, qb)
    Cw = np.hypot(p[:, 0] - q[:, 0], p[:, 1] - q[:, 1])
    return VertexGeometry(A, Cw, ETri)
