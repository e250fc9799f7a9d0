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
