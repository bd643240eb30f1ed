"""Tetrahedral meshes with the reference's node and element numbering.

``MeshTet.init_tensor`` and ``MeshTet2.from_mesh`` produce arrays identical
to python/skfemntv/basis.py:599-616 and :644-658 (vectorised instead of
Python loops, so 64^3-96^3 Kuhn meshes build in seconds).
"""

from __future__ import annotations

import numpy as np

# Kuhn split of one cube, corners numbered x + 2y + 4z (basis.py:607-608).
KUHN_TETS = np.array([(0, 1, 3, 7), (0, 3, 2, 7), (0, 2, 6, 7), (0, 6, 4, 7), (0, 4, 5, 7), (0, 5, 1, 7)],
                     dtype=np.int64)
# Local faces of a tetrahedron (basis.py:45-46).
TET_FACES = ((0, 1, 2), (0, 1, 3), (0, 2, 3), (1, 2, 3))
# Local edges of a tetrahedron (basis.py:652, edge_topology.cpp:40-42).
TET_EDGES = np.array([(0, 1), (1, 2), (2, 0), (0, 3), (1, 3), (2, 3)], dtype=np.int64)


class MeshTet:
    """Linear tetrahedral mesh: p (3, nodes) float64, t (4, elements) int64."""

    def __init__(self, p=None, t=None):
        self.p = np.asarray(p if p is not None else [[0., 1., 0., 0.], [0., 0., 1., 0.], [0., 0., 0., 1.]],
                            dtype=np.float64)
        self.t = np.asarray(t if t is not None else [[0], [1], [2], [3]], dtype=np.int64)
        if self.p.ndim != 2 or self.p.shape[0] != 3:
            raise ValueError("p must have shape (3, nodes)")
        if self.t.ndim != 2 or self.t.shape[0] != 4:
            raise ValueError("t must have shape (4, elements)")
        self._boundaries = {}

    @classmethod
    def init_tensor(cls, x, y, z):
        """Kuhn-split tensor grid; node id i + nx*(j + ny*k), x fastest."""
        x, y, z = (np.asarray(a, dtype=float) for a in (x, y, z))
        nx, ny, nz = len(x), len(y), len(z)
        K, J, I = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
        points = np.vstack((x[I.ravel()], y[J.ravel()], z[K.ravel()]))
        ci, cj, ck = np.meshgrid(np.arange(nx - 1), np.arange(ny - 1), np.arange(nz - 1), indexing="ij")
        # cube order: k outer, j, i inner
        ci, cj, ck = (a.transpose(2, 1, 0).ravel() for a in (ci, cj, ck))
        corners = np.empty((8, ci.size), dtype=np.int64)
        for bit in range(8):
            di, dj, dk = bit & 1, (bit >> 1) & 1, (bit >> 2) & 1
            corners[bit] = (ci + di) + nx * ((cj + dj) + ny * (ck + dk))
        # (4, cubes, 6) -> tets numbered cube*6 + m
        t = corners[KUHN_TETS.T]  # (4, 6, cubes)
        t = t.transpose(0, 2, 1).reshape(4, -1)
        return cls(points, np.ascontiguousarray(t))

    @property
    def nelements(self):
        return self.t.shape[1]

    @property
    def nvertices(self):
        return self.p.shape[1]

    def dim(self):
        return 3

    @property
    def boundaries(self):
        return dict(getattr(self, "_boundaries", {}))

    # ---------------------------------------------------------- facets
    def _build_topology(self):
        """Facets as the reference numbers them (basis.py:96-148): the local
        faces (0,1,2),(0,1,3),(0,2,3),(1,2,3) of every element, face-major,
        sorted per column; facets are the unique sorted triples in
        lexicographic order; f2t holds the first and second element met in
        that face-major order (-1 on the boundary)."""
        if hasattr(self, "_facets"):
            return
        t4 = np.asarray(self.t[:4], dtype=np.int64)
        E = t4.shape[1]
        indexing = np.hstack(tuple(t4[np.asarray(face)] for face in TET_FACES))  # (3, 4E), face-major
        canonical, inverse = np.unique(np.sort(indexing, axis=0), axis=1, return_inverse=True)
        inverse = np.asarray(inverse).ravel()
        F = canonical.shape[1]
        elements = np.tile(np.arange(E, dtype=np.int64), 4)
        order = np.argsort(inverse, kind="stable")          # occurrences of each facet in face-major order
        counts = np.bincount(inverse, minlength=F)
        if counts.max(initial=0) > 2:
            raise ValueError("non-manifold facet has more than two cells")
        start = np.concatenate([[0], np.cumsum(counts)[:-1]])
        f2t = np.full((2, F), -1, dtype=np.int64)
        f2t[0] = elements[order[start]]
        two = counts == 2
        f2t[1, two] = elements[order[start[two] + 1]]
        self._facets = np.ascontiguousarray(canonical, dtype=np.int64)
        self._t2f = np.ascontiguousarray(inverse.reshape(4, E), dtype=np.int64)
        self._f2t = f2t

    @property
    def facets(self):
        self._build_topology()
        return self._facets

    @property
    def t2f(self):
        self._build_topology()
        return self._t2f

    @property
    def f2t(self):
        self._build_topology()
        return self._f2t

    def boundary_facets(self):
        return np.flatnonzero(self.f2t[1] == -1)

    def interior_facets(self):
        return np.flatnonzero(self.f2t[1] != -1)


def first_seen_edges(t4):
    """Number the edges of tetrahedra by first occurrence over
    (element, local edge).  Returns (lo, hi) per edge id and the (6, E) edge
    id of every local edge."""
    t4 = np.asarray(t4, dtype=np.int64)
    E = t4.shape[1]
    a = t4[TET_EDGES[:, 0]]  # (6,E)
    b = t4[TET_EDGES[:, 1]]
    lo = np.minimum(a, b).T.ravel()  # element-major item order e*6+l
    hi = np.maximum(a, b).T.ravel()
    nv = int(t4.max(initial=-1)) + 1
    keys = lo * max(nv, 1) + hi
    uniq, first, inverse = np.unique(keys, return_index=True, return_inverse=True)
    order = np.argsort(first, kind="stable")
    rank = np.empty_like(order)
    rank[order] = np.arange(order.size)
    ids = rank[inverse.ravel()].reshape(E, 6).T
    edge_lo = lo[first[order]]
    edge_hi = hi[first[order]]
    return edge_lo, edge_hi, np.ascontiguousarray(ids)


class MeshTet2(MeshTet):
    """Quadratic tetrahedral mesh: t (10, elements), rows 4..9 are the edge
    nodes of local edges (0,1),(1,2),(2,0),(0,3),(1,3),(2,3)."""

    @classmethod
    def from_mesh(cls, mesh: MeshTet):
        p = np.asarray(mesh.p, dtype=np.float64)
        t4 = np.asarray(mesh.t[:4], dtype=np.int64)
        lo, hi, ids = first_seen_edges(t4)
        V = p.shape[1]
        mid = 0.5 * (p[:, lo] + p[:, hi])
        points = np.hstack((p, mid))
        t = np.vstack((t4, V + ids))
        return cls(points, t)

    def __init__(self, p=None, t=None):
        if p is None or t is None:
            generated = type(self).from_mesh(MeshTet())
            self.p, self.t = generated.p, generated.t
            self._boundaries = {}
            return
        self.p = np.asarray(p, dtype=np.float64)
        self.t = np.asarray(t, dtype=np.int64)
        if self.p.ndim != 2 or self.p.shape[0] != 3 or self.t.ndim != 2 or self.t.shape[0] != 10:
            raise ValueError("quadratic tetra mesh requires p (3,n) and t (10,e)")
        self._boundaries = {}


__all__ = ["MeshTet", "MeshTet2", "first_seen_edges", "KUHN_TETS", "TET_EDGES", "TET_FACES"]
