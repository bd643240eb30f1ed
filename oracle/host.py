"""Host-side CPU oracle (numpy) -- TEST INFRASTRUCTURE ONLY.

Restates, in plain numpy/Python, the parts of skfemntv 0.3.1 that run in
Python rather than in its C++ extension: Kuhn tetrahedral meshes, the P2
edge-node numbering, DOF numbering, quadrature rules, reference
tabulation, the affine Nedelec-N1 mapping and the coefficient layout rules.
Only tests/, the smoke check and the bench's CPU-baseline leg import this
module.  Each function cites the reference code (paths relative to the
reference repository root).
"""

from __future__ import annotations

import numpy as np

# Local edge order of a tetrahedron (python/skfemntv/basis.py:652,
# cpp/src/assembly/edge_topology.cpp:40-42).
TET_EDGES = ((0, 1), (1, 2), (2, 0), (0, 3), (1, 3), (2, 3))
# Kuhn split of a cube with corners numbered x + 2y + 4z (basis.py:607-608).
KUHN_TETS = ((0, 1, 3, 7), (0, 3, 2, 7), (0, 2, 6, 7), (0, 6, 4, 7), (0, 4, 5, 7), (0, 5, 1, 7))


def kuhn_mesh(x, y, z):
    """MeshTet.init_tensor (basis.py:599-616) with explicit Python loops."""
    x, y, z = map(np.asarray, (x, y, z))
    points = np.array([[a, b, c] for c in z for b in y for a in x], dtype=float).T
    nx, ny = len(x), len(y)

    def node(i, j, k):
        return i + nx * (j + ny * k)

    cells = []
    for k in range(len(z) - 1):
        for j in range(len(y) - 1):
            for i in range(len(x) - 1):
                cube = (node(i, j, k), node(i + 1, j, k), node(i, j + 1, k), node(i + 1, j + 1, k),
                        node(i, j, k + 1), node(i + 1, j, k + 1), node(i, j + 1, k + 1),
                        node(i + 1, j + 1, k + 1))
                cells.extend(tuple(cube[q] for q in tet) for tet in KUHN_TETS)
    return points, np.asarray(cells, dtype=np.int64).reshape(-1, 4).T.copy()


def p2_mesh(p, t):
    """MeshTet2.from_mesh (basis.py:644-658): first-seen edge nodes."""
    points = [p[:, i].copy() for i in range(p.shape[1])]
    found = {}
    rows = [list(map(int, t[i])) for i in range(4)]
    edge_rows = [[] for _ in TET_EDGES]
    for vertices in t.T:
        for row, (a, b) in zip(edge_rows, TET_EDGES):
            edge = tuple(sorted((int(vertices[a]), int(vertices[b]))))
            if edge not in found:
                found[edge] = len(points)
                points.append(0.5 * (p[:, edge[0]] + p[:, edge[1]]))
            row.append(found[edge])
    return np.asarray(points).T.copy(), np.asarray(rows + edge_rows, dtype=np.int64)


def jitter_interior(p, h, amount, seed):
    """Build-only synthetic input: move interior vertices of the unit cube by
    U(-amount*h, amount*h) per coordinate (no reference counterpart)."""
    rng = np.random.default_rng(seed)
    q = p.copy()
    interior = np.all((p > 1e-12) & (p < 1 - 1e-12), axis=0)
    q[:, interior] += rng.uniform(-amount * h, amount * h, size=(3, int(interior.sum())))
    return q


def element_dofs(conn, npoints, components):
    """DOF numbering for continuous nodal elements (basis.py:1749-1777).

    Returns (N, element_dofs (n*C, E))."""
    used = np.zeros(npoints, dtype=bool)
    used[conn] = True
    active = np.flatnonzero(used)
    N = len(active) * components
    nodal = np.arange(N).reshape(-1, components).T
    pos = np.full(npoints, -1, dtype=np.int64)
    pos[active] = np.arange(len(active))
    local = pos[conn]
    E = conn.shape[1]
    return N, nodal[:, local].transpose(2, 1, 0).reshape(E, conn.shape[0] * components).T.copy()


def tet_rule_4():
    """Default P1 tetrahedral rule (basis.py:1728-1734)."""
    a, b = .5854101966249685, .1381966011250105
    X = np.array([[a, b, b, b], [b, a, b, b], [b, b, a, b]])
    return X, np.full(4, 1. / 24.)


def tet_rule_keast11():
    """Default P2 tetrahedral rule, 11-point Keast (basis.py:1716-1727)."""
    X = np.array(
        [[.25, .7857142857142857, .0714285714285714, .0714285714285714, .0714285714285714,
          .1005964238332008, .3994035761667992, .3994035761667992, .3994035761667992,
          .1005964238332008, .1005964238332008],
         [.25, .0714285714285714, .0714285714285714, .0714285714285714, .7857142857142857,
          .3994035761667992, .1005964238332008, .3994035761667992, .1005964238332008,
          .3994035761667992, .1005964238332008],
         [.25, .0714285714285714, .0714285714285714, .7857142857142857, .0714285714285714,
          .3994035761667992, .3994035761667992, .1005964238332008, .1005964238332008,
          .1005964238332008, .3994035761667992]])
    W = np.array([-.01315555555555555] + [.00762222222222222] * 4 + [.02488888888888888] * 6)
    return X, W


def tet_rule_duffy(order):
    """Duffy product rule on the unit tetrahedron (hcurl_basis.py:37-56)."""
    count = max(1, (int(order) + 4) // 2)
    pts, wts = np.polynomial.legendre.leggauss(count)
    pts = (pts + 1.) / 2.
    wts = wts / 2.
    X, W = [], []
    for i in range(count):
        for j in range(count):
            for k in range(count):
                x = pts[i]
                y = (1. - x) * pts[j]
                z = (1. - x) * (1. - pts[j]) * pts[k]
                X.append((x, y, z))
                W.append((1. - x) ** 2 * (1. - pts[j]) * wts[i] * wts[j] * wts[k])
    return np.asarray(X).T, np.asarray(W)


def tabulate_p1(X):
    """Reference P1 values/gradients (basis.py:2236-2239)."""
    shape = np.vstack((1. - X.sum(axis=0), X)).T
    ref = np.array([[-1., -1., -1.], [1., 0., 0.], [0., 1., 0.], [0., 0., 1.]])
    return shape, np.broadcast_to(ref, (X.shape[1], 4, 3)).copy()


def tabulate_p2(X):
    """Reference P2 values/gradients with the (0,1),(1,2),(2,0),(0,3),(1,3),(2,3)
    edge order (basis.py:2224-2235)."""
    bary = np.vstack((1. - X.sum(axis=0), X)).T
    dl = np.array([[-1., -1., -1.], [1., 0., 0.], [0., 1., 0.], [0., 0., 1.]])
    nq = X.shape[1]
    shape = np.empty((nq, 10))
    grad = np.empty((nq, 10, 3))
    for q, L in enumerate(bary):
        for i in range(4):
            shape[q, i] = L[i] * (2. * L[i] - 1.)
            grad[q, i] = (4. * L[i] - 1.) * dl[i]
        for k, (i, j) in enumerate(TET_EDGES, start=4):
            shape[q, k] = 4. * L[i] * L[j]
            grad[q, k] = 4. * (L[j] * dl[i] + L[i] * dl[j])
    return shape, grad


def anisotropic_tensor(raw, coefficient_shape, dim=3):
    """_anisotropic_tensor_coefficient (forms.py:377-404), returning the
    entity-first (E,Q,dim,dim) array."""
    tensor = np.asarray(raw, dtype=np.float64)
    if tensor.shape == (dim, dim):
        tensor = np.broadcast_to(tensor, tuple(coefficient_shape) + (dim, dim))
    elif tensor.shape == (dim, dim) + tuple(coefficient_shape):
        tensor = np.moveaxis(tensor, (0, 1), (-2, -1))
    elif tensor.shape == tuple(coefficient_shape) + (dim, dim):
        pass
    else:
        raise ValueError("anisotropic tensor must use the recommended scikit-fem shape")
    return np.ascontiguousarray(tensor)


# ---------------------------------------------------------------- Nedelec N1
_DL = np.array(((-1., -1., -1.), (1., 0., 0.), (0., 1., 0.), (0., 0., 1.)))


def n1_reference(X):
    """tet_n1_basis / tet_n1_curl (_nedelec_reference.py:72-107):
    values (6,3,Q) and constant curls (6,3)."""
    lam = np.vstack((1. - X.sum(axis=0), X))
    values = np.empty((6, 3, X.shape[1]))
    curls = np.empty((6, 3))
    for k, (i, j) in enumerate(TET_EDGES):
        values[k] = lam[i][None, :] * _DL[j][:, None] - lam[j][None, :] * _DL[i][:, None]
        curls[k] = 2. * np.cross(_DL[i], _DL[j])
    return values, curls


def n1_basis(p, t, X, W, signs):
    """AffineTetN1Basis mapping (hcurl_basis.py:297-364, _hcurl_mapping.py:21-97).

    Returns dx (E,Q), values (E,6,3,Q) and curls (E,6,3,Q) with the edge signs
    applied; per-element loops with numpy.linalg.inv as in the reference."""
    v = p[:, t[:4]]  # (3,4,E)
    J = np.stack([v[:, k] - v[:, 0] for k in (1, 2, 3)], axis=1)  # (3,3,E)
    Jm = np.moveaxis(J, -1, 0)  # (E,3,3)
    det = np.linalg.det(Jm)
    inv = np.linalg.inv(Jm)  # (E,3,3)
    rv, rc = n1_reference(X)
    values = np.einsum("erp,brq->ebpq", inv, rv)  # J^{-T} phi: sum_r inv[r,p] phi_r
    curls = np.einsum("epr,br->ebp", Jm, rc)[..., None] / det[:, None, None, None]
    curls = np.broadcast_to(curls, values.shape)
    s = signs.T[:, :, None, None].astype(np.float64)
    dx = np.abs(det)[:, None] * W[None, :]
    return dx, np.ascontiguousarray(values * s), np.ascontiguousarray(curls * s), det


def composite_dofs(npoints, conns, components):
    """Composite-basis numbering (basis.py:2102-2127), restated as the
    reference's loop: nodes in increasing id, at each node the fields that use
    it in order, each with its consecutive components.  conns: per field the
    (n_f, E) node connectivity.  Returns (N, [element dofs (E, n_f, C_f)])."""
    active = [set(np.unique(c).tolist()) for c in conns]
    field_node = [dict() for _ in conns]
    nxt = 0
    for node in range(npoints):
        for f, (act, C) in enumerate(zip(active, components)):
            if node in act:
                field_node[f][node] = nxt
                nxt += C
    out = []
    for f, (c, C) in enumerate(zip(conns, components)):
        first = np.vectorize(lambda v: field_node[f][int(v)])(c.T)  # (E, n_f)
        out.append(first[:, :, None] + np.arange(C)[None, None, :])
    return nxt, out


TET_FACES = ((0, 1, 2), (0, 1, 3), (0, 2, 3), (1, 2, 3))


def facet_topology(t):
    """Facets, t2f, f2t restated as the reference builds them
    (basis.py:96-148): face-major indexing, unique sorted triples, f2t filled
    by a loop over (local face, element)."""
    t4 = np.asarray(t[:4], dtype=np.int64)
    E = t4.shape[1]
    indexing = np.hstack(tuple(t4[np.asarray(face)] for face in TET_FACES))
    canonical, first, inverse = np.unique(np.sort(indexing, axis=0), axis=1, return_index=True,
                                          return_inverse=True)
    t2f = np.asarray(inverse).reshape(4, E)
    f2t = np.full((2, canonical.shape[1]), -1, dtype=np.int64)
    for facet, element in zip(t2f.ravel(order="C"), np.tile(np.arange(E), 4)):
        row = 0 if f2t[0, facet] == -1 else 1
        f2t[row, facet] = element
    return canonical, t2f, f2t


def facet_basis(p, t, quadratic, components=1, intorder=2, side=0, interior=False):
    """Tetrahedral FacetBasis on the boundary facets (or, with interior=True,
    on one side of the interior facets: InteriorFacetBasis), restated per
    facet and point as the reference computes it (basis.py:2575-2685): the
    rule is placed by the parent's local corners of the facet's corners in
    the facet's own order, and side 1 of an interior facet flips its
    normal.  Returns a dict with facets, parents, local faces, shape
    (F,Q,n), dx (F,Q), x (F,Q,3), normals (F,Q,3), dofs (F,n,C) and N."""
    facets, t2f, f2t = facet_topology(t)
    boundary = np.flatnonzero(f2t[1] >= 0) if interior else np.flatnonzero(f2t[1] == -1)
    if intorder >= 4:
        a, b = .445948490915965, .091576213509771
        bary = np.array([[a, a, 1 - 2 * a], [a, 1 - 2 * a, a], [1 - 2 * a, a, a],
                         [b, b, 1 - 2 * b], [b, 1 - 2 * b, b], [1 - 2 * b, b, b]])
        wts = np.array([.1116907948390055] * 3 + [.054975871827661] * 3)
    else:
        bary = np.array([[2 / 3, 1 / 6, 1 / 6], [1 / 6, 2 / 3, 1 / 6], [1 / 6, 1 / 6, 2 / 3]])
        wts = np.full(3, 1 / 6)
    pts = bary[:, 1:]
    V = np.array([[0., 0., 0.], [1., 0., 0.], [0., 1., 0.], [0., 0., 1.]])
    nodes = 10 if quadratic else 4
    N, edofs = element_dofs(t[:nodes], p.shape[1], components)
    E = t.shape[1]
    dofs_e = edofs.T.reshape(E, nodes, components)
    F, Q = boundary.size, len(wts)
    out = {"shape": np.zeros((F, Q, nodes)), "dx": np.zeros((F, Q)), "x": np.zeros((F, Q, 3)),
           "normals": np.zeros((F, Q, 3)), "parents": np.zeros(F, dtype=np.int64), "local": np.zeros(F, dtype=int)}
    for f, facet in enumerate(boundary):
        e = int(f2t[side, facet])
        li = int(np.flatnonzero(t2f[:, e] == facet)[0])
        corner = [int(np.flatnonzero(t[:4, e] == node)[0]) for node in facets[:, facet]]
        rc = V[corner]
        x = p[:, t[:, e]]
        for q, (r, s) in enumerate(pts):
            ref = (1 - r - s) * rc[0] + r * rc[1] + s * rc[2]
            der = np.stack((rc[1] - rc[0], rc[2] - rc[0]), axis=1)
            sh, _ = tabulate_p2(ref[:, None]) if quadratic else tabulate_p1(ref[:, None])
            gsh, ggr = tabulate_p2(ref[:, None]) if t.shape[0] == 10 else tabulate_p1(ref[:, None])
            jac = x @ ggr[0]
            tang = jac @ der
            point = gsh[0] @ x.T
            nrm = np.cross(tang[:, 0], tang[:, 1])
            nrm /= np.linalg.norm(nrm)
            if np.dot(nrm, point - x.mean(axis=1)) < 0.:
                nrm = -nrm
            if interior and side == 1:
                nrm = -nrm
            out["shape"][f, q] = sh[0]
            out["x"][f, q] = point
            out["normals"][f, q] = nrm
            out["dx"][f, q] = np.linalg.norm(np.cross(tang[:, 0], tang[:, 1])) * wts[q]
        out["parents"][f] = e
        out["local"][f] = li
    out["dofs"] = dofs_e[out["parents"]]
    out["N"] = N
    out["facets"] = boundary
    return out
