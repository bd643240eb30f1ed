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
