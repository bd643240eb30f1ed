"""End-to-end CPU oracle for each assembled system -- TEST INFRASTRUCTURE ONLY.

Composes the numpy restatement (``oracle.host``) and the C restatement
(``oracle.liboracle``) in the order the reference runs them:
Basis geometry (basis.py:2194-2294) -> pattern + map (assembler constructor)
-> element kernel + scatter (assembler.assemble).
"""

from __future__ import annotations

import numpy as np

from . import host
import oracle as _o


def basis_geometry(p, t, X, W, quadratic):
    """Tabulated shape/gradients/dx/x of a tetrahedral P1/P2 basis."""
    shape, grad = host.tabulate_p2(X) if quadratic else host.tabulate_p1(X)
    gshape, ggrad = host.tabulate_p2(X) if t.shape[0] == 10 else host.tabulate_p1(X)
    out, bad = _o.tabulate_geometry(p, t, shape, grad, gshape, ggrad, W)
    return out, bad


def dofs_of(t, npoints, nodes, components):
    N, edofs = host.element_dofs(t[:nodes], npoints, components)
    E = t.shape[1]
    return N, edofs.T.reshape(E, nodes, components)


def bilinear(p, t, X, W, quadratic, components=1, value=None, gradient=None, symgrad=None, divergence=None,
             threads=1):
    """Scalar-kind bilinear assembly (value/gradient/symgrad/div), coefficients (E,Q)."""
    (shape, grads, dx, x, det, tol, cond), bad = basis_geometry(p, t, X, W, quadratic)
    nodes = 10 if quadratic else 4
    N, dofs = dofs_of(t, p.shape[1], nodes, components)
    E = t.shape[1]
    ip, ix, sc = _o.pattern(dofs.reshape(E, -1), nrows=N)
    colour = None
    if threads > 1:
        colour, _ = _o.colouring(dofs.reshape(E, -1), N)
    vals = _o.assemble_bilinear(shape, grads, dx, sc, len(ix), components, value=value, gradient=gradient,
                                symgrad=symgrad, divergence=divergence, threads=threads, colour=colour)
    return ip, ix, vals, N


def tensor(p, t, X, W, quadratic, A_eq, threads=1):
    """Anisotropic gradient tensor assembly, A_eq (E,Q,3,3)."""
    (shape, grads, dx, x, det, tol, cond), bad = basis_geometry(p, t, X, W, quadratic)
    nodes = 10 if quadratic else 4
    N, dofs = dofs_of(t, p.shape[1], nodes, 1)
    E = t.shape[1]
    ip, ix, sc = _o.pattern(dofs.reshape(E, -1), nrows=N)
    colour = None
    if threads > 1:
        colour, _ = _o.colouring(dofs.reshape(E, -1), N)
    vals = _o.assemble_tensor(grads, dx, A_eq, sc, len(ix), threads=threads, colour=colour)
    return ip, ix, vals, N


def cross(p, t, X, W, row, col, row_kind, col_kind, coefficient=None, tensor=False):
    """Cross block between two fields on one mesh: row = test, col = trial,
    each (quadratic, components).  Returns (indptr, indices, values, rows,
    columns) in the reference's rectangular layout."""
    (rq, Cr), (cq, Cc) = row, col
    (rshape, rgrads, dx, x, det, tol, cond), _ = basis_geometry(p, t, X, W, rq)
    (cshape, cgrads, _, _, _, _, _), _ = basis_geometry(p, t, X, W, cq)
    nr, nc = (10 if rq else 4), (10 if cq else 4)
    Nr, rdofs = dofs_of(t, p.shape[1], nr, Cr)
    Nc, cdofs = dofs_of(t, p.shape[1], nc, Cc)
    E = t.shape[1]
    ip, ix, sc = _o.pattern(rdofs.reshape(E, -1), cdofs.reshape(E, -1), nrows=Nr)
    vals = _o.assemble_cross(rshape, rgrads, cshape, cgrads, dx, sc, len(ix), Cr, Cc, row_kind, col_kind,
                             coefficient=coefficient, tensor=tensor)
    return ip, ix, vals, Nr, Nc


def composite_block(p, t, X, W, fields, row_field, col_field, kind, coefficient=None):
    """One block of a composite basis as the reference assembles it
    (bilinear_form.py:370-455): the cross kernel on the composite DOFs of the
    two fields, divergence kinds through the (Cr,[d],Cc,[d]) unit tensor
    times the scalar.  fields: [(quadratic, components), ...].  Returns
    (indptr, indices, values, N) with N x N shape."""
    conns = [t[:(10 if q else 4)] for q, _ in fields]
    N, dofs = host.composite_dofs(p.shape[1], conns, [C for _, C in fields])
    (rq, Cr), (cq, Cc) = fields[row_field], fields[col_field]
    (rs, rg, dx, *_), _ = basis_geometry(p, t, X, W, rq)
    (cs, cg, *_), _ = basis_geometry(p, t, X, W, cq)
    E, Q = dx.shape
    scalar = np.broadcast_to(np.asarray(1. if coefficient is None else coefficient, dtype=float), (E, Q))
    if kind in ("row_divergence", "column_divergence"):
        if kind == "row_divergence":
            unit = np.zeros((Cr, 3, Cc))
            for c in range(Cr):
                unit[c, c, 0] = 1.
            rk, ck = "gradient", "value"
        else:
            unit = np.zeros((Cr, Cc, 3))
            for c in range(Cc):
                unit[0, c, c] = 1.
            rk, ck = "value", "gradient"
        coef, tensor = scalar[(...,) + (None,) * 3] * unit, True
    else:
        coef, tensor, rk, ck = np.ascontiguousarray(scalar), False, kind, kind
    rows = dofs[row_field].reshape(E, -1)
    cols = dofs[col_field].reshape(E, -1)
    ip, ix, sc = _o.pattern(rows, cols, nrows=N)
    vals = _o.assemble_cross(rs, rg, cs, cg, dx, sc, len(ix), Cr, Cc, rk, ck, coefficient=coef, tensor=tensor)
    return ip, ix, vals, N


def load_vector(p, t, X, W, quadratic, f_const, components=1):
    (shape, grads, dx, x, det, tol, cond), bad = basis_geometry(p, t, X, W, quadratic)
    nodes = 10 if quadratic else 4
    N, dofs = dofs_of(t, p.shape[1], nodes, components)
    return _o.assemble_linear(shape, grads, dx, dofs, N, components, value=np.asarray(f_const, dtype=float))


def n1_maxwell(p, t, kappa_eq, mu_eq, intorder=2, include_mass=True, include_curl=True):
    """TetN1 maxwell assembly: edge topology + pattern + integrate + scatter."""
    edges, ee, signs = _o.edge_topology(t[:4])
    X, W = host.tet_rule_duffy(intorder)
    dx, values, curls, det = host.n1_basis(p, t, X, W, signs)
    ip, ix, sc = _o.pattern(ee.T.copy(), nrows=edges.shape[1])
    el = _o.integrate_tet_n1(values, curls, dx, kappa_eq, mu_eq, include_mass, include_curl)
    data = _o.scatter_add(sc, el, len(ix))
    return ip, ix, data, edges, ee, signs, (X, W)


def relative_frobenius(a, b):
    a = np.asarray(a, dtype=float)
    b = np.asarray(b, dtype=float)
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.))
