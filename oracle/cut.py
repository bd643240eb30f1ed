"""Cut-cell (ragged quadrature) assembly oracle (numpy) -- TEST INFRASTRUCTURE ONLY.

Restates the reference's cut-cell assemblers in plain numpy: the pattern,
the load vector, the square value/gradient form and the cross form with a
scalar or tensor coefficient.  Only tests/ import it.  Paths are relative
to the reference repository root.

Cell c owns the flat points [offsets[c], offsets[c+1]); cells without
points stay out of the pattern (cpp/src/assembly/cut_form_assembler.cpp:208-220,
cpp/src/assembly/cut_cross_assembler.cpp:148-153).
"""

from __future__ import annotations

import numpy as np


def cut_pattern(row_dofs, col_dofs, offsets, nrows=None):
    """Sorted unique columns per row over every (row DOF, column DOF) pair of
    every cell that has points (cut_form_assembler.cpp:208-220,
    cut_cross_assembler.cpp:148-153).  rows = max row DOF + 1 over all cells."""
    row_dofs = np.asarray(row_dofs, dtype=np.int64)
    col_dofs = np.asarray(col_dofs, dtype=np.int64)
    cells = row_dofs.shape[0]
    if nrows is None:
        nrows = int(row_dofs.max()) + 1 if row_dofs.size else 0
    active = np.diff(np.asarray(offsets)) > 0
    r = row_dofs.reshape(cells, -1)[active]
    c = col_dofs.reshape(cells, -1)[active]
    rows = np.repeat(r, c.shape[1], axis=1).ravel()
    cols = np.tile(c, (1, r.shape[1])).ravel()
    ncols = int(col_dofs.max()) + 1 if col_dofs.size else 1
    keys = np.unique(rows * max(ncols, 1) + cols)
    kr, kc = keys // max(ncols, 1), keys % max(ncols, 1)
    indptr = np.zeros(nrows + 1, dtype=np.int64)
    np.add.at(indptr, kr + 1, 1)
    return np.cumsum(indptr), kc.astype(np.int64)


def _position(indptr, indices, row, col):
    lo, hi = indptr[row], indptr[row + 1]
    return lo + int(np.searchsorted(indices[lo:hi], col))


def cut_linear(dofs, offsets, shape, gradients, weights, value=None, gradient=None):
    """CutLinearFormAssembler::assemble (cut_form_assembler.cpp:67-134):
    out[dofs[c,a,k]] += w_q (phi_a(q) f_k + grad phi_a(q) . G_k), with f of
    shape (C,) or (P, C) and G of shape (C, D) or (P, C, D)."""
    dofs = np.asarray(dofs, dtype=np.int64)
    cells, nodes, C = dofs.shape
    ndofs = int(dofs.max()) + 1 if dofs.size else 0
    out = np.zeros(ndofs)
    for cell in range(cells):
        for q in range(offsets[cell], offsets[cell + 1]):
            for a in range(nodes):
                for k in range(C):
                    s = 0.
                    if value is not None:
                        f = value[k] if np.ndim(value) == 1 else value[q, k]
                        s += shape[q, a] * f
                    if gradient is not None:
                        G = gradient[k] if np.ndim(gradient) == 2 else gradient[q, k]
                        s += float(np.dot(gradients[q, a], G))
                    out[dofs[cell, a, k]] += weights[q] * s
    return out


def cut_bilinear(dofs, offsets, shape, gradients, weights, value=None, gradient=None):
    """CutBilinearFormAssembler (cut_form_assembler.cpp:136-239): entry
    (a,k),(b,k) += w_q (v_q phi_a phi_b + g_q grad phi_a . grad phi_b) over
    the cell's points; returns (indptr, indices, values)."""
    dofs = np.asarray(dofs, dtype=np.int64)
    cells, nodes, C = dofs.shape
    indptr, indices = cut_pattern(dofs, dofs, offsets)
    values = np.zeros(len(indices))
    for cell in range(cells):
        for q in range(offsets[cell], offsets[cell + 1]):
            for a in range(nodes):
                for b in range(nodes):
                    e = 0.
                    if value is not None:
                        e += value[q] * shape[q, a] * shape[q, b]
                    if gradient is not None:
                        e += gradient[q] * float(np.dot(gradients[q, a], gradients[q, b]))
                    e *= weights[q]
                    for k in range(C):
                        values[_position(indptr, indices, dofs[cell, a, k], dofs[cell, b, k])] += e
    return indptr, indices, values


def cut_cross(row_dofs, col_dofs, offsets, row_shape, col_shape, weights, row_gradients, col_gradients,
              coefficient=None, row_kind="value", column_kind="value"):
    """CutCrossAssembler::assemble (cut_cross_assembler.cpp:70-137) with the
    contractions of cpp/include/native_fem/cross_contraction.hpp:42-97: a
    scalar coefficient (P,) (or None = 1) couples equal components through
    phi_a psi_b or grad phi_a . grad psi_b; a tensor (P, Cr, [Dr], Cc, [Dc])
    couples every component pair."""
    row_dofs = np.asarray(row_dofs, dtype=np.int64)
    col_dofs = np.asarray(col_dofs, dtype=np.int64)
    cells, nr, Cr = row_dofs.shape
    _, nc, Cc = col_dofs.shape
    indptr, indices = cut_pattern(row_dofs, col_dofs, offsets)
    values = np.zeros(len(indices))
    rg, cg = row_kind == "gradient", column_kind == "gradient"
    tensor = coefficient is not None and np.ndim(coefficient) == 3 + int(rg) + int(cg)

    def h(table_v, table_g, grad, q, a):
        return table_g[q, a] if grad else np.array([table_v[q, a]])

    for cell in range(cells):
        for q in range(offsets[cell], offsets[cell + 1]):
            for a in range(nr):
                ha = h(row_shape, row_gradients, rg, q, a)
                for b in range(nc):
                    hb = h(col_shape, col_gradients, cg, q, b)
                    if not tensor:
                        c = 1. if coefficient is None else coefficient[q]
                        e = weights[q] * c * float(np.dot(ha, hb))
                        for k in range(Cr):
                            values[_position(indptr, indices, row_dofs[cell, a, k], col_dofs[cell, b, k])] += e
                    else:
                        T = np.asarray(coefficient[q]).reshape(Cr, len(ha), Cc, len(hb))
                        for r in range(Cr):
                            for k in range(Cc):
                                e = float(ha @ T[r, :, k, :] @ hb)
                                values[_position(indptr, indices, row_dofs[cell, a, r],
                                                 col_dofs[cell, b, k])] += weights[q] * e
    return indptr, indices, values


def random_case(quadratic, components, seed, n=3, dim=3, empty_fraction=0.4, max_points=6):
    """Synthetic cut data on a Kuhn cube of n^3 cells: scikit-fem DOFs of a
    P1 or P2 field, a random subset of cells without points, 1..max_points
    points in the others, and random per-point shape values, gradients and
    weights (the assemblers treat the tables as data).  Returns
    (dofs, offsets, shape, gradients, weights)."""
    from . import host
    from .pipeline import dofs_of
    ax = np.linspace(0., 1., n + 1)
    p, t = host.kuhn_mesh(ax, ax, ax)
    if quadratic:
        p, t = host.p2_mesh(p, t)
    nodes = 10 if quadratic else 4
    _, dofs = dofs_of(t, p.shape[1], nodes, components)
    rng = np.random.default_rng(seed)
    E = dofs.shape[0]
    counts = rng.integers(1, max_points + 1, E)
    counts[rng.random(E) < empty_fraction] = 0
    offsets = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
    P = int(offsets[-1])
    shape = rng.standard_normal((P, nodes))
    gradients = rng.standard_normal((P, nodes, dim))
    weights = rng.uniform(0.05, 1., P)
    return dofs, offsets, shape, gradients, weights


def _cell_sums(per_point, offsets, cells):
    """Sum per-point arrays over each cell's point range, in point order per
    cell (cells without points give zeros)."""
    offsets = np.asarray(offsets, dtype=np.int64)
    active = np.diff(offsets) > 0
    out = np.zeros((cells,) + per_point.shape[1:])
    if active.any():
        out[active] = np.add.reduceat(per_point, offsets[:-1][active], axis=0)
    return out, active


def cut_bilinear_vectorized(dofs, offsets, shape, gradients, weights, value=None, gradient=None):
    """cut_bilinear with numpy array operations, for meshes where the loops
    are too slow: the per-point entries w_q (v_q phi_a phi_b + g_q grad phi_a
    . grad phi_b) are summed per cell, then each cell's node matrix is added
    at its CSR positions (component-diagonal pairs).  Summation order differs
    from cut_bilinear by rounding only (pinned against it on small cases)."""
    dofs = np.asarray(dofs, dtype=np.int64)
    cells, nodes, C = dofs.shape
    indptr, indices = cut_pattern(dofs, dofs, offsets)
    values = np.zeros(len(indices))
    P = len(weights)
    e = np.zeros((P, nodes, nodes))
    if value is not None:
        e += np.asarray(value)[:, None, None] * shape[:, :, None] * shape[:, None, :]
    if gradient is not None:
        e += np.asarray(gradient)[:, None, None] * np.einsum("pad,pbd->pab", gradients, gradients)
    e *= np.asarray(weights)[:, None, None]
    M, active = _cell_sums(e, offsets, cells)
    del e
    ncols = int(dofs.max()) + 1 if dofs.size else 1
    rows_of = np.repeat(np.arange(len(indptr) - 1, dtype=np.int64), np.diff(indptr))
    keys = rows_of * ncols + indices
    for k in range(C):
        d = dofs[active, :, k]
        key = d[:, :, None] * ncols + d[:, None, :]
        np.add.at(values, np.searchsorted(keys, key.ravel()), M[active].ravel())
    return indptr, indices, values


def cut_linear_vectorized(dofs, offsets, shape, gradients, weights, value=None, gradient=None):
    """cut_linear with numpy array operations (per-point contributions summed
    per cell, then added at the cells' DOFs); pinned against cut_linear on
    small cases."""
    dofs = np.asarray(dofs, dtype=np.int64)
    cells, nodes, C = dofs.shape
    ndofs = int(dofs.max()) + 1 if dofs.size else 0
    P = len(weights)
    s = np.zeros((P, nodes, C))
    if value is not None:
        f = np.broadcast_to(np.asarray(value), (P, C)) if np.ndim(value) == 1 else np.asarray(value)
        s += shape[:, :, None] * f[:, None, :]
    if gradient is not None:
        G = np.broadcast_to(np.asarray(gradient), (P,) + np.shape(gradient)) if np.ndim(gradient) == 2 \
            else np.asarray(gradient)
        s += np.einsum("pad,pkd->pak", gradients, G)
    s *= np.asarray(weights)[:, None, None]
    b, _ = _cell_sums(s, offsets, cells)
    out = np.zeros(ndofs)
    np.add.at(out, dofs.ravel(), b.ravel())
    return out
