"""Basis over tetrahedral P1/P2 (scalar or vector) fields on one HIP device.

Mirrors ``skfemntv.Basis`` (python/skfemntv/basis.py:1620-1802,2160-2356):
same quadrature selection, DOF numbering, geometry validation messages and
``GeometryDiagnostics``.  Host code builds the integer numbering and the
small reference tables; everything per element or per quadrature point is
computed on the GPU.  ``tabulated_shape``, ``tabulated_gradients``, ``dx``
and ``global_coordinates`` are computed on first access (on the GPU, then
copied to host) instead of eagerly, because at 10^6-10^7 elements they are
gigabytes that assembly never reads: the assembly kernels recompute the
geometry from vertex coordinates.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import _native
from .element import ElementComposite, ElementTetP1, ElementTetP2, ElementVector
from .mesh import MeshTet, MeshTet2

_EPS = np.finfo(np.float64).eps


@dataclass(frozen=True)
class GeometryDiagnostics:
    element_count: int
    quadrature_points_per_element: int
    minimum_determinant: float
    maximum_determinant: float
    minimum_scaled_determinant: float
    worst_element: int
    worst_quadrature_point: int
    determinant_tolerance: float
    negative_orientation_elements: int
    maximum_condition_number: float
    worst_condition_element: int
    worst_condition_quadrature_point: int


# ------------------------------------------------------------------ quadrature
def tet_rule_4():
    """4-point degree-2 rule (basis.py:1728-1734)."""
    a, b = .5854101966249685, .1381966011250105
    return np.array([[a, b, b, b], [b, a, b, b], [b, b, a, b]]), np.full(4, 1. / 24.)


def tet_rule_keast11():
    """11-point Keast rule (basis.py:1716-1727)."""
    s, t, u, v = .7857142857142857, .0714285714285714, .1005964238332008, .3994035761667992
    X = np.array([[.25, s, t, t, t, u, v, v, v, u, u],
                  [.25, t, t, t, s, v, u, v, u, v, u],
                  [.25, t, t, s, t, v, v, u, u, u, v]])
    W = np.array([-.01315555555555555] + [.00762222222222222] * 4 + [.02488888888888888] * 6)
    return X, W


def simplex_rule(intorder):
    """Collapsed Gauss rule on the unit tetrahedron (basis.py:1580-1598)."""
    count = max(1, (int(intorder) + 4) // 2)
    g, gw = np.polynomial.legendre.leggauss(count)
    g = (g + 1.) / 2.
    gw = gw / 2.
    X, W = [], []
    for index in np.ndindex(count, count, count):
        par = g[list(index)]
        coord = np.empty(3)
        remaining, jac = 1., 1.
        for axis, value in enumerate(par):
            coord[axis] = remaining * value
            if axis < 2:
                jac *= (1. - value) ** (3 - axis - 1)
                remaining *= 1. - value
        X.append(coord)
        W.append(jac * np.prod(gw[list(index)]))
    return np.asarray(X).T, np.asarray(W)


def validate_quadrature(quadrature, dimension=3):
    """basis.py:1547-1561."""
    points, weights = quadrature
    points = np.asarray(points, dtype=np.float64)
    weights = np.asarray(weights, dtype=np.float64)
    if points.ndim != 2 or points.shape[0] != dimension:
        raise ValueError(f"quadrature points must have shape ({dimension}, nq)")
    if weights.ndim != 1 or weights.shape[0] != points.shape[1]:
        raise ValueError("quadrature weights must have shape (nq,)")
    if not np.all(np.isfinite(points)) or not np.all(np.isfinite(weights)):
        raise ValueError("quadrature points and weights must be finite")
    return np.ascontiguousarray(points), np.ascontiguousarray(weights)


_DL = np.array([[-1., -1., -1.], [1., 0., 0.], [0., 1., 0.], [0., 0., 1.]])
_P2_PAIRS = ((0, 1), (1, 2), (2, 0), (0, 3), (1, 3), (2, 3))


def tabulate(X, quadratic):
    """Reference values (Q,n) and gradients (Q,n,3) of tetrahedral P1/P2
    (basis.py:2224-2239)."""
    X = np.asarray(X, dtype=np.float64)
    bary = np.vstack((1. - X.sum(axis=0), X)).T  # (Q,4)
    if not quadratic:
        return bary.copy(), np.broadcast_to(_DL, (X.shape[1], 4, 3)).copy()
    Q = X.shape[1]
    shape = np.empty((Q, 10))
    grad = np.empty((Q, 10, 3))
    for i in range(4):
        shape[:, i] = bary[:, i] * (2. * bary[:, i] - 1.)
        grad[:, i] = (4. * bary[:, i] - 1.)[:, None] * _DL[i]
    for k, (i, j) in enumerate(_P2_PAIRS, start=4):
        shape[:, k] = 4. * bary[:, i] * bary[:, j]
        grad[:, k] = 4. * (bary[:, j][:, None] * _DL[i] + bary[:, i][:, None] * _DL[j])
    return shape, grad


def pack_tables(W, shape, grad, gshape, ggrad):
    """FP64 table buffer consumed by the kernels:
    w[Q] | phi[Q][n] | dphi[Q][n][3] | gphi[Q][gn] | gdphi[Q][gn][3]."""
    return np.concatenate([np.ravel(W), np.ravel(shape), np.ravel(grad), np.ravel(gshape), np.ravel(ggrad)]).astype(
        np.float64)


def reduced_tables(X, W, quadratic):
    """Quadrature tables of the element-constant fast path.

    For affine elements with coefficients constant per element, the element
    matrix depends on the rule only through G = sum_q w_q dphi_a(q) dphi_b(q)^T
    and M = sum_q w_q phi_a(q) phi_b(q).  Any rule reproducing those sums gives
    the same matrix, so the gradient part uses one point when the reference
    gradients are constant (P1) or the 4-point rule when its sums equal the
    basis rule's, and the mass part likewise; otherwise the basis rule.
    Returns (tables wg|dg|wm|pm FP64, Qg, Qm)."""
    s, g = tabulate(X, quadratic)
    X4, W4 = tet_rule_4()
    s4, g4 = tabulate(X4, quadratic)

    def gsum(w, gg):
        return np.einsum("q,qai,qbj->aibj", w, gg, gg)

    def msum(w, ss):
        return np.einsum("q,qa,qb->ab", w, ss, ss)

    G = gsum(W, g)
    if not quadratic and np.array_equal(g, np.broadcast_to(g[:1], g.shape)):
        wg, dg, Xg = np.array([W.sum()]), g[:1], None
    elif np.allclose(gsum(W4, g4), G, rtol=1e-14, atol=1e-15 * np.abs(G).max()):
        wg, dg, Xg = W4, g4, X4
    else:
        wg, dg, Xg = W, g, X
    M = msum(W, s)
    if np.allclose(msum(W4, s4), M, rtol=1e-14, atol=1e-15 * np.abs(M).max()):
        wm, pm = W4, s4
    else:
        wm, pm = W, s
    tab = np.concatenate([wg, dg.ravel(), wm, pm.ravel()]).astype(np.float64)
    if quadratic:
        c = barycentric_coefficients(Xg, dg)
        tab = np.concatenate([tab, c.ravel()])
        # the pair table starts at a multiple of 4 entries (16-byte rows in
        # the FP32 copy, 32-byte in FP64)
        tab = np.concatenate([tab, np.zeros(-len(tab) % 4), pair_tables(wg, dg, wm, pm, c).ravel()])
    return tab, len(wg), len(wm)


def barycentric_coefficients(X, dg):
    """(16, Q) coefficients c_bm(q) of the P2 reference gradients on the
    barycentric gradients, grad phi_b(q) = sum_m c_bm(q) grad L_m: vertex k
    (pair k) 4 L_k - 1; edge e = (i, j) (pairs 4 + 2e, 5 + 2e) 4 L_j on
    grad L_i and 4 L_i on grad L_j.  Checked against the tabulated gradients."""
    X = np.asarray(X, dtype=np.float64)
    bary = np.vstack((1. - X.sum(axis=0), X)).T  # (Q,4)
    c = np.empty((16, X.shape[1]))
    for k in range(4):
        c[k] = 4. * bary[:, k] - 1.
    for e, (i, j) in enumerate(_P2_PAIRS):
        c[4 + 2 * e] = 4. * bary[:, j]
        c[5 + 2 * e] = 4. * bary[:, i]
    rec = np.empty_like(dg)
    for k in range(4):
        rec[:, k] = c[k][:, None] * _DL[k]
    for e, (i, j) in enumerate(_P2_PAIRS):
        rec[:, 4 + e] = c[4 + 2 * e][:, None] * _DL[i] + c[5 + 2 * e][:, None] * _DL[j]
    if not np.allclose(rec, dg, rtol=0., atol=1e-14 * max(1., np.abs(dg).max())):
        raise AssertionError("barycentric coefficients do not reproduce the P2 gradients")
    return c


def _pair_table_offset(nqg, nqm):
    """Offset of the pair table in the 10-node reduced tables: after
    wg | dg | wm | pm | c, rounded up to a multiple of 4."""
    n = nqg + nqg * 30 + nqm + nqm * 10 + 16 * nqg
    return n + (-n % 4)


def _p2_pair_slots(a):
    """(barycentric index m, coefficient row of c) of the two terms of
    grad phi_a = sum_i c_(a,i)(q) grad L_m(a,i); a vertex has one term (the
    second row is None)."""
    if a < 4:
        return (a, a), (a, None)
    i, j = _P2_PAIRS[a - 4]
    return (i, 4 + 2 * (a - 4)), (j, 5 + 2 * (a - 4))


def pair_tables(wg, dg, wm, pm, c):
    """(10, 10, 8) per node pair (a, b) of the 10-node element, for gathers
    that keep one 3x3 block per pair:
      [0:6]  q_(ab,i) = sum_q wg_q c_(a,i)(q) grad phi_b(q), i = 0, 1, which
             factor the gradient sum  sum_q wg_q grad phi_a grad phi_b^T =
             sum_i grad L_m(a,i) q_(ab,i)^T  (rank two at most),
      [6]    M_ab = sum_q wm_q phi_a(q) phi_b(q),
      [7]    0 (16-byte rows).
    Checked against the sums taken directly from the tabulated gradients."""
    n = dg.shape[1]
    out = np.zeros((n, n, 8))
    for a in range(n):
        for i, (_, row) in enumerate(_p2_pair_slots(a)):
            if row is not None:
                out[a, :, 3 * i:3 * i + 3] = np.einsum("q,q,qbk->bk", wg, c[row], dg)
    out[:, :, 6] = np.einsum("q,qa,qb->ab", wm, pm, pm)
    G = np.einsum("q,qai,qbj->abij", wg, dg, dg)
    rec = np.zeros_like(G)
    for a in range(n):
        for i, (m, _) in enumerate(_p2_pair_slots(a)):
            rec[a] += np.einsum("i,bj->bij", _DL[m], out[a, :, 3 * i:3 * i + 3])
    if not np.allclose(rec, G, rtol=0., atol=1e-14 * max(1., np.abs(G).max())):
        raise AssertionError("pair tables do not reproduce the gradient sums")
    return out


def _is_affine_p2(p, t):
    """True when every edge node of a 10-node mesh is the exact midpoint of
    its edge (as MeshTet2.from_mesh creates it), so the isoparametric map is
    affine and the element Jacobian can be taken from the corners."""
    pairs = np.array(_P2_PAIRS)
    a = p[:, t[pairs[:, 0]]]
    b = p[:, t[pairs[:, 1]]]
    mid = p[:, t[4:10]]
    return bool(np.array_equal(mid, 0.5 * (a + b)))


class Basis:
    """H1 basis on a tetrahedral mesh (MI355X build)."""

    def __init__(self, mesh: MeshTet, element: ElementVector, intorder=2, quadrature=None, elements=None):
        intorder = 2 if intorder is None else int(intorder)
        self.intorder = intorder
        if isinstance(element, ElementComposite):
            if elements is not None:
                raise NotImplementedError("element restriction is outside the MI355X assembly build")
            self._init_composite(mesh, element, intorder, quadrature)
            return
        scalar = element.elem if isinstance(element, ElementVector) else None
        if not isinstance(element, ElementVector) or not isinstance(scalar, (ElementTetP1, ElementTetP2)):
            raise NotImplementedError(
                "independent Basis supports P0 and nodal H1/DG elements; the MI355X build assembles "
                "ElementVector(ElementTetP1|ElementTetP2)")
        if elements is not None:
            raise NotImplementedError("element restriction is outside the MI355X assembly build")
        if mesh.t.shape[0] not in (4, 10):
            raise NotImplementedError("the MI355X build supports tetrahedral meshes")
        self.mesh, self.elem = mesh, element
        quadratic = isinstance(scalar, ElementTetP2)
        self._quadratic = quadratic
        if quadrature is not None:
            self.X, self.W = validate_quadrature(quadrature, 3)
        elif intorder > 4:
            self.X, self.W = simplex_rule(intorder)
        elif quadratic or intorder >= 4:
            self.X, self.W = tet_rule_keast11()
        else:
            self.X, self.W = tet_rule_4()
        self.quadrature = (self.X.copy(), self.W.copy())
        nodes = 10 if quadratic else 4
        if mesh.t.shape[0] < nodes:
            raise ValueError("mesh geometry nodes are incompatible with the element")
        connectivity = mesh.t[:nodes]
        components = element._dim
        used = np.zeros(mesh.p.shape[1], dtype=bool)
        used[connectivity] = True
        active = np.flatnonzero(used)
        self.N = len(active) * components
        self.nodal_dofs = np.arange(self.N).reshape(-1, components).T
        node_pos = np.full(mesh.p.shape[1], -1, dtype=np.int64)
        node_pos[active] = np.arange(len(active))
        local = node_pos[connectivity]  # (n,E) node-level DOF ids
        self._node_local = local
        self.element_dofs = self.nodal_dofs[:, local].transpose(2, 1, 0).reshape(mesh.nelements, nodes * components).T
        self.doflocs = np.repeat(mesh.p[:, active], components, axis=1)
        self.active_nodes = active
        self.field_connectivity = connectivity
        self.tind = np.arange(mesh.nelements, dtype=np.int64)
        self.normals = None
        # reference tables
        shape, grad = tabulate(self.X, quadratic)
        gn = mesh.t.shape[0]
        gshape, ggrad = tabulate(self.X, gn == 10)
        self._ref_shape, self._ref_grad = shape, grad
        self._gshape, self._ggrad = gshape, ggrad
        self._affine = gn == 4 or _is_affine_p2(mesh.p, mesh.t)
        self._device = None
        self._lazy = {}
        self._check_geometry()

    # ---------------------------------------------------------- composite
    def _init_composite(self, mesh, element, intorder, quadrature=None):
        """Composite basis (basis.py:2076-2158): one sub-basis per field and a
        node-interleaved numbering -- nodes in increasing id, at each node the
        fields in order, each with its consecutive components.  Sub-bases keep
        their field-local node numbering for the kernels; their
        ``element_dofs`` / ``nodal_dofs`` / ``N`` follow the composite numbering
        as in the reference."""
        fields = tuple(f if isinstance(f, ElementVector) else ElementVector(f, dim=1) for f in element.elems)
        subbases = [Basis(mesh, f, intorder=intorder, quadrature=quadrature) for f in fields]
        shapes = {(b.X.shape, b.W.shape) for b in subbases}
        if len(shapes) != 1 or any(not np.array_equal(subbases[0].X, b.X) for b in subbases[1:]):
            raise ValueError("composite fields require a common quadrature rule; "
                             "increase intorder for mixed-order elements")
        components = tuple(f._dim for f in fields)
        npoints = mesh.p.shape[1]
        active = np.zeros((len(subbases), npoints), dtype=bool)
        for k, b in enumerate(subbases):
            active[k, b.active_nodes] = True
        per_field = active * np.asarray(components)[:, None]          # DOFs of field k at each node
        node_start = np.concatenate([[0], np.cumsum(per_field.sum(axis=0))])[:-1]
        within = np.cumsum(per_field, axis=0) - per_field              # DOFs of earlier fields at the node
        N = int(per_field.sum())
        for k, (b, C) in enumerate(zip(subbases, components)):
            first = node_start[b.active_nodes] + within[k, b.active_nodes]
            nodal = first[None, :] + np.arange(C)[:, None]             # (C, active nodes)
            b._field_N = b.N
            b._composite_dofs = np.ascontiguousarray(nodal.T).ravel()  # field-local DOF -> composite DOF
            b._composite_parent = self
            b.nodal_dofs = nodal
            b.N = N
            b.element_dofs = nodal[:, b._node_local].transpose(2, 1, 0).reshape(
                mesh.nelements, b._node_local.shape[0] * C).T
        common = np.flatnonzero(active.all(axis=0))
        self.mesh, self.elem = mesh, element
        self.subbases = tuple(subbases)
        self.field_components = components
        self.N = N
        self.nodal_dofs = np.concatenate(
            [b.nodal_dofs[:, np.searchsorted(b.active_nodes, common)] for b in subbases], axis=0)
        self.element_dofs = np.concatenate([b.element_dofs for b in subbases], axis=0)
        doflocs = np.empty((mesh.p.shape[0], N))
        for b in subbases:
            for c in range(b.nodal_dofs.shape[0]):
                doflocs[:, b.nodal_dofs[c]] = mesh.p[:, b.active_nodes]
        self.doflocs = doflocs
        first = subbases[0]
        self.X, self.W = first.X, first.W
        self.quadrature = (self.X.copy(), self.W.copy())
        self.normals = None
        self.basis = tuple()
        self.tind = first.tind
        self.geometry_diagnostics = first.geometry_diagnostics
        self._composite_first = first

    # ---------------------------------------------------------- device data
    def device_data(self):
        """Device-resident arrays used by the kernels (built once)."""
        if self._device is None:
            import torch
            L = _native.lib()  # noqa: F841  (fails loudly without a device)
            dev = _native.device()
            mesh = self.mesh
            xyz = np.zeros((mesh.p.shape[1], 4))
            xyz[:, :3] = mesh.p.T
            gn = mesh.t.shape[0]
            geo_tables = pack_tables(self.W, self._ref_shape, self._ref_grad, self._gshape, self._ggrad)
            p1s, p1g = tabulate(self.X, False)
            affine_tables = pack_tables(self.W, self._ref_shape, self._ref_grad, p1s, p1g)
            rtab, nqg, nqm = reduced_tables(self.X, self.W, self._quadratic)
            d = {
                "xyz": torch.from_numpy(xyz).to(dev),
                "vconn": torch.from_numpy(np.ascontiguousarray(mesh.t[:4].T, dtype=np.int32)).to(dev),
                "gconn": (torch.from_numpy(np.ascontiguousarray(mesh.t.T, dtype=np.int32)).to(dev)
                          if gn == 10 else None),
                "gnodes": gn,
                "node_conn": torch.from_numpy(np.ascontiguousarray(self._node_local.T, dtype=np.int32)).to(dev),
                "nnodes": len(self.active_nodes),
                # tables for isoparametric geometry (gphi over the mesh's geometry nodes)
                "tables_iso": torch.from_numpy(geo_tables).to(dev),
                # tables for affine assembly (gphi over the 4 corners)
                "tables": torch.from_numpy(affine_tables).to(dev),
                "rtab": torch.from_numpy(rtab).to(dev),
                # 10-node rtab carries the barycentric coefficients after pm
                "rtab_barycentric": self._quadratic and len(rtab) == _pair_table_offset(nqg, nqm) + 800,
                "rtab32": torch.from_numpy(rtab.astype(np.float32)).to(dev),
                "nqg": nqg,
                "nqm": nqm,
            }
            self._device = d
        return self._device

    def _geometry_call(self, det=None, tol=None, cond=None, dx=None, x=None, shape=None, grads=None):
        d = self.device_data()
        L = _native.lib()
        iso = self.mesh.t.shape[0] == 10
        _native.check(L.skfa_tabulate_geometry(
            _native.ptr(d["xyz"]), _native.ptr(d["vconn"]), _native.ptr(d["gconn"]) if iso else None,
            d["gnodes"] if iso else 4, self.mesh.nelements,
            _native.ptr(d["tables_iso"] if iso else d["tables"]), len(self.W), self._ref_shape.shape[1],
            _native.ptr(det), _native.ptr(tol), _native.ptr(cond), _native.ptr(dx), _native.ptr(x),
            _native.ptr(shape), _native.ptr(grads), _native.stream()))

    def _check_geometry(self):
        import torch
        E, Q = self.mesh.nelements, len(self.W)
        dev = None
        try:
            dev = _native.device() if torch.cuda.is_available() else None
        except Exception:
            dev = None
        if dev is None:
            _native.lib()  # raises NativeUnavailable
        det = torch.empty((E, Q), dtype=torch.float64, device=dev)
        tol = torch.empty_like(det)
        cond = torch.empty_like(det)
        self._geometry_call(det=det, tol=tol, cond=cond)
        det, tol, cond = det.cpu().numpy(), tol.cpu().numpy(), cond.cpu().numpy()
        # validation, basis_geometry.cpp:178-205
        ok = np.isfinite(det) & np.isfinite(tol) & (np.abs(det) > tol)
        neg = det < 0.
        flip = neg != neg[:, :1]
        bad_rows = np.flatnonzero((~ok).any(axis=1) | flip.any(axis=1))
        if bad_rows.size:
            e = int(bad_rows[0])
            if not ok[e].all():
                q = int(np.flatnonzero(~ok[e])[0])
                reason = "near_singular_or_non_finite"
            else:
                q = int(np.flatnonzero(flip[e])[0])
                reason = "orientation_change"
            raise ValueError(
                f"invalid element geometry: cell={e}, quadrature_point={q}, determinant={det[e, q]:.17e}, "
                f"tolerance={tol[e, q]:.17e}, reason={reason}")
        self._geometry_determinants = det
        self._geometry_tolerances = tol
        self._geometry_condition_numbers = cond
        self._update_geometry_diagnostics()

    def _update_geometry_diagnostics(self):
        """basis.py:2296-2356."""
        determinants = self._geometry_determinants
        tolerances = self._geometry_tolerances
        if determinants.size == 0:
            self.geometry_diagnostics = GeometryDiagnostics(
                element_count=0, quadrature_points_per_element=determinants.shape[1] if determinants.ndim == 2 else 0,
                minimum_determinant=float("inf"), maximum_determinant=float("-inf"),
                minimum_scaled_determinant=float("inf"), worst_element=-1, worst_quadrature_point=-1,
                determinant_tolerance=0., negative_orientation_elements=0, maximum_condition_number=0.,
                worst_condition_element=-1, worst_condition_quadrature_point=-1)
            return
        scale_power = tolerances / (64. * _EPS)
        scaled = np.abs(determinants) / scale_power
        flat = int(np.argmin(scaled))
        e, q = np.unravel_index(flat, determinants.shape)
        cflat = int(np.argmax(self._geometry_condition_numbers))
        ce, cq = np.unravel_index(cflat, self._geometry_condition_numbers.shape)
        self.geometry_diagnostics = GeometryDiagnostics(
            element_count=int(determinants.shape[0]),
            quadrature_points_per_element=int(determinants.shape[1]),
            minimum_determinant=float(np.min(determinants)),
            maximum_determinant=float(np.max(determinants)),
            minimum_scaled_determinant=float(scaled.flat[flat]),
            worst_element=int(self.tind[e]),
            worst_quadrature_point=int(q),
            determinant_tolerance=float(tolerances.flat[flat]),
            negative_orientation_elements=int(np.count_nonzero(determinants[:, 0] < 0.)),
            maximum_condition_number=float(self._geometry_condition_numbers.flat[cflat]),
            worst_condition_element=int(self.tind[ce]),
            worst_condition_quadrature_point=int(cq))

    # ---------------------------------------------------------- lazy tabulations
    def _tabulate_all(self):
        if "shape" in self._lazy:
            return
        import torch
        E, Q, n = self.mesh.nelements, len(self.W), self._ref_shape.shape[1]
        dev = _native.device()
        out = {
            "shape": torch.empty((E, Q, n), dtype=torch.float64, device=dev),
            "grads": torch.empty((E, Q, n, 3), dtype=torch.float64, device=dev),
            "dx": torch.empty((E, Q), dtype=torch.float64, device=dev),
            "x": torch.empty((E, Q, 3), dtype=torch.float64, device=dev),
        }
        self._geometry_call(dx=out["dx"], x=out["x"], shape=out["shape"], grads=out["grads"])
        for k, v in out.items():
            self._lazy[k] = v.cpu().numpy()

    def _coordinates_only(self):
        if "x" not in self._lazy:
            import torch
            E, Q = self.mesh.nelements, len(self.W)
            x = torch.empty((E, Q, 3), dtype=torch.float64, device=_native.device())
            dx = torch.empty((E, Q), dtype=torch.float64, device=_native.device())
            self._geometry_call(x=x, dx=dx)
            self._lazy["x"] = x.cpu().numpy()
            self._lazy["dx"] = dx.cpu().numpy()
        return self._lazy["x"]

    @property
    def tabulated_shape(self):
        self._tabulate_all()
        return self._lazy["shape"]

    @property
    def tabulated_gradients(self):
        self._tabulate_all()
        return self._lazy["grads"]

    @property
    def dx(self):
        if "_composite_first" in self.__dict__:
            return self._composite_first.dx
        if "dx" not in self._lazy:
            self._coordinates_only()
        return self._lazy["dx"]

    @property
    def global_coordinates(self):
        if "_composite_first" in self.__dict__:
            return self._composite_first.global_coordinates
        return self._coordinates_only()

    @property
    def coefficient_shape(self):
        """(E, Q): the shape every scalar coefficient broadcasts to."""
        return (self.mesh.nelements, len(self.W))

    @property
    def nelems(self):
        return self.mesh.nelements


__all__ = ["Basis", "GeometryDiagnostics", "tet_rule_4", "tet_rule_keast11", "simplex_rule", "tabulate",
           "pack_tables", "validate_quadrature"]
