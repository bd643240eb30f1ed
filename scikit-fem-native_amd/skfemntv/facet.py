"""Boundary (facet) bases of tetrahedral meshes and their linear forms.

``FacetBasis(mesh, element, facets=None, intorder=2, quadrature=None)``
mirrors the tetrahedral branch of python/skfemntv/basis.py:2442-2685: the
same facet selection (boundary facets by default, in facet-id order), parent
elements, triangle quadrature rules, quadrature-point placement on the
parent's local face, outward normals, ``dx = |t1 x t2| w`` and
``global_coordinates``.  The per-facet geometry is computed vectorised on
the host (facets are O(n^2) against O(n^3) elements); the facet load vector
``asm(LinearForm, facet_basis)`` is reduced on the GPU per DOF node in facet
order (skfa_assemble_facet_linear).
"""

from __future__ import annotations

import numpy as np

from . import _native
from .basis import Basis, tabulate, validate_quadrature
from .element import ElementTetP1, ElementTetP2, ElementVector
from .mesh import TET_FACES

_REF_VERTICES = np.array([[0., 0., 0.], [1., 0., 0.], [0., 1., 0.], [0., 0., 1.]])


def triangle_rule(intorder=2, quadrature=None):
    """Face rule (basis.py:2518-2528, 2557-2566): 3 points (intorder < 4),
    the 6-point rule (intorder 4), collapsed Gauss above, or a given
    (points (2,Q), weights (Q,)).  Returns (r, s) points (Q, 2), weights."""
    if quadrature is not None:
        points, weights = validate_quadrature(quadrature, 2)
        return points.T, weights
    if intorder > 4:
        count = max(1, (int(intorder) + 3) // 2)
        g, gw = np.polynomial.legendre.leggauss(count)
        g, gw = (g + 1.) / 2., gw / 2.
        pts, wts = [], []
        for i in range(count):
            for j in range(count):
                a, b = g[i], g[j]
                pts.append((a, (1. - a) * b))
                wts.append((1. - a) * gw[i] * gw[j])
        return np.asarray(pts), np.asarray(wts)
    if intorder >= 4:
        a, b = .445948490915965, .091576213509771
        bary = np.array([[a, a, 1 - 2 * a], [a, 1 - 2 * a, a], [1 - 2 * a, a, a],
                         [b, b, 1 - 2 * b], [b, 1 - 2 * b, b], [1 - 2 * b, b, b]])
        return bary[:, 1:], np.array([.1116907948390055] * 3 + [.054975871827661] * 3)
    bary = np.array([[2 / 3, 1 / 6, 1 / 6], [1 / 6, 2 / 3, 1 / 6], [1 / 6, 1 / 6, 2 / 3]])
    return bary[:, 1:], np.full(3, 1 / 6)


class FacetBasis:
    """Tetrahedral facet basis (basis.py:2442-2685, tetrahedral branch)."""

    def __init__(self, mesh, element, facets=None, intorder=2, quadrature=None, *, _side=0, _interior=False):
        intorder = 2 if intorder is None else int(intorder)
        if mesh.t.shape[0] not in (4, 10):
            raise NotImplementedError("the MI355X build supports tetrahedral facet bases")
        scalar = element.elem if isinstance(element, ElementVector) else None
        if not isinstance(scalar, (ElementTetP1, ElementTetP2)):
            raise NotImplementedError("FacetBasis supports ElementVector(ElementTetP1|ElementTetP2)")
        if isinstance(facets, str):
            try:
                facets = mesh.boundaries[facets]
            except KeyError as error:
                raise KeyError(f"unknown boundary {facets!r}") from error
        volume = Basis(mesh, element, intorder=intorder)
        if facets is None:
            facet_ids = mesh.interior_facets() if _interior else mesh.boundary_facets()
        else:
            facet_ids = np.asarray(facets, dtype=np.int64)
        if facet_ids.ndim != 1:
            raise NotImplementedError("the MI355X FacetBasis takes facet ids")
        F = facet_ids.size
        # side 1 of an interior facet is its second neighbour (basis.py:2746-2752)
        parents = mesh.f2t[_side, facet_ids]
        if np.any(parents < 0):
            bad = int(facet_ids[np.flatnonzero(parents < 0)[0]])
            raise ValueError(f"facet {bad} does not have side {_side}")
        # local face of the parent that is this facet (first match)
        match = mesh.t2f[:, parents] == facet_ids[None, :]
        local = np.argmax(match, axis=0)
        face_points, face_weights = triangle_rule(intorder, quadrature)
        Q = len(face_weights)
        quadratic = isinstance(scalar, ElementTetP2)
        # reference points (1-r-s) V[c0] + r V[c1] + s V[c2], with c0, c1, c2
        # the parent's local corners of the facet's corners in the facet's own
        # (sorted) order (basis.py:2634-2645), so both sides of an interior
        # facet place a rule point at the same physical point.  The distinct
        # corner orders are tabulated once each.
        face_nodes = mesh.facets[:, facet_ids].T                         # (F, 3)
        tt = np.asarray(mesh.t[:4, parents]).T                           # (F, 4)
        corner_of = np.argmax(tt[:, None, :] == face_nodes[:, :, None], axis=2)  # (F, 3)
        combos, orient = np.unique(corner_of, axis=0, return_inverse=True)
        orient = np.asarray(orient).ravel()
        K = combos.shape[0]
        Vc = _REF_VERTICES[combos]                                       # (K, 3, 3)
        r, s_ = face_points[:, 0], face_points[:, 1]
        ref = ((1 - r - s_)[None, :, None] * Vc[:, None, 0] + r[None, :, None] * Vc[:, None, 1]
               + s_[None, :, None] * Vc[:, None, 2])                     # (K, Q, 3)
        derivatives = np.stack((Vc[:, 1] - Vc[:, 0], Vc[:, 2] - Vc[:, 0]), axis=2)  # (K, 3, 2)
        shape, refgrad = tabulate(ref.reshape(-1, 3).T, quadratic)
        n = shape.shape[1]
        shape = shape.reshape(K, Q, n)
        refgrad = refgrad.reshape(K, Q, n, 3)
        gn = mesh.t.shape[0]
        gshape, ggrad = tabulate(ref.reshape(-1, 3).T, gn == 10)
        gshape, ggrad = gshape.reshape(K, Q, gn), ggrad.reshape(K, Q, gn, 3)
        x = mesh.p[:, mesh.t[:, parents]].transpose(2, 0, 1)            # (F, 3, gn)
        J = np.einsum("fig,fqgj->fqij", x, ggrad[orient])                # (F, Q, 3, 3)
        tangents = np.einsum("fqij,fjk->fqik", J, derivatives[orient])   # (F, Q, 3, 2)
        points = np.einsum("fqg,fig->fqi", gshape[orient], x)            # (F, Q, 3)
        cross = np.cross(tangents[..., 0], tangents[..., 1])
        area = np.linalg.norm(cross, axis=-1)
        normal = cross / area[..., None]
        centre = x.mean(axis=2)
        outward = np.einsum("fqi,fqi->fq", normal, points - centre[:, None, :])
        normal = np.where((outward < 0.)[..., None], -normal, normal)
        if _interior and _side == 1:
            # both traces of an interior facet share side 0's normal (basis.py:2794-2795)
            normal = -normal
        self.mesh, self.elem = mesh, element
        self.X, self.W = face_points.T.copy(), np.asarray(face_weights, dtype=np.float64)
        self.quadrature = (self.X.copy(), self.W.copy())
        self.N = volume.N
        self.doflocs = volume.doflocs
        self.element_dofs = volume.element_dofs[:, parents]
        self.tabulated_shape = shape[orient]                             # (F, Q, n)
        self.tabulated_gradients = np.einsum("fqak,fqkj->fqaj", refgrad[orient], np.linalg.inv(J))
        self.dx = area * self.W[None, :]
        self.global_coordinates = points
        self.normals = normal
        self.parent_elements = parents
        self.facet_ids = facet_ids.copy()
        self.facet_sides = np.full(F, _side, dtype=np.int8)
        self.normal_signs = np.ones(F, dtype=np.int8)
        self.local_faces = [tuple(TET_FACES[k]) for k in local]
        self.volume_basis = volume
        self._local_face = orient.astype(np.int32)                       # row of _face_shape
        self._face_shape = shape                                         # (K, Q, n)
        self._device = None

    @property
    def coefficient_shape(self):
        return self.dx.shape

    def ragged(self):
        """The facets as cut-style cells of Q points each: DOFs (F, n, C),
        point offsets, per-point shape values, physical gradients, weights
        and normals, facet-major (the arguments of the ragged-quadrature
        assemblers in ``skfemntv.cut``)."""
        F, Q = self.dx.shape
        n = self.tabulated_shape.shape[2]
        C = self.elem._dim
        return {"dofs": np.ascontiguousarray(self.element_dofs.T.reshape(F, n, C), dtype=np.int64),
                "offsets": np.arange(F + 1, dtype=np.int64) * Q,
                "shape": np.ascontiguousarray(self.tabulated_shape.reshape(F * Q, n)),
                "gradients": np.ascontiguousarray(self.tabulated_gradients.reshape(F * Q, n, 3)),
                "weights": np.ascontiguousarray(self.dx.reshape(F * Q)),
                "normals": np.ascontiguousarray(self.normals.reshape(F * Q, 3))}

    @property
    def nelems(self):
        return self.dx.shape[0]

    def device_data(self):
        if self._device is None:
            import torch
            _native.lib()
            dev = _native.device()
            local = self.volume_basis._node_local[:, self.parent_elements]   # (n, F) node ids
            self._device = {
                "node_conn": torch.from_numpy(np.ascontiguousarray(local.T, dtype=np.int32)).to(dev),
                "nnodes": len(self.volume_basis.active_nodes),
                "face": torch.from_numpy(self._local_face).to(dev),
                "face_shape": torch.from_numpy(np.ascontiguousarray(self._face_shape)).to(dev),
            }
        return self._device


class InteriorFacetBasis(FacetBasis):
    """Traces on one side of every selected interior facet
    (basis.py:2993-3013): side 0 is the facet's first neighbour, side 1 its
    second; both sides carry side 0's unit normal."""

    def __init__(self, mesh, element, mapping=None, intorder=2, quadrature=None, facets=None, dofs=None, side=0,
                 disable_doflocs=False):
        if mapping is not None or dofs is not None:
            raise NotImplementedError("custom mapping and dofs are not implemented")
        if side not in (0, 1):
            raise ValueError("side must be 0 or 1")
        super().__init__(mesh, element, facets=facets, intorder=intorder, quadrature=quadrature, _side=side,
                         _interior=True)
        self.side = side
        if disable_doflocs:
            self.doflocs = None


class NativeFacetLinearForm:
    """f . v on a FacetBasis (linear_form.py:17-107 over facet DOFs); returns
    (vector of length N, LinearFormDiagnostics) like NativeLinearForm."""

    def __init__(self, basis):
        from ._assembly import Plan
        import torch
        self.basis = basis
        d = basis.device_data()
        F, Q = basis.dx.shape
        C = basis.elem._dim
        self._shape = (F, Q, C)
        self._gradient_shape = (F, Q, C, 3)
        self.plan = Plan(d["node_conn"], F, d["node_conn"].shape[1], d["nnodes"], 1)
        self._out = torch.zeros(max(d["nnodes"] * C, 1), dtype=torch.float64, device=_native.device())

    @property
    def ndofs(self):
        return self.basis.N

    def _assemble_ragged(self, value, gradient):
        """Gradient terms G:grad(v): the facets as cells of Q points through
        the ragged-quadrature load-vector kernel."""
        import time
        from ._assembly import LinearFormDiagnostics
        from .cut import CutLinearFormAssembler
        F, Q, C = self._shape
        if getattr(self, "_cut", None) is None:
            r = self.basis.ragged()
            self._cut = CutLinearFormAssembler(r["dofs"], r["offsets"], r["shape"], r["gradients"], r["weights"])
        start = time.perf_counter()
        f = None
        if value is not None:
            try:
                f = np.broadcast_to(np.asarray(value, dtype=np.float64), (F, Q, C)).reshape(F * Q, C)
            except ValueError as error:
                raise ValueError(f"value coefficient must broadcast to {(F, Q, C)}") from error
        try:
            G = np.broadcast_to(np.asarray(gradient, dtype=np.float64), (F, Q, C, 3)).reshape(F * Q, C, 3)
        except ValueError as error:
            raise ValueError(f"gradient coefficient must broadcast to {(F, Q, C, 3)}") from error
        out, _ = self._cut.assemble(f, G)
        result = np.zeros(self.basis.N, dtype=np.float64)
        result[:out.shape[0]] = out
        seconds = time.perf_counter() - start
        return result, LinearFormDiagnostics(entity_count=F, quadrature_point_count=Q, assembly_seconds=seconds)

    def assemble(self, *, value=None, gradient=None, num_threads=0):
        import time
        from ._assembly import LinearFormDiagnostics, to_host, upload
        if gradient is not None:
            return self._assemble_ragged(value, gradient)
        if value is None:
            raise ValueError("at least one coefficient is required")
        F, Q, C = self._shape
        f = np.asarray(value, dtype=np.float64)
        try:
            f = np.broadcast_to(f, (F, Q, C))
        except ValueError as error:
            raise ValueError(f"value coefficient must broadcast to {(F, Q, C)}") from error
        start = time.perf_counter()
        weights = upload(np.ascontiguousarray(self.basis.dx[..., None] * f))
        d = self.basis.device_data()
        L = _native.lib()
        _native.check(L.skfa_assemble_facet_linear(
            self.plan.handle, _native.ptr(d["face"]), _native.ptr(d["face_shape"]), Q, _native.ptr(weights),
            _native.ptr(self._out), C, _native.stream()))
        nodal = to_host(self._out[:d["nnodes"] * C])
        # node-level rows are the volume basis's DOF order (node * C + c)
        result = np.zeros(self.basis.N, dtype=np.float64)
        result[:nodal.shape[0]] = nodal
        seconds = time.perf_counter() - start
        return result, LinearFormDiagnostics(entity_count=F, quadrature_point_count=Q, assembly_seconds=seconds)


class NativeFacetBilinearForm:
    """Value-kind bilinear forms on a FacetBasis (bilinear_form.py:13-99 over
    facet DOFs): boundary mass / Robin terms.  ``assemble`` returns the same
    CSR object on every call, shape (N, N), with the reference's facet
    pattern (every pair of DOFs of each facet's parent element)."""

    def __init__(self, basis):
        from scipy.sparse import csr_matrix
        from ._assembly import Plan
        import torch
        self.basis = basis
        d = basis.device_data()
        F, Q = basis.dx.shape
        C = basis.elem._dim
        self.plan = Plan(d["node_conn"], F, d["node_conn"].shape[1], d["nnodes"], C)
        self._values = torch.zeros(max(self.plan.nnz, 1), dtype=torch.float64, device=_native.device())
        ip, ix = self.plan.host_csr_arrays()
        self._matrix = csr_matrix((np.zeros(self.plan.nnz), ix, ip), shape=(self.plan.nrows, self.plan.nrows),
                                  copy=False)
        self._matrix.resize((basis.N, basis.N))

    def assemble(self, *, value=None, gradient=None, symmetric_gradient=None, divergence=None, num_threads=0,
                 precision="fp64"):
        from ._assembly import to_host, upload
        from ._symbols import UnsupportedNativeForm
        if precision != "fp64":
            raise UnsupportedNativeForm("facet bilinear forms are assembled in FP64")
        if value is None and gradient is None and symmetric_gradient is None and divergence is None:
            raise ValueError("at least one coefficient is required")
        F, Q = self.basis.dx.shape
        if symmetric_gradient is not None or divergence is not None:
            return self._assemble_tensor(value, gradient, symmetric_gradient, divergence)
        if gradient is not None:
            return self._assemble_ragged(value, gradient)
        c = np.asarray(value, dtype=np.float64)
        try:
            c = np.broadcast_to(c, (F, Q))
        except ValueError as error:
            raise ValueError(f"value coefficient must broadcast to {(F, Q)}") from error
        weights = upload(np.ascontiguousarray(self.basis.dx * c))
        d = self.basis.device_data()
        L = _native.lib()
        _native.check(L.skfa_assemble_facet_bilinear(
            self.plan.handle, _native.ptr(d["node_conn"]), _native.ptr(d["face"]), _native.ptr(d["face_shape"]), Q,
            _native.ptr(weights), _native.ptr(self._values), _native.stream()))
        np.copyto(self._matrix.data, to_host(self._values[:self.plan.nnz]))
        return self._matrix

    def _assemble_ragged(self, value, gradient):
        """Gradient terms g grad(u).grad(v) (with an optional value term):
        the facets as cells of Q points through the ragged-quadrature kernel,
        whose pattern is this form's facet pattern."""
        from .cut import CutBilinearFormAssembler
        F, Q = self.basis.dx.shape
        if getattr(self, "_cut", None) is None:
            r = self.basis.ragged()
            cut = CutBilinearFormAssembler(r["dofs"], r["offsets"], r["shape"], r["gradients"], r["weights"])
            ip, ix = self.plan.host_csr_arrays()
            if not (np.array_equal(cut.indices, ix) and np.array_equal(cut.indptr, ip[:len(cut.indptr)])):
                raise RuntimeError("facet gradient pattern differs from the facet plan's pattern")
            self._cut = cut
        coefs = []
        for name, c in (("value", value), ("gradient", gradient)):
            if c is None:
                coefs.append(None)
                continue
            try:
                coefs.append(np.ascontiguousarray(np.broadcast_to(np.asarray(c, dtype=np.float64),
                                                                  (F, Q))).reshape(F * Q))
            except ValueError as error:
                raise ValueError(f"{name} coefficient must broadcast to {(F, Q)}") from error
        self._cut.assemble(coefs[0], coefs[1])
        np.copyto(self._matrix.data, self._cut.values)
        return self._matrix

    def _per_point(self, name, c):
        F, Q = self.basis.dx.shape
        if c is None:
            return None
        try:
            return np.ascontiguousarray(np.broadcast_to(np.asarray(c, dtype=np.float64), (F, Q))).reshape(F * Q)
        except ValueError as error:
            raise ValueError(f"{name} coefficient must broadcast to {(F, Q)}") from error

    def _assemble_tensor(self, value, gradient, symmetric_gradient, divergence):
        """Symmetric-gradient and divergence kinds
        (bilinear_form_assembler.cpp:76-100): with T[c,i,d,j] =
        g d_cd d_ij + s/2 (d_id d_jc + d_cd d_ij) + k d_ic d_jd the entry
        (a,c),(b,d) is sum_ij dphi_a/dx_i T dphi_b/dx_j, run as a
        gradient-gradient contraction on the ragged cross kernel; a value
        term is a second, scalar launch on the same pattern."""
        from .cut import CutCrossAssembler
        C = self.basis.elem._dim
        if C != 3:
            raise ValueError("symmetric-gradient and divergence kinds need a 3-component field")
        if getattr(self, "_cutx", None) is None:
            r = self.basis.ragged()
            cx = CutCrossAssembler(r["dofs"], r["dofs"], r["offsets"], r["shape"], r["shape"], r["weights"],
                                   r["gradients"], r["gradients"])
            ip, ix = self.plan.host_csr_arrays()
            if not (np.array_equal(cx.indices, ix) and np.array_equal(cx.indptr, ip[:len(cx.indptr)])):
                raise RuntimeError("facet tensor pattern differs from the facet plan's pattern")
            self._cutx = cx
        v, g = self._per_point("value", value), self._per_point("gradient", gradient)
        sg, dv = self._per_point("symmetric_gradient", symmetric_gradient), self._per_point("divergence", divergence)
        P = len(self.basis.dx.ravel())
        I3 = np.eye(3)
        T = np.zeros((P, 3, 3, 3, 3))
        # axes (c, i, d, j)
        unit_grad = np.einsum("cd,ij->cidj", I3, I3)
        if g is not None:
            T += g[:, None, None, None, None] * unit_grad
        if sg is not None:
            T += (0.5 * sg)[:, None, None, None, None] * (np.einsum("id,jc->cidj", I3, I3) + unit_grad)
        if dv is not None:
            T += dv[:, None, None, None, None] * np.einsum("ic,jd->cidj", I3, I3)
        cx = self._cutx
        cx.assemble(T, "gradient", "gradient")
        data = cx.values.copy()
        if v is not None:
            cx.assemble(v, "value", "value")
            data += cx.values
        np.copyto(self._matrix.data, data)
        return self._matrix


__all__ = ["FacetBasis", "InteriorFacetBasis", "NativeFacetBilinearForm", "NativeFacetLinearForm",
           "triangle_rule"]
