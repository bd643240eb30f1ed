"""Lowest-order Nedelec (N1) edge elements on affine tetrahedra.

Mirrors ``AffineTetN1Basis`` / ``TetN1Assembler``
(python/skfemntv/hcurl_basis.py:292-445, hcurl_assembler.py:62-97,160-212,
edge_topology.py:33-63, edge_dof_map.py:12-55): first-seen global edge ids
and +-1 orientation signs (built on the GPU), covariant Piola mapping, and a
reusable CSR matrix overwritten by every assembly.  The element integration
and the scatter into CSR run as one GPU kernel; the (E,6,3,Q) mapped basis
arrays of the reference are only materialised (on the host) if a caller
reads ``values`` / ``curls``.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np
from scipy.sparse import csr_matrix

from . import _native
from ._assembly import to_host, Plan, AssemblyMemoryEstimate, enforce_memory_budget, field_layout, upload, to_per_qp
from .basis import tet_rule_4

TET_DIRECTED_EDGES = ((0, 1), (1, 2), (2, 0), (0, 3), (1, 3), (2, 3))
_DL = np.array(((-1., -1., -1.), (1., 0., 0.), (0., 1., 0.), (0., 0., 1.)))


def tetrahedron_quadrature(order):
    """Duffy product rule (hcurl_basis.py:37-56)."""
    count = max(1, (int(order) + 4) // 2)
    g, gw = np.polynomial.legendre.leggauss(count)
    g = (g + 1.) / 2.
    gw = gw / 2.
    X, W = [], []
    for i in range(count):
        for j in range(count):
            for k in range(count):
                x = g[i]
                y = (1. - x) * g[j]
                z = (1. - x) * (1. - g[j]) * g[k]
                X.append((x, y, z))
                W.append((1. - x) ** 2 * (1. - g[j]) * gw[i] * gw[j] * gw[k])
    return np.asarray(X).T, np.asarray(W)


def tet_n1_basis(points):
    """Whitney functions lambda_i grad lambda_j - lambda_j grad lambda_i as
    (6, 3, ...) (_nedelec_reference.py:72-89)."""
    points = np.asarray(points, dtype=np.float64)
    if points.ndim < 1 or points.shape[0] != 3:
        raise ValueError("tetrahedron reference points must have shape (3, ...)")
    lam = np.concatenate(((1. - np.sum(points, axis=0))[None, ...], points), axis=0)
    out = np.empty((6, 3) + points.shape[1:])
    ax = (slice(None),) + (None,) * (points.ndim - 1)
    for k, (i, j) in enumerate(TET_DIRECTED_EDGES):
        out[k] = lam[i][None, ...] * _DL[j][ax] - lam[j][None, ...] * _DL[i][ax]
    return out


def tet_n1_curl_constants():
    """Reference curls 2 grad lambda_i x grad lambda_j, (6,3) (_nedelec_reference.py:92-107)."""
    return np.asarray([2. * np.cross(_DL[i], _DL[j]) for i, j in TET_DIRECTED_EDGES])


@dataclass(frozen=True)
class OrientedEdgeTopology:
    edges: np.ndarray
    element_edges: np.ndarray
    element_edge_signs: np.ndarray
    local_edges: tuple


@dataclass(frozen=True)
class OrientedEdgeDofMap:
    topology: OrientedEdgeTopology

    @property
    def ndofs(self):
        return int(self.topology.edges.shape[1])

    @property
    def element_dofs(self):
        return self.topology.element_edges

    @property
    def basis_signs(self):
        return self.topology.element_edge_signs


def build_oriented_edge_topology(mesh):
    """GPU first-seen edge numbering (edge_topology.cpp:30-84)."""
    import ctypes
    import torch
    if mesh.t.shape[0] < 4:
        raise ValueError("mesh connectivity does not contain all corner vertices")
    L = _native.lib()
    dev = _native.device()
    E = mesh.nelements
    vconn = torch.from_numpy(np.ascontiguousarray(mesh.t[:4].T, dtype=np.int32)).to(dev)
    ee = torch.empty((max(E, 1), 6), dtype=torch.int32, device=dev)
    sg = torch.empty((max(E, 1), 6), dtype=torch.int8, device=dev)
    lo = torch.empty(max(6 * E, 1), dtype=torch.int64, device=dev)
    hi = torch.empty(max(6 * E, 1), dtype=torch.int64, device=dev)
    ne = ctypes.c_int64()
    bad = ctypes.c_int64()
    rc = L.skfa_edge_topology(_native.ptr(vconn), E, mesh.p.shape[1], _native.ptr(ee), _native.ptr(sg),
                              _native.ptr(lo), _native.ptr(hi), ctypes.byref(ne), ctypes.byref(bad), _native.stream())
    _native.check(rc)
    n = ne.value
    edges = np.vstack((lo[:n].cpu().numpy(), hi[:n].cpu().numpy()))
    topo = OrientedEdgeTopology(edges, ee[:E].cpu().numpy().T.astype(np.int64).copy(),
                                sg[:E].cpu().numpy().T.copy(), TET_DIRECTED_EDGES)
    return topo, vconn, ee[:E]


@dataclass(frozen=True)
class HcurlTetGeometryDiagnostics:
    minimum_signed_determinant: float
    minimum_absolute_determinant: float
    minimum_volume: float
    maximum_aspect_ratio: float
    inverted_cell_count: int


class AffineTetN1Basis:
    """Lowest-order H(curl) basis on affine tetrahedral meshes (hcurl_basis.py:292-364)."""

    def __init__(self, mesh, *, intorder=2, max_aspect_ratio=None):
        if mesh.dim() != 3 or mesh.t.shape[0] not in (4, 10):
            raise TypeError("AffineTetN1Basis requires a tetrahedral mesh")
        self.mesh = mesh
        topo, vconn, ee = build_oriented_edge_topology(mesh)
        self.dof_map = OrientedEdgeDofMap(topo)
        self._vconn_dev, self._edges_dev = vconn, ee
        self.N = self.dof_map.ndofs
        self.element_dofs = self.dof_map.element_dofs
        self.basis_signs = self.dof_map.basis_signs
        self.X, self.W = tetrahedron_quadrature(intorder)
        v = mesh.p[:, mesh.t[:4]]
        self.jacobians = np.stack(tuple(v[:, k] - v[:, 0] for k in (1, 2, 3)), axis=1)
        J = self.jacobians
        self.detJ = (J[0, 0] * (J[1, 1] * J[2, 2] - J[1, 2] * J[2, 1])
                     - J[0, 1] * (J[1, 0] * J[2, 2] - J[1, 2] * J[2, 0])
                     + J[0, 2] * (J[1, 0] * J[2, 1] - J[1, 1] * J[2, 0]))
        if np.any(np.isclose(self.detJ, 0.)):
            raise ValueError("H(curl) basis requires nonsingular tetrahedra")
        lengths = [np.linalg.norm(v[:, b] - v[:, a], axis=0) for a, b in TET_DIRECTED_EDGES]
        longest = np.max(lengths, axis=0)
        aspect = longest ** 3 / np.abs(self.detJ)
        self.geometry_diagnostics = HcurlTetGeometryDiagnostics(
            minimum_signed_determinant=float(np.min(self.detJ)),
            minimum_absolute_determinant=float(np.min(np.abs(self.detJ))),
            minimum_volume=float(np.min(np.abs(self.detJ)) / 6.),
            maximum_aspect_ratio=float(np.max(aspect)),
            inverted_cell_count=int(np.count_nonzero(self.detJ < 0.)))
        if max_aspect_ratio is not None:
            if (isinstance(max_aspect_ratio, bool) or not np.isscalar(max_aspect_ratio)
                    or not np.isfinite(max_aspect_ratio) or max_aspect_ratio <= 0.):
                raise ValueError("max_aspect_ratio must be a positive finite scalar")
            if self.geometry_diagnostics.maximum_aspect_ratio > max_aspect_ratio:
                raise ValueError(f"H(curl) tetrahedron aspect ratio exceeds max_aspect_ratio={max_aspect_ratio}")
        self._mapped = None

    @property
    def dx(self):
        return np.abs(self.detJ)[:, None] * self.W[None, :]

    def _map(self):
        if self._mapped is None:
            Jm = np.moveaxis(self.jacobians, -1, 0)
            inv = np.linalg.inv(Jm)
            rv = tet_n1_basis(self.X)
            rc = tet_n1_curl_constants()
            values = np.einsum("erp,brq->ebpq", inv, rv)
            curls = np.einsum("epr,br->ebp", Jm, rc)[..., None] / self.detJ[:, None, None, None]
            curls = np.broadcast_to(curls, values.shape)
            s = self.basis_signs.T[:, :, None, None].astype(np.float64)
            self._mapped = (np.ascontiguousarray(values * s), np.ascontiguousarray(curls * s))
        return self._mapped

    @property
    def _element_values(self):
        return self._map()[0]

    @property
    def _element_curls(self):
        return self._map()[1]

    @property
    def values(self):
        return self._element_values.transpose(1, 2, 0, 3)

    @property
    def curls(self):
        return self._element_curls.transpose(1, 2, 0, 3)

    def element_mass_matrices(self):
        return np.einsum("ebiq,eciq,eq->ebc", self._element_values, self._element_values, self.dx)

    def element_curl_curl_matrices(self):
        return np.einsum("ebiq,eciq,eq->ebc", self._element_curls, self._element_curls, self.dx)

    @property
    def global_coordinates(self):
        vertices = self.mesh.p[:, self.mesh.t[0]]
        return vertices[:, :, None] + np.einsum("ire,rq->ieq", self.jacobians, self.X)


def _n1_tables(X, W):
    return np.concatenate([W, np.transpose(tet_n1_basis(X), (2, 0, 1)).ravel(), tet_n1_curl_constants().ravel()])


def _moments(X, W):
    """sum_q w_q phi_a,i(q) phi_b,j(q): the only way the quadrature enters the
    mass integral for element-constant coefficients."""
    phi = tet_n1_basis(X)  # (6,3,Q)
    return np.einsum("q,aiq,bjq->aibj", W, phi, phi)


def estimate_tet_n1_assembly_memory(basis):
    E = basis.mesh.nelements
    nnz = min(basis.N * basis.N, E * 36)
    return AssemblyMemoryEstimate("hcurl_tet_n1", basis.N, basis.N, nnz,
                                  int(8 * (basis.N + 1) + 12 * nnz + 2 * 16 * 6 * E + 24 * E))


class TetN1Assembler:
    """Reusable mass / curl-curl assembler for affine tetrahedra
    (hcurl_assembler.py:160-212).  Assembly methods overwrite and return one
    CSR matrix; call ``.copy()`` when a result must survive the next call."""

    def __init__(self, basis, *, memory_limit_bytes=None, memory_safety_factor=1.25):
        if not isinstance(basis, AffineTetN1Basis):
            raise TypeError("basis must be AffineTetN1Basis")
        self.basis = basis
        self.memory_estimate = estimate_tet_n1_assembly_memory(basis)
        enforce_memory_budget(self.memory_estimate, memory_limit_bytes, memory_safety_factor)
        E = basis.mesh.nelements
        self.plan = Plan(basis._edges_dev, E, 6, basis.N, 1)
        import torch
        dev = _native.device()
        xyz = np.zeros((basis.mesh.p.shape[1], 4))
        xyz[:, :3] = basis.mesh.p.T
        self._xyz = torch.from_numpy(xyz).to(dev)
        self._values = torch.zeros(max(self.plan.nnz, 1), dtype=torch.float64, device=dev)
        ip, ix = self.plan.host_csr_arrays()
        # the returned CSR object's values live in pinned host memory, which
        # the device fills directly on every call
        self._pinned = torch.zeros(max(self.plan.nnz, 1), dtype=torch.float64, pin_memory=True)
        self._matrix = csr_matrix((self._pinned[:self.plan.nnz].numpy(), ix, ip), shape=(basis.N, basis.N),
                                  copy=False)
        if not np.shares_memory(self._matrix.data, self._pinned.numpy()):
            self._pinned = None
        self._tables_basis = torch.from_numpy(_n1_tables(basis.X, basis.W)).to(dev)
        X4, W4 = tet_rule_4()
        m_basis, m_4 = _moments(basis.X, basis.W), _moments(X4, W4)
        # element-constant coefficients: a rule with the same moments gives
        # the same matrix; use the 4-point one when it matches
        self._equivalent_rule = bool(np.allclose(m_basis, m_4, rtol=1e-14, atol=1e-15))
        self._tables4 = torch.from_numpy(_n1_tables(X4, W4)).to(dev) if self._equivalent_rule else None
        self._shape = (E, len(basis.W))

    def _coefficient(self, name, value):
        if value is None:
            value = 1.
        return field_layout(value, self._shape, name)

    def _assemble(self, mass, curl, include_mass, include_curl, precision="fp64"):
        (m, mq), (c, cq) = mass, curl
        per_qp = mq or cq
        md, cd = upload(m), upload(c)
        Q = self._shape[1]
        if per_qp:
            if not mq:
                md = to_per_qp(md, Q)
            if not cq:
                cd = to_per_qp(cd, Q)
        self.assemble_device(md, cd, per_qp, include_mass, include_curl)
        if self._pinned is not None:
            to_host(self._values[:self.plan.nnz], self._pinned[:self.plan.nnz])
        else:
            np.copyto(self._matrix.data, to_host(self._values[:self.plan.nnz]))
        return self._matrix

    def assemble_device(self, kappa, mu, per_qp, include_mass=True, include_curl=True, out=None):
        """Device-coefficient assembly into a device tensor of CSR values."""
        L = _native.lib()
        vals = self._values if out is None else out
        if not per_qp and self._equivalent_rule:
            tables, nq = self._tables4, 4
        else:
            tables, nq = self._tables_basis, len(self.basis.W)
        def launch():
            _native.check(L.skfa_assemble_tet_n1(
                self.plan.handle, _native.ptr(self._xyz), _native.ptr(self.basis._vconn_dev), _native.ptr(tables),
                nq, _native.ptr(kappa), _native.ptr(mu), int(per_qp), int(include_mass), int(include_curl),
                _native.ptr(vals), _native.stream()))

        if not per_qp:
            from . import autotune
            autotune.maybe_tune(self.plan, launch)
        launch()
        return vals

    def assemble_mass(self, coefficient=None):
        m = self._coefficient("mass", coefficient)
        return self._assemble(m, m, True, False)

    def assemble_curl_curl(self, coefficient=None):
        c = self._coefficient("curl", coefficient)
        return self._assemble(c, c, False, True)

    def assemble_maxwell(self, *, mass_coefficient=None, curl_coefficient=None):
        m = self._coefficient("mass", mass_coefficient)
        c = self._coefficient("curl", curl_coefficient)
        return self._assemble(m, c, True, True)


__all__ = ["AffineTetN1Basis", "TetN1Assembler", "HcurlTetGeometryDiagnostics", "OrientedEdgeTopology",
           "OrientedEdgeDofMap", "build_oriented_edge_topology", "tet_n1_basis", "tet_n1_curl_constants",
           "tetrahedron_quadrature", "estimate_tet_n1_assembly_memory"]
