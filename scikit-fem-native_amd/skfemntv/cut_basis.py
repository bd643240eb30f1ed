"""Cut-cell quadrature container and the H1 basis tabulated on it.

``CutCellQuadrature`` holds the fields of the reference's container
(python/skfemntv/levelset.py:47-62): flat points with their cells, CSR-like
``cell_offsets`` over all mesh cells, reference coordinates, weights and
normals.  Generating it from a level set (``LevelSet.cut_quadrature``) is
not part of this package; callers build it from arrays, e.g. with
``CutCellQuadrature.from_reference_points``.

``CutCellBasis`` follows python/skfemntv/cut_basis.py:12-150 for
ElementVector(ElementTetP1 | ElementTetP2) on straight-sided tetrahedra: the
shape values at each point's reference coordinates, physical gradients
through the cell's affine Jacobian, the per-point and per-cell DOFs, and
the one-point-per-entity adapter arrays that the form wrappers read.  The
tabulation is vectorised host work, O(points); the assemblers that consume
it run on the GPU (``skfemntv.cut``).
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .basis import Basis, tabulate
from .element import ElementTetP1, ElementTetP2, ElementVector


class DiscreteField:
    """Values and physical gradients at the cut points (basis.py:1121-1133)."""

    def __init__(self, value, grad):
        self.value = np.asarray(value)
        self.grad = None if grad is None else np.asarray(grad)

    def __array__(self, dtype=None):
        return np.asarray(self.value, dtype=dtype)

    @property
    def shape(self):
        return self.value.shape


@dataclass(frozen=True)
class CutCellQuadrature:
    """Per-point cut quadrature over the cells of one mesh
    (levelset.py:47-62).  ``cells[p]`` is the cell of point p and cell c owns
    points [cell_offsets[c], cell_offsets[c+1])."""

    cell_offsets: np.ndarray
    points: np.ndarray
    reference_points: np.ndarray
    weights: np.ndarray
    cells: np.ndarray
    normals: np.ndarray
    side: str = "negative"
    diagnostics: object = None

    def cell_slice(self, cell: int) -> slice:
        if cell < 0 or cell + 1 >= len(self.cell_offsets):
            raise IndexError("cut-quadrature cell index is out of bounds")
        return slice(int(self.cell_offsets[cell]), int(self.cell_offsets[cell + 1]))

    @classmethod
    def from_reference_points(cls, mesh, cell_offsets, reference_points, weights, normals=None, side="negative"):
        """Build the container from per-cell reference coordinates (P, 3) and
        physical weights (P,): points are the affine images in their cells."""
        off = np.ascontiguousarray(cell_offsets, dtype=np.int64)
        if off.ndim != 1 or len(off) != mesh.nelements + 1 or off[0] != 0 or np.any(np.diff(off) < 0):
            raise ValueError("cell offsets must be nondecreasing from 0 over every mesh cell")
        X = np.ascontiguousarray(reference_points, dtype=np.float64)
        w = np.ascontiguousarray(weights, dtype=np.float64)
        if X.ndim != 2 or X.shape != (off[-1], 3) or w.shape != (off[-1],):
            raise ValueError("reference points must be (points, 3) and weights (points,)")
        cells = np.repeat(np.arange(mesh.nelements, dtype=np.int64), np.diff(off))
        corners = mesh.p[:, mesh.t[:4, cells]]  # (3, 4, P)
        J = corners[:, 1:4] - corners[:, [0]]  # (3, 3, P)
        x = corners[:, 0].T + np.einsum("ijp,pj->pi", J, X)
        n = np.zeros((len(w), 3)) if normals is None else np.ascontiguousarray(normals, dtype=np.float64)
        return cls(off, x, X, w, cells, n, side)


class CutCellBasis:
    """H1 basis tabulated on cut-cell quadrature points (cut_basis.py:12-150)."""

    def __init__(self, basis: Basis, quadrature: CutCellQuadrature):
        if not isinstance(basis, Basis):
            raise TypeError("CutCellBasis requires a Basis")
        if not isinstance(quadrature, CutCellQuadrature):
            raise TypeError("CutCellBasis requires CutCellQuadrature")
        if len(quadrature.cell_offsets) != basis.mesh.nelements + 1:
            raise ValueError("basis and cut quadrature have different cell counts")
        scalar = basis.elem.elem if isinstance(basis.elem, ElementVector) else None
        node_count = 10 if isinstance(scalar, ElementTetP2) else 4
        if not isinstance(scalar, (ElementTetP1, ElementTetP2)):
            raise NotImplementedError("CutCellBasis supports vector-wrapped TetP1 and straight-sided TetP2 "
                                      "in the MI355X build")
        components = basis.elem._dim
        positions = np.full(basis.mesh.nelements, -1, dtype=np.int64)
        positions[basis.tind] = np.arange(len(basis.tind))
        qcells = np.asarray(quadrature.cells, dtype=np.int64)
        local_cells = positions[qcells] if len(qcells) else np.zeros(0, dtype=np.int64)
        if np.any(local_cells < 0):
            bad = int(qcells[np.flatnonzero(local_cells < 0)[0]])
            raise ValueError(f"cut quadrature cell {bad} is absent from Basis")

        self.mesh = basis.mesh
        self.elem = basis.elem
        self.N = basis.N
        self.parent_basis = basis
        self.quadrature = quadrature
        self.cell_offsets = np.asarray(quadrature.cell_offsets, dtype=np.int64)
        self.cells = qcells
        self.points = np.asarray(quadrature.points, dtype=np.float64)
        self.reference_coordinates = np.asarray(quadrature.reference_points, dtype=np.float64)
        self.weights = np.asarray(quadrature.weights, dtype=np.float64)
        self.normal_vectors = np.asarray(quadrature.normals, dtype=np.float64)
        self.tind = np.flatnonzero(np.diff(self.cell_offsets) > 0)
        shape, ref_grad = tabulate(self.reference_coordinates.T, node_count == 10)
        self.shape = shape
        # physical gradients: reference gradients times the inverse of the
        # cell's affine Jacobian (columns = edge vectors from corner 0)
        corners = self.mesh.p[:, self.mesh.t[:4, self.tind]]  # (3, 4, A)
        J = np.moveaxis(corners[:, 1:4] - corners[:, [0]], -1, 0)  # (A, 3, 3)
        Jinv = np.linalg.inv(J) if len(J) else np.zeros((0, 3, 3))
        which = np.repeat(np.arange(len(self.tind)), np.diff(self.cell_offsets)[self.tind])
        self.gradients = np.einsum("pnk,pki->pni", ref_grad, Jinv[which]) if len(which) else \
            np.empty((0, node_count, 3))
        self.quadrature_dofs = (basis.element_dofs[:, local_cells].T.copy() if len(local_cells)
                                else np.empty((0, basis.element_dofs.shape[0]), dtype=np.int64))
        self.cell_dofs = np.zeros((basis.mesh.nelements, node_count, components), dtype=np.int64)
        self.cell_dofs[np.asarray(basis.tind, dtype=np.int64)] = basis.element_dofs.T.reshape(
            len(basis.tind), node_count, components)
        counts = np.diff(self.cell_offsets)[self.tind]
        self.active_cell_offsets = np.concatenate((np.array([0], dtype=np.int64),
                                                   np.cumsum(counts, dtype=np.int64)))
        self.active_cell_dofs = np.ascontiguousarray(self.cell_dofs[self.tind])
        # every point is an entity with one quadrature point
        self.dx = self.weights[:, None]
        self.tabulated_shape = self.shape[:, None, :]
        self.tabulated_gradients = self.gradients[:, None, :, :]
        self.global_coordinates = self.points[:, None, :]
        self.normals = self.normal_vectors[:, None, :]
        self.element_dofs = self.quadrature_dofs.T
        self.basis = ()
        for array in (self.shape, self.gradients, self.quadrature_dofs, self.tind, self.dx, self.tabulated_shape,
                      self.tabulated_gradients, self.global_coordinates, self.element_dofs, self.cell_dofs,
                      self.normals, self.active_cell_offsets, self.active_cell_dofs):
            array.flags.writeable = False

    @property
    def coefficient_shape(self):
        return (len(self.weights), 1)

    @property
    def npoints(self):
        return len(self.weights)

    @property
    def nelems(self):
        return len(self.tind)

    def cell_slice(self, cell):
        return self.quadrature.cell_slice(cell)

    def interpolate(self, coefficients):
        """Values (C, P) and gradients (C, 3, P) of a field at the cut points
        (squeezed for one component)."""
        coefficients = np.asarray(coefficients, dtype=np.float64)
        if coefficients.shape != (self.N,):
            raise ValueError(f"coefficients must have shape ({self.N},)")
        components = self.elem._dim
        nodes = self.shape.shape[1]
        local = coefficients[self.quadrature_dofs].reshape(self.npoints, nodes, components)
        value = np.einsum("qn,qnc->cq", self.shape, local)
        gradient = np.einsum("qnd,qnc->cdq", self.gradients, local)
        if components == 1:
            value, gradient = value[0], gradient[0]
        return DiscreteField(value, gradient)

    def integrate(self, values):
        values = np.asarray(values, dtype=np.float64)
        if values.shape != (self.npoints,):
            raise ValueError(f"cut integrand must have shape ({self.npoints},)")
        return float(self.weights @ values)
