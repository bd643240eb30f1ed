"""Cut-cell (ragged quadrature) assemblers on the GPU.

Drop-in counterparts of the reference's ``CutLinearFormAssembler``,
``CutBilinearFormAssembler`` (cpp/src/assembly/cut_form_assembler.cpp:22-257)
and ``CutCrossAssembler`` (cpp/src/assembly/cut_cross_assembler.cpp:16-181):
same constructor arguments, ``assemble`` signatures, properties and error
messages.  A cut cell owns the flat points [offsets[c], offsets[c+1]) with
tabulated shape values, physical gradients and weights per point.

The pattern comes from a device rectangular plan of the cells' node
connectivity (cells without points left out, as in the reference), and the
values from ``skfa_assemble_cut_bilinear`` / ``skfa_assemble_cut_linear``:
every row sums its cells in cell order, so results are deterministic.  The
reference colours cells for its threads; the GPU path needs no colouring,
so ``color_count`` is 0.  ``num_threads`` is accepted and ignored.

DOFs must number the components of a node consecutively
(dof = node * components + component), which every ``CutCellBasis`` and
``ElementVector`` basis does; other numberings raise ``ValueError``.
"""

from __future__ import annotations

import ctypes
import time

import numpy as np

from . import _native
from ._assembly import upload

_KIND = {"value": 0, "gradient": 1}
CUT_VALUE, CUT_GRADIENT, CUT_TENSOR = 1, 2, 4


def _i64(a):
    return np.ascontiguousarray(a, dtype=np.int64)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _node_connectivity(dofs):
    """(cells, nodes, C) DOFs -> (cells, nodes) node ids and the node count
    (max DOF // C + 1)."""
    if dofs.size and int(dofs.min()) < 0:
        raise ValueError("negative dof index")
    C = dofs.shape[2]
    base = dofs[:, :, 0]
    if C > 1 and dofs.size and (np.any(dofs != base[:, :, None] + np.arange(C)) or np.any(base % C)):
        raise ValueError("cut cell dofs must number the components of a node consecutively "
                         "(dof = node * components + component)")
    nnodes = int(dofs.max()) // C + 1 if dofs.size else 0
    return base // C, nnodes


class _CutPlan:
    """Device pattern of the cells that have points: a rectangular plan of
    their row and column node connectivity, their point ranges and the CSR
    arrays (the pattern half of the reference's constructors,
    cut_form_assembler.cpp:208-220, cut_cross_assembler.cpp:148-153)."""

    def __init__(self, row_nodes, nrow_nodes, Cr, col_nodes, ncol_nodes, Cc, offsets):
        import torch
        active = np.diff(offsets) > 0
        self.cells = int(np.count_nonzero(active))
        self.nrows, self.ncols = nrow_nodes * Cr, ncol_nodes * Cc
        self._h = None
        dev = _native.device()
        if self.cells == 0:
            self.nnz = 0
            self.indptr = torch.zeros(self.nrows + 1, dtype=torch.int64, device=dev)
            self.indices = torch.zeros(1, dtype=torch.int32, device=dev)
            self.offsets = None
            self.repeats_column = False
            return
        rfull, cfull = row_nodes[active], col_nodes[active]
        # The plan runs on the nodes that the cut cells touch, renumbered by
        # increasing maps: rows keep their order and columns stay sorted, so
        # the compact CSR values are the full ones, and no wave is spent on
        # the (many) rows without cut cells.
        self.row_used, rloc = np.unique(rfull, return_inverse=True)
        self.col_used, cloc = np.unique(cfull, return_inverse=True)
        rconn = np.ascontiguousarray(np.asarray(rloc).reshape(rfull.shape), dtype=np.int32)
        cconn = np.ascontiguousarray(np.asarray(cloc).reshape(cfull.shape), dtype=np.int32)
        s = np.sort(cconn, axis=1)
        self.repeats_column = bool(np.any(s[:, 1:] == s[:, :-1]))
        self.offsets = upload(np.concatenate([offsets[:-1][active], offsets[-1:]]).astype(np.int64))
        self._rconn, self._cconn = upload(rconn), upload(cconn)
        L = _native.lib()
        h = ctypes.c_void_p()
        _native.check(L.skfa_cross_plan_create(
            _native.ptr(self._rconn), rconn.shape[1], len(self.row_used), Cr, _native.ptr(self._cconn),
            cconn.shape[1], len(self.col_used), Cc, self.cells, _native.stream(), ctypes.byref(h)))
        self._h = h
        info = (ctypes.c_int64 * 6)()
        _native.check(L.skfa_cross_plan_info(h, info))
        self.nnz = int(info[2])
        ip_c = torch.empty(len(self.row_used) * Cr + 1, dtype=torch.int64, device=dev)
        ix_c = torch.empty(max(self.nnz, 1), dtype=torch.int32, device=dev)
        _native.check(L.skfa_cross_plan_export_csr(h, _native.ptr(ip_c), _native.ptr(ix_c), _native.stream()))
        self.compact_rows = len(self.row_used) * Cr
        # full numbering: row DOF used[i] * Cr + c holds compact row i * Cr + c
        ip_c, ix_c = ip_c.cpu().numpy(), ix_c[:self.nnz].cpu().numpy().astype(np.int64)
        self.row_dofs = (self.row_used[:, None] * Cr + np.arange(Cr)).ravel()
        counts = np.zeros(self.nrows, dtype=np.int64)
        counts[self.row_dofs] = np.diff(ip_c)
        ip = np.zeros(self.nrows + 1, dtype=np.int64)
        np.cumsum(counts, out=ip[1:])
        ix = self.col_used[ix_c // Cc] * Cc + ix_c % Cc
        self.indptr = torch.from_numpy(ip).to(dev)
        self.indices = torch.from_numpy(np.ascontiguousarray(ix if len(ix) else np.zeros(1, np.int64),
                                                             dtype=np.int32)).to(dev)

    @property
    def handle(self):
        return self._h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                _native.load_library().skfa_cross_plan_destroy(h)
            except Exception:
                pass
            self._h = None


class _CutFormData:
    """Validated per-point data of one cut basis (CutFormData,
    cut_form_assembler.cpp:22-65), uploaded once."""

    def __init__(self, dofs, offsets, shape, gradients, weights):
        dofs, offsets = _i64(dofs), _i64(offsets)
        shape, gradients, weights = _f64(shape), _f64(gradients), _f64(weights)
        if dofs.ndim != 3:
            raise ValueError("cell dofs must have shape (cells, nodes, components)")
        cells, nodes, components = dofs.shape
        if offsets.ndim != 1 or offsets.shape[0] != cells + 1:
            raise ValueError("cell offsets must have shape (cells + 1,)")
        if shape.ndim != 2 or shape.shape[1] != nodes:
            raise ValueError("cut shape values have an invalid shape")
        points = shape.shape[0]
        if gradients.ndim != 3 or gradients.shape[0] != points or gradients.shape[1] != nodes:
            raise ValueError("cut shape gradients have an invalid shape")
        if weights.ndim != 1 or weights.shape[0] != points:
            raise ValueError("cut weights have an invalid shape")
        if offsets[0] != 0 or offsets[-1] != points:
            raise ValueError("cell offsets do not span cut points")
        if np.any(np.diff(offsets) < 0):
            raise ValueError("cell offsets must be nondecreasing")
        if not 1 <= gradients.shape[2] <= 3:
            raise ValueError("cut shape gradients must have 1 to 3 spatial dimensions")
        self._cells, self._nodes, self._components = cells, nodes, components
        self._points, self._dimension = points, gradients.shape[2]
        node_conn, nnodes = _node_connectivity(dofs)
        self._ndofs = nnodes * components
        self._plan = _CutPlan(node_conn, nnodes, components, node_conn, nnodes, components, offsets)
        self._dev = {k: upload(v) for k, v in
                     (("shape", shape), ("gradients", gradients), ("weights", weights))} if points else None

    @property
    def ndofs(self):
        return self._ndofs


class CutLinearFormAssembler(_CutFormData):
    """Cut-cell load vector (CutLinearFormAssembler,
    cut_form_assembler.cpp:67-134): ``assemble(value=None, gradient=None,
    num_threads=0)`` returns (vector over ndofs, seconds)."""

    def assemble(self, value=None, gradient=None, num_threads=0):
        C, D, P = self._components, self._dimension, self._points
        vflag = gflag = 0
        if value is not None:
            value = _f64(value)
            vconst = value.ndim == 1 and value.shape[0] == C
            if not vconst and (value.ndim != 2 or value.shape != (P, C)):
                raise ValueError("cut value coefficient has an invalid shape")
            vflag = 0 if vconst else 1
        if gradient is not None:
            gradient = _f64(gradient)
            gconst = gradient.ndim == 2 and gradient.shape == (C, D)
            if not gconst and (gradient.ndim != 3 or gradient.shape != (P, C, D)):
                raise ValueError("cut gradient coefficient has an invalid shape")
            gflag = 0 if gconst else 1
        if value is None and gradient is None:
            raise ValueError("at least one coefficient is required")
        import torch
        t0 = time.perf_counter()
        p = self._plan
        out = torch.zeros(max(p.compact_rows if p.cells else 0, 1), dtype=torch.float64, device=_native.device())
        if p.cells:
            dv = upload(value) if value is not None else None
            dg = upload(gradient) if gradient is not None else None
            d = self._dev
            _native.check(_native.lib().skfa_assemble_cut_linear(
                p.handle, _native.ptr(p.offsets), _native.ptr(d["weights"]), _native.ptr(d["shape"]),
                _native.ptr(d["gradients"]), D, _native.ptr(dv), vflag, _native.ptr(dg), gflag, _native.ptr(out),
                _native.stream()))
        result = np.zeros(self._ndofs)
        if p.cells:
            # compact row i * C + c is DOF used[i] * C + c
            result[p.row_dofs] = out[:p.compact_rows].cpu().numpy()
        return result, time.perf_counter() - t0

    @property
    def cell_count(self):
        return self._cells

    @property
    def point_count(self):
        return self._points


class _CutMatrix:
    """Pattern on the host and the values buffer that ``assemble`` refills in
    place (the reference's ``values`` view keeps its identity across calls)."""

    def _init_matrix(self, plan):
        ip = plan.indptr.cpu().numpy()
        self._indptr = ip
        self._indices = plan.indices[:plan.nnz].cpu().numpy().astype(np.int64)
        self._values = np.zeros(plan.nnz)
        self.values_device = None

    def _finish(self, dev_values):
        self.values_device = dev_values
        if len(self._values):
            self._values[:] = dev_values.cpu().numpy()

    @property
    def indptr(self):
        return self._indptr

    @property
    def indices(self):
        return self._indices

    @property
    def values(self):
        return self._values

    @property
    def color_count(self):
        return 0


class CutBilinearFormAssembler(_CutFormData, _CutMatrix):
    """Cut-cell square form (CutBilinearFormAssembler,
    cut_form_assembler.cpp:136-239): ``assemble(value=None, gradient=None,
    num_threads=0)`` with per-point coefficients (points,) refills
    ``values`` on the pattern ``indptr`` / ``indices``."""

    def __init__(self, dofs, offsets, shape, gradients, weights):
        _CutFormData.__init__(self, dofs, offsets, shape, gradients, weights)
        self._init_matrix(self._plan)

    def _coefficient(self, obj, name):
        if obj is None:
            return None
        a = _f64(obj)
        if a.ndim != 1 or a.shape[0] != self._points:
            raise ValueError(f"cut {name} coefficient has an invalid shape")
        return a

    def assemble(self, value=None, gradient=None, num_threads=0):
        value = self._coefficient(value, "value")
        gradient = self._coefficient(gradient, "gradient")
        if value is None and gradient is None:
            raise ValueError("at least one coefficient is required")
        import torch
        p = self._plan
        vals = torch.zeros(max(p.nnz, 1), dtype=torch.float64, device=_native.device())
        if p.cells:
            terms = (CUT_VALUE if value is not None else 0) | (CUT_GRADIENT if gradient is not None else 0)
            dv = upload(value) if value is not None else None
            dg = upload(gradient) if gradient is not None else None
            d = self._dev
            _native.check(_native.lib().skfa_assemble_cut_bilinear(
                p.handle, _native.ptr(p.offsets), _native.ptr(d["weights"]), _native.ptr(d["shape"]),
                _native.ptr(d["gradients"]), self._dimension, _native.ptr(d["shape"]), _native.ptr(d["gradients"]),
                self._dimension, terms, _native.ptr(dv), _native.ptr(dg), 0, 0, None, int(p.repeats_column),
                _native.ptr(vals), _native.stream()))
        self._finish(vals[:p.nnz])


class CutCrossAssembler(_CutMatrix):
    """Cut-cell rectangular form between a test (row) and a trial (column)
    basis on the same cut points (CutCrossAssembler,
    cut_cross_assembler.cpp:16-165): ``assemble(coefficient=None,
    row_kind="value", column_kind="value", num_threads=0)``."""

    def __init__(self, row_dofs, column_dofs, offsets, row_shape, column_shape, weights, row_gradients,
                 column_gradients):
        rd, cd, offsets = _i64(row_dofs), _i64(column_dofs), _i64(offsets)
        rs, cs, w = _f64(row_shape), _f64(column_shape), _f64(weights)
        rg, cg = _f64(row_gradients), _f64(column_gradients)
        if rd.ndim != 3 or cd.ndim != 3 or rd.shape[0] != cd.shape[0]:
            raise ValueError("cut cross cell dofs have incompatible shapes")
        cells, nr, Cr = rd.shape
        _, nc, Cc = cd.shape
        if offsets.ndim != 1 or offsets.shape[0] != cells + 1:
            raise ValueError("cut cross offsets have an invalid shape")
        if rs.ndim != 2 or rs.shape[1] != nr or cs.ndim != 2 or cs.shape[0] != rs.shape[0] or cs.shape[1] != nc:
            raise ValueError("cut cross shape values have an invalid shape")
        points = rs.shape[0]
        if w.ndim != 1 or w.shape[0] != points:
            raise ValueError("cut cross weights have an invalid shape")
        if (rg.ndim != 3 or rg.shape[0] != points or rg.shape[1] != nr or cg.ndim != 3 or cg.shape[0] != points
                or cg.shape[1] != nc):
            raise ValueError("cut cross gradients have an invalid shape")
        if offsets[0] != 0 or offsets[-1] != points:
            raise ValueError("cut cross offsets do not span points")
        if np.any(np.diff(offsets) < 0):
            raise ValueError("cell offsets must be nondecreasing")
        if not (1 <= rg.shape[2] <= 3 and 1 <= cg.shape[2] <= 3):
            raise ValueError("cut shape gradients must have 1 to 3 spatial dimensions")
        self._points, self._Cr, self._Cc = points, Cr, Cc
        self._Dr, self._Dc = rg.shape[2], cg.shape[2]
        rnodes, nrn = _node_connectivity(rd)
        cnodes, ncn = _node_connectivity(cd)
        self._rows, self._columns = nrn * Cr, ncn * Cc
        self._plan = _CutPlan(rnodes, nrn, Cr, cnodes, ncn, Cc, offsets)
        self._dev = {k: upload(v) for k, v in (("rs", rs), ("cs", cs), ("w", w), ("rg", rg), ("cg", cg))} \
            if points else None
        self._init_matrix(self._plan)

    @staticmethod
    def _gradient_kind(kind):
        if kind not in _KIND:
            raise ValueError("cut cross kind must be value or gradient")
        return kind == "gradient"

    def assemble(self, coefficient=None, row_kind="value", column_kind="value", num_threads=0):
        rg, cg = self._gradient_kind(row_kind), self._gradient_kind(column_kind)
        tensor = False
        if coefficient is not None:
            coefficient = _f64(coefficient)
            expect = (self._points, self._Cr) + ((self._Dr,) if rg else ()) + (self._Cc,) + \
                ((self._Dc,) if cg else ())
            tensor = coefficient.ndim == 3 + int(rg) + int(cg) and coefficient.shape == expect
            if not tensor and (coefficient.ndim != 1 or coefficient.shape[0] != self._points):
                raise ValueError("cut cross coefficient has an invalid shape")
        if not tensor and (self._Cr != self._Cc or rg != cg or (rg and self._Dr != self._Dc)):
            raise ValueError("scalar cut cross coefficient has incompatible bases")
        import torch
        p = self._plan
        vals = torch.zeros(max(p.nnz, 1), dtype=torch.float64, device=_native.device())
        if p.cells:
            d = self._dev
            dc = upload(coefficient) if coefficient is not None else None
            if tensor:
                terms, dv, dg, T = CUT_TENSOR, None, None, dc
            else:
                terms = CUT_GRADIENT if rg else CUT_VALUE
                dv, dg, T = (None, dc, None) if rg else (dc, None, None)
            _native.check(_native.lib().skfa_assemble_cut_bilinear(
                p.handle, _native.ptr(p.offsets), _native.ptr(d["w"]), _native.ptr(d["rs"]), _native.ptr(d["rg"]),
                self._Dr, _native.ptr(d["cs"]), _native.ptr(d["cg"]), self._Dc, terms, _native.ptr(dv),
                _native.ptr(dg), _KIND[row_kind], _KIND[column_kind], _native.ptr(T), int(p.repeats_column),
                _native.ptr(vals), _native.stream()))
        self._finish(vals[:p.nnz])

    @property
    def rows(self):
        return self._rows

    @property
    def columns(self):
        return self._columns
