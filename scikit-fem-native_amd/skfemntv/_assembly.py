"""Host side of the native assemblers: plans, coefficient upload, kernel calls.

``NativeBilinearForm`` / ``NativeCrossBilinearForm`` / ``NativeLinearForm``
mirror python/skfemntv/bilinear_form.py:13-99,102-178,336-367 and
linear_form.py:17-107: same result conventions (the scalar-kind form
returns the same scipy CSR object on every call, overwritten; the tensor
path returns a fresh copy; the linear form returns a vector and
diagnostics), with the work done by libskfem_amd on the GPU.
"""

from __future__ import annotations

import ctypes
import time
from dataclasses import dataclass

import numpy as np
from scipy.sparse import csr_matrix

from . import _native
from ._symbols import UnsupportedNativeForm


# ------------------------------------------------------------------ plans
_TUNE_KEYS = {"xcd_chunk": 1, "lds_atomic": 2, "placement": 3, "streaming": 4, "element_table": 5,
              "row_split": 6, "mfma": 7, "row_form": 8}


class Plan:
    """Device CSR pattern + gather plan of one (basis, components) pair."""

    def __init__(self, node_conn_dev, E, n, nnodes, components, lds_budget=None, min_waves=2048, row_lanes=None):
        import torch
        from . import autotune
        if lds_budget is None:
            # the layout the autotuner chose for a plan of this shape, else
            # the default (scalar fields 12 KiB, 3-component fields 16 KiB)
            lds_budget, cached_lanes = autotune.cached_layout(E, n, nnodes, components,
                                                              16 * 1024 if components == 3 else 12 * 1024)
            if row_lanes is None:
                row_lanes = cached_lanes
        self.row_lanes = int(row_lanes or 0)
        self._conn = node_conn_dev
        self._h = None
        self.E, self.n, self.nnodes, self.components = E, n, nnodes, components
        self.min_waves = int(min_waves)
        self.tuning = {"xcd_chunk": 16, "lds_atomic": 1, "placement": 2, "streaming": 1, "element_table": -1,
                       "row_split": 0, "mfma": -1, "row_form": -1}
        self._create(lds_budget)
        L = _native.lib()
        dev = _native.device()
        self.indptr = torch.empty(self.nrows + 1, dtype=torch.int64, device=dev)
        self.indices = torch.empty(max(self.nnz, 1), dtype=torch.int32, device=dev)
        _native.check(L.skfa_plan_export_csr(self._h, _native.ptr(self.indptr), _native.ptr(self.indices),
                                             _native.stream()))
        self._host = None

    def _create(self, lds_budget):
        L = _native.lib()
        h = ctypes.c_void_p()
        _native.check(L.skfa_plan_create_layout(_native.ptr(self._conn), self.E, self.n, self.nnodes,
                                                self.components, int(lds_budget), self.min_waves, self.row_lanes,
                                                _native.stream(), ctypes.byref(h)))
        old = self._h
        self._h = h
        self.lds_budget = int(lds_budget)
        info = (ctypes.c_int64 * 9)()
        _native.check(L.skfa_plan_info(h, info))
        self.nrows, self.nnz, self.nnz_node = int(info[0]), int(info[1]), int(info[2])
        self.max_row_len, self.max_inc = int(info[3]), int(info[4])
        self.waves, self.steps, self.rec_bytes, self.buckets = int(info[5]), int(info[6]), int(info[7]), int(info[8])
        for name, value in self.tuning.items():
            _native.check(L.skfa_plan_set_tuning(h, _TUNE_KEYS[name], int(value)))
        if old is not None and old.value:
            import torch
            torch.cuda.synchronize()   # no launch may still use the old plan
            L.skfa_plan_destroy(old)

    def rebuild(self, lds_budget, row_lanes=None):
        """Re-layout the gather waves for another LDS span budget (and,
        given, another number of lanes per node row: 0 default, 4 the
        16-row waves of the vector-P2 MFMA gather).  The pattern (indptr,
        indices) is the same for every layout; the tuning settings are kept."""
        nnz = self.nnz
        if row_lanes is not None:
            self.row_lanes = int(row_lanes)
        self._create(lds_budget)
        if self.nnz != nnz:
            raise RuntimeError("plan rebuild changed the pattern")

    @property
    def handle(self):
        return self._h

    def set_tuning(self, **settings):
        """Launch tuning (library key in parentheses).

        Bit-identical results for every value: ``xcd_chunk`` (1, waves per
        XCD chunk), ``placement`` (3: 0 plain, 1 XCD segments, 2 XCD
        chunks), ``streaming`` (4), ``row_split`` (6: how the lanes of one
        row share it: 0 incidences, 1 entries, 2 columns).  Different
        rounding only: ``row_split`` 3 (concurrent incidences),
        ``element_table`` (5: -1 auto, 0, 1), ``mfma`` (7: 0 off, 1 all
        waves, 2 waves of at most 16 rows) and ``row_form`` (8: P2 rows and
        vector-P2 blocks from the barycentric structure, 1, or from the
        quadrature tables, 0; -1 auto).  ``lds_atomic`` (2) accepts only 1:
        the span adds are always ds_add_f64."""
        L = _native.lib()
        for name, value in settings.items():
            if name not in _TUNE_KEYS:
                raise ValueError(f"unknown tuning setting {name!r}")
            v = int(value)
            _native.check(L.skfa_plan_set_tuning(self._h, _TUNE_KEYS[name], v))
            self.tuning[name] = v

    def host_csr_arrays(self):
        """(indptr, indices) on the host, int32 when they fit (as scipy does)."""
        if self._host is None:
            ip = self.indptr.cpu().numpy()
            ix = self.indices[:self.nnz].cpu().numpy()
            if self.nnz < 2 ** 31 - 1:
                ip = ip.astype(np.int32)
            else:
                ix = ix.astype(np.int64)
            self._host = (ip, ix)
        return self._host

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                _native.load_library().skfa_plan_destroy(h)
            except Exception:
                pass
            self._h = None


def basis_plan(basis):
    """Plan of a Basis (cached on the basis)."""
    plan = getattr(basis, "_plan", None)
    if plan is None:
        d = basis.device_data()
        C = basis.elem._dim
        # the LDS span budget comes from the autotuner's cache for this mesh
        # shape, else the default of Plan
        plan = Plan(d["node_conn"], basis.mesh.nelements, d["node_conn"].shape[1], d["nnodes"], C)
        basis._plan = plan
    return plan


# ------------------------------------------------------------------ coefficients
def field_layout(coefficient, shape, name, trailing=()):
    """Broadcast a coefficient to shape+trailing and decide whether it is
    constant per element.  Returns (host array, per_qp) where the array is
    (E,)+trailing or (E,Q)+trailing, C-contiguous float64."""
    arr = np.asarray(coefficient)
    if arr.dtype != np.float64:
        arr = arr.astype(np.float64)
    target = tuple(shape) + tuple(trailing)
    try:
        full = np.broadcast_to(arr, target)
    except ValueError as error:
        raise ValueError(f"{name} coefficient must broadcast to {target}") from error
    E, Q = shape
    if Q == 1 or E == 0 or full.strides[1] == 0:
        return np.ascontiguousarray(full[:, 0]), False
    first = full[:, :1]
    if np.array_equal(full, np.broadcast_to(first, full.shape)):
        return np.ascontiguousarray(full[:, 0]), False
    return np.ascontiguousarray(full), True


def upload(host):
    """Host array -> device tensor.  A read-only contiguous array (a
    broadcast or moved-axes view of the caller's data) is wrapped without a
    host copy: torch only reads it during the transfer."""
    import torch
    import warnings
    arr = np.ascontiguousarray(host)
    if arr.flags.writeable:
        return torch.from_numpy(arr).to(_native.device())
    with warnings.catch_warnings():
        warnings.filterwarnings("ignore", message="The given NumPy array is not writable")
        return torch.from_numpy(arr).to(_native.device())


def to_host(dev, out=None):
    """Device tensor -> numpy array through pinned host memory (the copy runs
    at full PCIe rate instead of through pageable staging).  ``out`` is a
    pinned torch tensor to fill; by default a new one (torch's pinned
    allocator reuses freed blocks) backs the returned array."""
    import torch
    host = torch.empty(dev.shape, dtype=dev.dtype, pin_memory=True) if out is None else out
    host.copy_(dev, non_blocking=True)
    torch.cuda.current_stream().synchronize()
    return host.numpy()


def host_copy(arr):
    """A fresh copy of a host array, made with torch's multi-threaded copy
    (2-3x faster than numpy's single-threaded one on the 10^8-entry index
    arrays of a fresh CSR)."""
    import torch
    return torch.from_numpy(np.ascontiguousarray(arr)).clone().numpy()


def to_per_qp(dev_tensor, Q):
    """(E,)+t -> (E,Q)+t on the device."""
    shape = dev_tensor.shape
    return dev_tensor.unsqueeze(1).expand((shape[0], Q) + tuple(shape[1:])).contiguous()


def basis_scalar_plan(basis):
    """Scalar (C = 1) plan over a vector basis's node connectivity, for
    value / gradient forms written as 3x3 diagonal node blocks."""
    plan = getattr(basis, "_plan_scalar", None)
    if plan is None:
        d = basis.device_data()
        plan = Plan(d["node_conn"], basis.mesh.nelements, d["node_conn"].shape[1], d["nnodes"], 1)
        basis._plan_scalar = plan
    return plan


def assemble_bilinear_device(basis, plan, values_dev, *, value=None, gradient=None, symmetric_gradient=None,
                             divergence=None, tensor=None, per_qp=False, precision="fp64", symmetric_tensor=False,
                             noalloc=False, load=None):
    """Launch skfa_assemble_bilinear with device coefficient tensors
    (``noalloc``: skfa_assemble_bilinear_noalloc, which keeps its scratch in
    a fixed device array of the library instead of allocating; ``load`` =
    (f, out) or (f, out, value_map, row_map): skfa_assemble_bilinear_load,
    which also writes the load vector of the constant source f (one device
    double) into out, scalar fields only, and with the int32 maps writes
    values and load through them, dropping entries mapped to -1)."""
    d = basis.device_data()
    L = _native.lib()
    iso = not basis._affine
    flags = _native.FLAG_SYMMETRIC_TENSOR if symmetric_tensor else 0
    if (plan.components == 3 and basis.elem._dim == 3 and not iso and not per_qp and symmetric_gradient is None
            and divergence is None and tensor is None):
        # value / gradient kinds of a 3-component field: K = [c == d] K_scalar,
        # assembled by the scalar gather and written as diagonal node blocks
        plan = basis_scalar_plan(basis)
        flags |= _native.FLAG_VECTOR3_DIAGONAL
    tables = d["tables_iso"] if iso else d["tables"]
    if d.get("rtab_barycentric"):
        flags |= _native.FLAG_BARYCENTRIC_TABLE
    prec = {"fp64": 0, "fp32": 1}[precision]
    if plan.row_lanes == 4 and (iso or per_qp or d["nqg"] != 4 or d["nqm"] > 4):
        # the 16-row MFMA layout covers element-constant forms on affine
        # elements only: other forms run on the default layout
        plan.rebuild(plan.lds_budget, row_lanes=0)
    if prec == 1 and (iso or per_qp):
        raise UnsupportedNativeForm("FP32 products are available for affine elements with element-constant "
                                    "coefficients")
    entry = L.skfa_assemble_bilinear_noalloc if noalloc else L.skfa_assemble_bilinear
    if load is not None:
        if noalloc:
            raise ValueError("the fused load vector has no no-allocation entry point")
        entry = L.skfa_assemble_bilinear_load
    if load is not None and len(load) == 2:
        load = (load[0], load[1], None, None)
    extra = () if load is None else tuple(_native.ptr(a) for a in load)

    def launch():
        _native.check(entry(
            plan.handle, _native.ptr(d["xyz"]), d["xyz"].shape[0], _native.ptr(d["vconn"]),
            _native.ptr(d["gconn"]) if iso else None, d["gnodes"] if iso else 4,
            _native.ptr(tables), len(basis.W), _native.ptr(d["rtab"]), _native.ptr(d["rtab32"]), d["nqg"],
            d["nqm"], _native.ptr(value), _native.ptr(gradient), _native.ptr(symmetric_gradient),
            _native.ptr(divergence), _native.ptr(tensor), int(per_qp), prec, flags, _native.ptr(values_dev),
            *extra, _native.stream()))

    if not noalloc and not iso and not per_qp:
        # element-constant fast path: pick the launch shape on first use
        from . import autotune
        vector_mfma = plan.components == 3 and plan.n == 10 and d["nqg"] == 4 and d["nqm"] <= 4
        autotune.maybe_tune(plan, launch, vector_mfma)
    launch()


def _prepare_scalar_kinds(basis, coefficients):
    """Host arrays -> device tensors with one common layout."""
    E, Q = basis.coefficient_shape
    laid = {}
    for name, coef in coefficients.items():
        if coef is None:
            continue
        trailing = (3, 3) if name == "tensor" else ()
        laid[name] = field_layout(coef, (E, Q), name, trailing)
    per_qp = any(p for _, p in laid.values())
    dev = {}
    for name, (arr, p) in laid.items():
        t = upload(arr)
        if per_qp and not p:
            t = to_per_qp(t, Q)
        dev[name] = t
    return dev, per_qp


class NativeBilinearForm:
    """Value / gradient / symmetric-gradient / divergence assembly of one
    basis (bilinear_form.py:13-99).  ``assemble`` returns the same CSR object
    on every call; its data are overwritten."""

    def __init__(self, basis, *, memory_limit_bytes=None, memory_safety_factor=1.25):
        self.basis = basis
        self._cut = None
        if hasattr(basis, "cell_offsets"):
            self._init_cut(basis, memory_limit_bytes, memory_safety_factor)
            return
        self.memory_estimate = estimate_memory_upper(basis)
        enforce_memory_budget(self.memory_estimate, memory_limit_bytes, memory_safety_factor)
        self.plan = basis_plan(basis)
        import torch
        self._values = torch.zeros(max(self.plan.nnz, 1), dtype=torch.float64, device=_native.device())
        ip, ix = self.plan.host_csr_arrays()
        # the returned CSR object's values live in pinned host memory, which
        # the device fills directly on every call
        self._pinned = torch.zeros(max(self.plan.nnz, 1), dtype=torch.float64, pin_memory=True)
        self._matrix = csr_matrix((self._pinned[:self.plan.nnz].numpy(), ix, ip),
                                  shape=(self.plan.nrows, self.plan.nrows), copy=False)
        self._matrix.resize((basis.N, basis.N))
        if not np.shares_memory(self._matrix.data, self._pinned.numpy()):
            self._pinned = None
        self._coefficient_shape = basis.coefficient_shape
        self.last_kernel_seconds = 0.

    def _init_cut(self, basis, memory_limit_bytes, memory_safety_factor):
        """Cut-cell basis (bilinear_form.py:33-41): the ragged-quadrature
        assembler over the cells that have points."""
        from .cut import CutBilinearFormAssembler
        self.memory_estimate = estimate_cut_memory(basis.active_cell_dofs, basis.active_cell_dofs, basis.N,
                                                   basis.N, "bilinear")
        enforce_memory_budget(self.memory_estimate, memory_limit_bytes, memory_safety_factor)
        self._cut = CutBilinearFormAssembler(basis.active_cell_dofs, basis.active_cell_offsets, basis.shape,
                                             basis.gradients, basis.weights)
        c = self._cut
        self._matrix = csr_matrix((c.values, c.indices, c.indptr), shape=(c.ndofs, c.ndofs), copy=False)
        self._matrix.resize((basis.N, basis.N))
        self._shares = np.shares_memory(self._matrix.data, c.values)
        self._coefficient_shape = basis.coefficient_shape
        self.last_kernel_seconds = 0.

    def _assemble_cut(self, value, gradient, symmetric_gradient, divergence):
        if symmetric_gradient is not None or divergence is not None:
            raise ValueError("cut-basis symmetric-gradient and divergence assembly is not supported")
        coefs = []
        for name, c in (("value", value), ("gradient", gradient)):
            if c is None:
                coefs.append(None)
                continue
            c = np.asarray(c, dtype=np.float64)
            try:
                c = np.broadcast_to(c, self._coefficient_shape)
            except ValueError as error:
                raise ValueError(f"{name} coefficient must broadcast to {self._coefficient_shape}") from error
            coefs.append(np.ascontiguousarray(c).reshape(-1))
        self._cut.assemble(coefs[0], coefs[1])
        if not self._shares:
            np.copyto(self._matrix.data, self._cut.values)
        return self._matrix

    def assemble(self, *, value=None, gradient=None, symmetric_gradient=None, divergence=None, num_threads=0,
                 precision="fp64"):
        coefs = {"value": value, "gradient": gradient, "symmetric_gradient": symmetric_gradient,
                 "divergence": divergence}
        if all(v is None for v in coefs.values()):
            raise ValueError("at least one coefficient is required")
        if self._cut is not None:
            return self._assemble_cut(value, gradient, symmetric_gradient, divergence)
        dev, per_qp = _prepare_scalar_kinds(self.basis, coefs)
        self.assemble_device(per_qp=per_qp, precision=precision, **dev)
        if self._pinned is not None:
            to_host(self._values[:self.plan.nnz], self._pinned[:self.plan.nnz])
        else:
            np.copyto(self._matrix.data, to_host(self._values[:self.plan.nnz]))
        return self._matrix

    def assemble_device(self, *, value=None, gradient=None, symmetric_gradient=None, divergence=None,
                        tensor=None, per_qp=False, precision="fp64", out=None, load=None):
        """Assemble into a device tensor (default: this form's buffer) with
        coefficient tensors already on the device; returns the values.
        ``load`` = (f, out): also the load vector of the constant source f
        (a one-element device tensor) into out (scalar fields), fused with
        the gather where it can be (skfa_assemble_bilinear_load)."""
        vals = self._values if out is None else out
        assemble_bilinear_device(self.basis, self.plan, vals, value=value, gradient=gradient,
                                 symmetric_gradient=symmetric_gradient, divergence=divergence, tensor=tensor,
                                 per_qp=per_qp, precision=precision, load=load)
        return vals


_CROSS_KIND = {"value": 0, "gradient": 1}
_CROSS_MODE = {"scalar": 0, "row_divergence": 1, "column_divergence": 2, "tensor": 3}


class CrossPlan:
    """Device rectangular pattern between a test (row) and a trial (column)
    basis on one mesh (CrossBilinearAssembler's constructor,
    cross_bilinear_assembler.cpp:19-54,318-341)."""

    def __init__(self, test_basis, trial_basis):
        import torch
        L = _native.lib()
        dr, dc = test_basis.device_data(), trial_basis.device_data()
        h = ctypes.c_void_p()
        _native.check(L.skfa_cross_plan_create(
            _native.ptr(dr["node_conn"]), dr["node_conn"].shape[1], dr["nnodes"], test_basis.elem._dim,
            _native.ptr(dc["node_conn"]), dc["node_conn"].shape[1], dc["nnodes"], trial_basis.elem._dim,
            test_basis.mesh.nelements, _native.stream(), ctypes.byref(h)))
        self._h = h
        info = (ctypes.c_int64 * 6)()
        _native.check(L.skfa_cross_plan_info(h, info))
        self.nrows, self.ncols, self.nnz, self.nnz_node, self.max_row_len, self.max_inc = (int(v) for v in info)
        dev = _native.device()
        self.indptr = torch.empty(self.nrows + 1, dtype=torch.int64, device=dev)
        self.indices = torch.empty(max(self.nnz, 1), dtype=torch.int32, device=dev)
        _native.check(L.skfa_cross_plan_export_csr(h, _native.ptr(self.indptr), _native.ptr(self.indices),
                                                   _native.stream()))
        self._host = None

    @property
    def handle(self):
        return self._h

    host_csr_arrays = Plan.host_csr_arrays

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                _native.load_library().skfa_cross_plan_destroy(h)
            except Exception:
                pass
            self._h = None


def cross_tables(test_basis, trial_basis):
    """FP64 tables of the cross kernel: w | test phi | test dphi | trial phi |
    trial dphi | geometry dphi (isoparametric meshes)."""
    from .basis import tabulate
    X, W = test_basis.X, test_basis.W
    rs, rg = tabulate(X, test_basis._quadratic)
    cs, cg = tabulate(X, trial_basis._quadratic)
    _, gg = tabulate(X, test_basis.mesh.t.shape[0] == 10)
    return np.concatenate([np.ravel(W), rs.ravel(), rg.ravel(), cs.ravel(), cg.ravel(), gg.ravel()]).astype(
        np.float64)


def cross_reduced_table(test_basis, trial_basis, row_kind, column_kind):
    """rt[a][b][k][l] = sum_q w_q h_ak(q) h_bl(q) with h = phi (value side) or
    dphi/dxi (gradient side), FP64 (nr, nc, [3], [3]) flattened: the element
    matrix of affine elements with element-constant coefficients depends on
    the rule only through these sums."""
    from .basis import tabulate
    X, W = test_basis.X, test_basis.W
    rs, rg = tabulate(X, test_basis._quadratic)
    cs, cg = tabulate(X, trial_basis._quadratic)
    hr = rg if row_kind == "gradient" else rs[:, :, None]
    hc = cg if column_kind == "gradient" else cs[:, :, None]
    return np.einsum("q,qak,qbl->abkl", W, hr, hc).astype(np.float64).ravel()


def _same_mesh(a, b):
    return a is b or (a.p.shape == b.p.shape and a.t.shape == b.t.shape and np.array_equal(a.p, b.p)
                      and np.array_equal(a.t, b.t))


class NativeCrossBilinearForm:
    """Assembly between a trial and a test basis (bilinear_form.py:100-362);
    every call returns a fresh CSR copy of shape (test N, trial N).

    Two volume bases on one tetrahedral mesh with one quadrature rule; rows
    follow the test basis, columns the trial basis.  The gradient x gradient
    tensor of one scalar basis with itself runs on the square row-gather
    kernels; every other pairing runs the rectangular cross kernel."""

    def __init__(self, test_basis, trial_basis, *, memory_limit_bytes=None, memory_safety_factor=1.25):
        self._cutx = None
        facet = hasattr(test_basis, "parent_elements") and hasattr(trial_basis, "parent_elements")
        if facet or (hasattr(test_basis, "cell_offsets") and hasattr(trial_basis, "cell_offsets")):
            self._init_cut(test_basis, trial_basis, memory_limit_bytes, memory_safety_factor, facet)
            return
        if test_basis is not trial_basis:
            if len(test_basis.W) != len(trial_basis.W) or test_basis.mesh.nelements != trial_basis.mesh.nelements:
                raise ValueError("trial and test quadrature shapes must match")
            if not np.allclose(test_basis.W, trial_basis.W):
                raise ValueError("trial and test quadrature weights must match")
            if not (_same_mesh(test_basis.mesh, trial_basis.mesh) and np.allclose(test_basis.X, trial_basis.X)):
                raise ValueError("trial and test quadrature coordinates must match")
        self.test_basis, self.trial_basis = test_basis, trial_basis
        self.basis = test_basis
        self.memory_estimate = (estimate_memory_upper(test_basis) if test_basis is trial_basis
                                else estimate_cross_memory(test_basis, trial_basis))
        enforce_memory_budget(self.memory_estimate, memory_limit_bytes, memory_safety_factor)
        self.shape = (test_basis.N, trial_basis.N)
        self.coefficient_shape = test_basis.coefficient_shape
        self.test_components = test_basis.elem._dim
        self.trial_components = trial_basis.elem._dim
        self.dimension = 3
        self._square = None
        self._cross = None
        self._rtabs = {}

    def _init_cut(self, test_basis, trial_basis, memory_limit_bytes, memory_safety_factor, facet=False):
        """Two cut-cell bases on the same points (bilinear_form.py:134-155),
        or two facet bases on the same facets, whose facets are cells of Q
        points each: both run the ragged-quadrature cross kernel."""
        from .cut import CutCrossAssembler
        if test_basis.dx.shape != trial_basis.dx.shape:
            raise ValueError("trial and test quadrature shapes must match")
        if not np.allclose(test_basis.dx, trial_basis.dx):
            raise ValueError("trial and test quadrature weights must match")
        if not np.allclose(test_basis.global_coordinates, trial_basis.global_coordinates):
            raise ValueError("trial and test quadrature coordinates must match")
        if facet:
            if not np.array_equal(test_basis.facet_ids, trial_basis.facet_ids):
                raise ValueError("trial and test facet bases must cover the same facets")
            rt, rc = test_basis.ragged(), trial_basis.ragged()
            rows = (rt["dofs"], rt["offsets"], rt["shape"], rt["gradients"], rt["weights"], rt["normals"])
            cols = (rc["dofs"], rc["shape"], rc["gradients"], rc["normals"])
        else:
            if not np.array_equal(test_basis.active_cell_offsets, trial_basis.active_cell_offsets):
                raise ValueError("cut cross bases have different cell offsets")
            rows = (test_basis.active_cell_dofs, test_basis.active_cell_offsets, test_basis.shape,
                    test_basis.gradients, test_basis.weights, test_basis.normal_vectors)
            cols = (trial_basis.active_cell_dofs, trial_basis.shape, trial_basis.gradients,
                    trial_basis.normal_vectors)
        self.test_basis, self.trial_basis = test_basis, trial_basis
        self.basis = test_basis
        self.memory_estimate = estimate_cut_memory(rows[0], cols[0], test_basis.N, trial_basis.N, "cross")
        enforce_memory_budget(self.memory_estimate, memory_limit_bytes, memory_safety_factor)
        self._cutx = CutCrossAssembler(rows[0], cols[0], rows[1], rows[2], cols[1], rows[4], rows[3], cols[2])
        self._cut_normals = (np.asarray(rows[5], dtype=np.float64).reshape(-1, 3),
                             np.asarray(cols[3], dtype=np.float64).reshape(-1, 3))
        self.shape = (test_basis.N, trial_basis.N)
        self.coefficient_shape = test_basis.coefficient_shape
        self.test_components = test_basis.elem._dim
        self.trial_components = trial_basis.elem._dim
        self.dimension = 3
        self._square = self._cross = None
        self._rtabs = {}

    def _cut_assemble(self, row_kind, column_kind, coefficient):
        """CutCrossAssembler.assemble with a (P,) scalar or a
        (P, Cr, [3], Cc, [3]) tensor; a fresh CSR copy per call."""
        c = self._cutx
        c.assemble(coefficient, row_kind, column_kind)
        m = csr_matrix((c.values.copy(), c.indices.copy(), c.indptr.copy()), shape=(c.rows, c.columns))
        m.resize(self.shape)
        return m

    def _cut_scalar(self, coefficient):
        arr = np.broadcast_to(np.asarray(coefficient, dtype=np.float64), self.coefficient_shape)
        return np.ascontiguousarray(arr).reshape(-1)

    def _cut_kinds(self, row_kind, column_kind, coefficient):
        """Value / gradient / normal-gradient contractions on ragged points
        (bilinear_form.py:243-325): the scalar times unit directions, or the
        basis normals for a normal-gradient side, as a contraction tensor."""
        s = self._cut_scalar(coefficient)
        rg, cg = row_kind != "value", column_kind != "value"
        rn = self._cut_normals[0] if row_kind == "normal_gradient" else None
        cn = self._cut_normals[1] if column_kind == "normal_gradient" else None
        if rg == cg and rn is None and cn is None:
            return self._cut_assemble(row_kind, column_kind, s)
        P, C, D = len(s), self.test_components, self.dimension
        T = np.zeros((P, C) + ((D,) if rg else ()) + (C,) + ((D,) if cg else ()))
        if rg and cg:
            # factor[i, j] = (n_i or delta_ij) * (m_j or 1)
            left = rn[:, :, None] * np.ones((1, 1, D)) if rn is not None else np.broadcast_to(np.eye(D), (P, D, D))
            factor = left * (cn[:, None, :] if cn is not None else 1.)
            for c in range(C):
                T[:, c, :, c, :] = s[:, None, None] * factor
        elif rg:
            direction = rn if rn is not None else np.ones((P, D))
            for c in range(C):
                T[:, c, :, c] = s[:, None] * direction
        else:
            direction = cn if cn is not None else np.ones((P, D))
            for c in range(C):
                T[:, c, c, :] = s[:, None] * direction
        return self._cut_assemble("gradient" if rg else "value", "gradient" if cg else "value", T)

    # ---------------------------------------------------------- plans
    @property
    def plan(self):
        """Square gather plan (gradient x gradient tensor of one scalar basis)."""
        if self._square is None:
            import torch
            plan = basis_plan(self.test_basis)
            self._square = (plan, torch.zeros(max(plan.nnz, 1), dtype=torch.float64, device=_native.device()))
        return self._square[0]

    @plan.setter
    def plan(self, plan):
        """Use another square plan of the same basis (e.g. another LDS budget)."""
        import torch
        self._square = (plan, torch.zeros(max(plan.nnz, 1), dtype=torch.float64, device=_native.device()))

    @property
    def _values(self):
        self.plan
        return self._square[1]

    def cross_plan(self):
        if self._cross is None:
            import torch
            plan = CrossPlan(self.test_basis, self.trial_basis)
            values = torch.zeros(max(plan.nnz, 1), dtype=torch.float64, device=_native.device())
            tables = upload(cross_tables(self.test_basis, self.trial_basis))
            self._cross = (plan, values, tables)
        return self._cross[0]

    # ---------------------------------------------------------- launches
    def assemble_cross_device(self, row_kind, column_kind, mode, coefficient=None, strides=(0, 0), out=None):
        """Launch the cross kernel with a device coefficient (scalar per
        (e,q) for modes scalar/row_divergence/column_divergence, tensor block
        for mode tensor) addressed as coefficient[e*strides[0] + q*strides[1]]."""
        plan = self.cross_plan()
        _, values, tables = self._cross
        vals = values if out is None else out
        b = self.test_basis
        d = b.device_data()
        iso = not b._affine
        rtab = None
        if not iso and mode != "tensor":
            key = (row_kind, column_kind)
            rtab = self._rtabs.get(key)
            if rtab is None:
                rtab = upload(cross_reduced_table(self.test_basis, self.trial_basis, row_kind, column_kind))
                self._rtabs[key] = rtab
        L = _native.lib()
        _native.check(L.skfa_assemble_cross(
            plan.handle, _native.ptr(d["xyz"]), _native.ptr(d["vconn"]), _native.ptr(d["gconn"]) if iso else None,
            d["gnodes"] if iso else 4, _native.ptr(tables), len(b.W), _native.ptr(rtab), _CROSS_KIND[row_kind],
            _CROSS_KIND[column_kind], _CROSS_MODE[mode], _native.ptr(coefficient), int(strides[0]),
            int(strides[1]), _native.ptr(vals), _native.stream()))
        return vals

    def _cross_matrix(self):
        plan, values, _ = self._cross
        ip, ix = plan.host_csr_arrays()
        data = to_host(values[:plan.nnz])
        m = csr_matrix((data, host_copy(ix), host_copy(ip)), shape=(plan.nrows, plan.ncols), copy=False)
        m.resize(self.shape)
        return m

    def _scalar_field(self, coefficient):
        """(device array, strides) of a scalar field broadcast to (E,Q)."""
        arr = np.asarray(coefficient, dtype=np.float64)
        E, Q = self.coefficient_shape
        if arr.ndim == 0:
            return upload(arr.reshape(1)), (0, 0)
        host, per_qp = field_layout(arr, (E, Q), "cross")
        return upload(host), ((Q, 1) if per_qp else (1, 0))

    def _tensor_field(self, coefficient, expected):
        """(device array, strides) of a tensor broadcast to expected =
        (E,Q)+block; element- or point-constant tensors upload one block per
        element or a single block."""
        full = np.broadcast_to(np.asarray(coefficient, dtype=np.float64), expected)
        block = int(np.prod(expected[2:]))
        if full.strides[0] == 0 and full.strides[1] == 0:
            return upload(np.ascontiguousarray(full[0, 0]).ravel()), (0, 0)
        if full.strides[1] == 0 or expected[1] == 1:
            return upload(np.ascontiguousarray(full[:, 0]).reshape(-1)), (block, 0)
        return upload(np.ascontiguousarray(full).reshape(-1)), (expected[1] * block, block)

    # ---------------------------------------------------------- reference API
    def assemble(self, kind, coefficient):
        """One term of a cross-basis form (bilinear_form.py:180-241): kind
        'value', 'gradient', 'row_divergence' or 'column_divergence' with a
        scalar coefficient broadcastable to (E,Q)."""
        if kind in ("row_divergence", "column_divergence"):
            if kind == "row_divergence" and (self.test_components != self.dimension or self.trial_components != 1):
                raise ValueError("row divergence requires vector test and scalar trial")
            if kind == "column_divergence" and (self.trial_components != self.dimension
                                                or self.test_components != 1):
                raise ValueError("column divergence requires scalar test and vector trial")
            row_kind, column_kind = (("gradient", "value") if kind == "row_divergence" else ("value", "gradient"))
            mode = kind
        elif kind in ("value", "gradient"):
            if self.test_components != self.trial_components:
                raise ValueError("value and gradient contractions require matching component counts")
            row_kind = column_kind = kind
            mode = "scalar"
        else:
            raise UnsupportedNativeForm(f"cross-basis assembly does not support the {kind!r} kind")
        if self._cutx is not None:
            scalar = self._cut_scalar(coefficient)
            if mode == "scalar":
                return self._cut_assemble(row_kind, column_kind, scalar)
            # the reference's divergence tensor: T[c,c,0] (row) or T[0,c,c] (column)
            C = self.dimension
            unit = np.zeros((C, C, 1) if mode == "row_divergence" else (1, C, C))
            for k in range(C):
                unit[(k, k, 0) if mode == "row_divergence" else (0, k, k)] = 1.
            return self._cut_assemble(row_kind, column_kind, scalar[:, None, None, None] * unit)
        coef, strides = self._scalar_field(coefficient)
        self.assemble_cross_device(row_kind, column_kind, mode, coef, strides)
        return self._cross_matrix()

    def assemble_kinds(self, row_kind, column_kind, coefficient=1., *, num_threads=0):
        """Value/gradient contractions with a scalar coefficient
        (bilinear_form.py:243-325, volume bases: no normal gradients)."""
        valid = {"value", "gradient", "normal_gradient"}
        if row_kind not in valid or column_kind not in valid:
            raise ValueError("cross kind must be value, gradient, or normal_gradient")
        if self.test_components != self.trial_components:
            raise ValueError("cross contractions require matching components")
        if self._cutx is not None:
            return self._cut_kinds(row_kind, column_kind, coefficient)
        if row_kind == "normal_gradient":
            raise ValueError("row normal_gradient requires test-basis normals")
        if column_kind == "normal_gradient":
            raise ValueError("column normal_gradient requires trial-basis normals")
        if row_kind == column_kind:
            if self._cutx is not None:
                return self._cut_assemble(row_kind, column_kind, self._cut_scalar(coefficient))
            coef, strides = self._scalar_field(coefficient)
            self.assemble_cross_device(row_kind, column_kind, "scalar", coef, strides)
            return self._cross_matrix()
        # value x gradient with a scalar: the reference's tensor has
        # T[c,i,c] = s (or T[c,c,j] = s) for every direction
        C = self.test_components
        shape = (C, 3, C) if row_kind == "gradient" else (C, C, 3)
        unit = np.zeros(shape)
        for c in range(C):
            if row_kind == "gradient":
                unit[c, :, c] = 1.
            else:
                unit[c, c, :] = 1.
        scalar = np.broadcast_to(np.asarray(coefficient, dtype=np.float64), self.coefficient_shape)
        return self.assemble_tensor(row_kind, column_kind, scalar[(...,) + (None,) * 3] * unit)

    def assemble_tensor(self, row_kind, column_kind, coefficient, *, num_threads=0, precision="fp64"):
        """Caller-supplied contraction tensor (bilinear_form.py:327-362),
        coefficient broadcastable to (E,Q,Cr,[3],Cc,[3])."""
        valid = {"value", "gradient"}
        if row_kind not in valid or column_kind not in valid:
            raise ValueError("tensor kind must be value or gradient")
        row_axes = (self.test_components,) + ((self.dimension,) if row_kind == "gradient" else ())
        column_axes = (self.trial_components,) + ((self.dimension,) if column_kind == "gradient" else ())
        expected = tuple(self.coefficient_shape) + row_axes + column_axes
        coefficient = np.asarray(coefficient, dtype=np.float64)
        try:
            coefficient = np.broadcast_to(coefficient, expected)
        except ValueError as error:
            raise ValueError(f"tensor coefficient must broadcast to {expected}") from error
        if self._cutx is not None:
            return self._cut_assemble(row_kind, column_kind,
                                      np.ascontiguousarray(coefficient).reshape((-1,) + expected[2:]))
        square = (self.test_basis is self.trial_basis and (row_kind, column_kind) == ("gradient", "gradient")
                  and self.test_components == 1)
        if square:
            arr, per_qp = field_layout(coefficient[:, :, 0, :, 0, :], self.coefficient_shape, "tensor", (3, 3))
            dev = upload(arr)
            # A == A^T per element: the three off-diagonal pairs agree (checked on the device)
            import torch
            sym = all(bool(torch.equal(dev[..., i, j], dev[..., j, i])) for i, j in ((0, 1), (0, 2), (1, 2)))
            self.assemble_tensor_device(dev, per_qp=per_qp, precision=precision, symmetric=sym)
            ip, ix = self.plan.host_csr_arrays()
            data = to_host(self._values[:self.plan.nnz])
            m = csr_matrix((data, host_copy(ix), host_copy(ip)), shape=(self.plan.nrows, self.plan.nrows),
                           copy=False)
            m.resize(self.shape)
            return m
        if precision != "fp64":
            raise UnsupportedNativeForm("FP32 products are available for the square gradient tensor path")
        coef, strides = self._tensor_field(coefficient, expected)
        self.assemble_cross_device(row_kind, column_kind, "tensor", coef, strides)
        return self._cross_matrix()

    def assemble_tensor_device(self, tensor, *, per_qp=False, precision="fp64", out=None, symmetric=False,
                               load=None):
        """Tensor (E,3,3) [(E,Q,3,3) when per_qp] already on the device;
        ``symmetric`` declares A == A^T per element (smaller payload);
        ``load`` = (f, out) also writes the load vector of the constant
        source f (a one-element device tensor) into out (nnodes FP64), fused
        with the gather where it can be."""
        vals = self._values if out is None else out
        assemble_bilinear_device(self.test_basis, self.plan, vals, tensor=tensor, per_qp=per_qp,
                                 precision=precision, symmetric_tensor=symmetric, load=load)
        return vals


def _composite_csr(ip, ix, data, row_map, col_map, N):
    """Field-local CSR -> composite numbering.  Both maps are strictly
    increasing (composite DOFs follow node order within a field), so rows
    keep their order and columns stay sorted: only indptr is expanded and the
    column ids relabelled."""
    counts = np.zeros(N, dtype=np.int64)
    counts[row_map] = np.diff(np.asarray(ip, dtype=np.int64))
    indptr = np.zeros(N + 1, dtype=np.int64)
    np.cumsum(counts, out=indptr[1:])
    indices = col_map[np.asarray(ix, dtype=np.int64)]
    if indptr[-1] < 2 ** 31 - 1:
        indptr, indices = indptr.astype(np.int32), indices.astype(np.int32)
    return csr_matrix((data, indices, indptr), shape=(N, N), copy=False)


class NativeCompositeBilinearForm:
    """Blocks between the subfields of a composite basis
    (bilinear_form.py:370-455); each call returns a fresh (N, N) CSR in the
    composite numbering.  Value and gradient couplings of a field with itself
    run on the square row-gather kernels, every other block on the cross
    kernels."""

    def __init__(self, basis):
        self.basis = basis
        self._assemblers = {}
        self._patterns = {}

    def _square(self, field):
        key = ("square", field)
        native = self._assemblers.get(key)
        if native is None:
            native = NativeBilinearForm(self.basis.subbases[field])
            self._assemblers[key] = native
        return native

    def _cross(self, row_field, column_field):
        key = ("cross", row_field, column_field)
        native = self._assemblers.get(key)
        if native is None:
            native = NativeCrossBilinearForm(self.basis.subbases[row_field], self.basis.subbases[column_field])
            self._assemblers[key] = native
        return native

    def assemble(self, row_field, column_field, *, kind, coefficient):
        row = self.basis.subbases[row_field]
        column = self.basis.subbases[column_field]
        rc, cc = row.elem._dim, column.elem._dim
        divergence = kind in ("row_divergence", "column_divergence")
        if kind not in ("value", "gradient") and not divergence:
            raise UnsupportedNativeForm(f"composite blocks do not support the {kind!r} kind")
        if rc != cc and not divergence:
            raise ValueError("component-mismatched composite coupling requires an explicit tensor contraction")
        if kind == "row_divergence" and (rc != 3 or cc != 1):
            raise ValueError("row divergence requires vector test and scalar trial")
        if kind == "column_divergence" and (cc != 3 or rc != 1):
            raise ValueError("column divergence requires scalar test and vector trial")
        if row_field == column_field and not divergence:
            native = self._square(row_field)
            dev, per_qp = _prepare_scalar_kinds(row, {kind: coefficient})
            native.assemble_device(per_qp=per_qp, **dev)
            plan, values = native.plan, native._values
        else:
            native = self._cross(row_field, column_field)
            if divergence:
                row_kind, column_kind = ("gradient", "value") if kind == "row_divergence" else ("value", "gradient")
                mode = kind
            else:
                row_kind = column_kind = kind
                mode = "scalar"
            coef, strides = native._scalar_field(coefficient)
            native.assemble_cross_device(row_kind, column_kind, mode, coef, strides)
            plan, values = native.cross_plan(), native._cross[1]
        key = (row_field, column_field, row_field == column_field and not divergence)
        pattern = self._patterns.get(key)
        if pattern is None:
            ip, ix = plan.host_csr_arrays()
            m = _composite_csr(ip, ix, np.zeros(len(ix)), row._composite_dofs, column._composite_dofs,
                               self.basis.N)
            pattern = (m.indptr, m.indices)
            self._patterns[key] = pattern
        data = to_host(values[:plan.nnz])
        # the relabelled pattern is fixed per block; each result owns a copy
        return csr_matrix((data, host_copy(pattern[1]), host_copy(pattern[0])), shape=(self.basis.N, self.basis.N),
                          copy=False)


@dataclass(frozen=True)
class LinearFormDiagnostics:
    entity_count: int
    quadrature_point_count: int
    assembly_seconds: float


class NativeLinearForm:
    """f.v + G:grad(v) assembly (linear_form.py:17-107)."""

    def __init__(self, basis):
        self.basis = basis
        self._cut = None
        if hasattr(basis, "cell_offsets"):
            # cut-cell basis (linear_form.py:33-41)
            from .cut import CutLinearFormAssembler
            self._cut = CutLinearFormAssembler(basis.active_cell_dofs, basis.active_cell_offsets, basis.shape,
                                               basis.gradients, basis.weights)
            P = len(basis.weights)
            C = basis.elem._dim
            self._shape = (P, 1, C)
            self._gradient_shape = (P, 1, C, 3)
            return
        self.plan = basis_plan(basis)
        E, Q = basis.coefficient_shape
        C = basis.elem._dim
        self._shape = (E, Q, C)
        self._gradient_shape = (E, Q, C, 3)
        import torch
        self._out = torch.zeros(max(self.plan.nrows, 1), dtype=torch.float64, device=_native.device())

    @property
    def ndofs(self):
        return self._cut.ndofs if self._cut is not None else self.plan.nrows

    @staticmethod
    def _coefficient(name, coefficient, shape):
        if coefficient is None:
            return None, False
        coefficient = np.asarray(coefficient)
        if coefficient.dtype != np.float64:
            raise TypeError(f"{name} coefficient must have dtype float64")
        if coefficient.shape == shape[2:]:
            return np.ascontiguousarray(coefficient), False
        try:
            coefficient = np.broadcast_to(coefficient, shape)
        except ValueError as error:
            raise ValueError(f"{name} coefficient must broadcast to {shape}") from error
        return np.ascontiguousarray(coefficient), True

    def assemble(self, *, value=None, gradient=None, num_threads=0):
        f, fq = self._coefficient("value", value, self._shape)
        G, Gq = self._coefficient("gradient", gradient, self._gradient_shape)
        if f is None and G is None:
            raise ValueError("at least one coefficient is required")
        if self._cut is not None:
            if f is not None and f.ndim == 3:
                f = f[:, 0, :]
            if G is not None and G.ndim == 4:
                G = G[:, 0, :, :]
            result, seconds = self._cut.assemble(f, G)
            return result, LinearFormDiagnostics(entity_count=self._cut.cell_count,
                                                 quadrature_point_count=self._cut.point_count,
                                                 assembly_seconds=seconds)
        start = time.perf_counter()
        out = self.assemble_device(upload(f) if f is not None else None, fq,
                                   upload(G) if G is not None else None, Gq)
        result = to_host(out[:self.plan.nrows])
        seconds = time.perf_counter() - start
        E, Q, _ = self._shape
        return result, LinearFormDiagnostics(entity_count=E, quadrature_point_count=Q, assembly_seconds=seconds)

    def assemble_device(self, f, f_per_qp, G, G_per_qp):
        basis = self.basis
        d = basis.device_data()
        iso = not basis._affine
        L = _native.lib()
        _native.check(L.skfa_assemble_linear(
            self.plan.handle, _native.ptr(d["xyz"]), _native.ptr(d["vconn"]),
            _native.ptr(d["gconn"]) if iso else None, d["gnodes"] if iso else 4,
            _native.ptr(d["tables_iso"] if iso else d["tables"]), len(basis.W),
            _native.ptr(f), int(f_per_qp), _native.ptr(G), int(G_per_qp), _native.ptr(self._out), _native.stream()))
        return self._out


# ------------------------------------------------------------------ memory budget
class AssemblyMemoryBudgetError(MemoryError):
    """Raised before allocation when an assembler would exceed the budget."""


@dataclass(frozen=True)
class AssemblyMemoryEstimate:
    kind: str
    rows: int
    columns: int
    nnz: int
    device_bytes: int


def estimate_memory_upper(basis):
    """Upper bound of the device bytes one assembler holds, known before any
    allocation: CSR pattern + values (nnz <= min(N^2, E L^2)), gather records
    (<= E n padded 16-byte records per lane) and incidence lists."""
    E = basis.mesh.nelements
    n = basis._node_local.shape[0]
    C = basis.elem._dim
    N = basis.N
    L = n * C
    nnz = min(N * N, E * L * L)
    b = 8 * (N + 1) + 12 * nnz + 2 * 32 * E * n + 4 * E * n + 8 * (N + 1)
    return AssemblyMemoryEstimate("bilinear", N, N, nnz, int(b))


def estimate_cross_memory(test_basis, trial_basis):
    """Upper bound of the device bytes of one cross assembler: rectangular
    pattern + values (nnz <= min(Nr Nc, E Lr Lc)), one record per row
    incidence and the incidence lists."""
    E = test_basis.mesh.nelements
    nr, nc = test_basis._node_local.shape[0], trial_basis._node_local.shape[0]
    Lr, Lc = nr * test_basis.elem._dim, nc * trial_basis.elem._dim
    Nr, Nc = test_basis.N, trial_basis.N
    nnz = min(Nr * Nc, E * Lr * Lc)
    b = 8 * (Nr + 1) + 12 * nnz + 32 * E * nr + 4 * E * nr + 8 * (Nr + 1)
    return AssemblyMemoryEstimate("cross", Nr, Nc, nnz, int(b))


def integrate_functional(values, weights):
    """sum(values * weights) over matching (entities, quadrature) arrays,
    reduced on the GPU (integrate_functional, functional_assembler.cpp:12-33)."""
    import torch
    v = np.ascontiguousarray(values, dtype=np.float64)
    w = np.ascontiguousarray(weights, dtype=np.float64)
    if v.ndim != 2 or w.ndim != 2 or v.shape != w.shape:
        raise ValueError("functional values and weights must have matching (entities, quadrature) shapes")
    if v.size == 0:
        return 0.0
    dv, dw = upload(v), upload(w)
    ptr_ = upload(np.array([0, v.size], dtype=np.int64))
    out = torch.zeros(1, dtype=torch.float64, device=_native.device())
    _native.check(_native.lib().skfa_sys_dot(_native.ptr(dv), _native.ptr(dw), _native.ptr(ptr_), 1,
                                             _native.ptr(out), None, _native.stream()))
    return float(out.item())


def estimate_cut_memory(row_dofs, col_dofs, N_rows, N_cols, kind):
    """Upper bound of the device bytes of one cut assembler: pattern +
    values (nnz <= cells Lr Lc), one record per row incidence and the
    incidence lists."""
    cells, nr, Cr = row_dofs.shape
    _, nc, Cc = col_dofs.shape
    nnz = min(N_rows * N_cols, cells * nr * Cr * nc * Cc)
    b = 8 * (N_rows + 1) + 12 * nnz + 36 * cells * nr + 8 * (N_rows + 1)
    return AssemblyMemoryEstimate(kind, N_rows, N_cols, nnz, int(b))


def enforce_memory_budget(estimate, limit, safety_factor=1.25):
    if limit is None:
        return
    need = int(np.ceil(estimate.device_bytes * float(safety_factor)))
    if need > int(limit):
        raise AssemblyMemoryBudgetError(
            f"assembly needs about {need} bytes (safety factor {safety_factor}) but the limit is {int(limit)} bytes")


class NativeCompositeLinearForm:
    """Linear assemblers per composite subfield (linear_form.py:110-132);
    ``assemble(field, value=, gradient=)`` returns the field's load vector in
    the composite numbering (length N, zeros at the other fields' DOFs)."""

    def __init__(self, basis):
        self.basis = basis
        self._assemblers = {}

    def assembler(self, field):
        native = self._assemblers.get(field)
        if native is None:
            native = NativeLinearForm(self.basis.subbases[field])
            self._assemblers[field] = native
        return native

    def assemble(self, field, *, value=None, gradient=None):
        vector, _ = self.assembler(field).assemble(value=value, gradient=gradient)
        result = np.zeros(self.basis.N, dtype=np.float64)
        sub = self.basis.subbases[field]
        result[sub._composite_dofs] = vector[:sub._field_N]
        return result
