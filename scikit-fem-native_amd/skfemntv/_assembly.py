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
              "row_split": 6, "mfma": 7}


class Plan:
    """Device CSR pattern + gather plan of one (basis, components) pair."""

    def __init__(self, node_conn_dev, E, n, nnodes, components, lds_budget=12 * 1024):
        import torch
        L = _native.lib()
        h = ctypes.c_void_p()
        _native.check(L.skfa_plan_create_budget(_native.ptr(node_conn_dev), E, n, nnodes, components,
                                                int(lds_budget), _native.stream(), ctypes.byref(h)))
        self._h = h
        self.lds_budget = int(lds_budget)
        self.tuning = {"xcd_chunk": 16, "lds_atomic": 1, "placement": 2, "streaming": 1, "element_table": -1,
                       "row_split": 0, "mfma": -1}
        info = (ctypes.c_int64 * 9)()
        _native.check(L.skfa_plan_info(h, info))
        self.nrows, self.nnz, self.nnz_node = int(info[0]), int(info[1]), int(info[2])
        self.max_row_len, self.max_inc = int(info[3]), int(info[4])
        self.waves, self.steps, self.rec_bytes, self.buckets = int(info[5]), int(info[6]), int(info[7]), int(info[8])
        dev = _native.device()
        self.indptr = torch.empty(self.nrows + 1, dtype=torch.int64, device=dev)
        self.indices = torch.empty(max(self.nnz, 1), dtype=torch.int32, device=dev)
        _native.check(L.skfa_plan_export_csr(h, _native.ptr(self.indptr), _native.ptr(self.indices), _native.stream()))
        self._host = None
        self.E, self.n, self.nnodes, self.components = E, n, nnodes, components

    @property
    def handle(self):
        return self._h

    def set_tuning(self, **settings):
        """Launch tuning (library key in parentheses).

        Bit-identical results for every value: ``xcd_chunk`` (1, waves per
        XCD chunk), ``lds_atomic`` (2), ``placement`` (3: 0 plain, 1 XCD
        segments, 2 XCD chunks), ``streaming`` (4), ``row_split`` (6: how
        the lanes of one row share it: 0 incidences, 1 entries, 2 columns).  Different rounding only:
        ``element_table`` (5: -1 auto, 0, 1) and ``mfma`` (7: 0 off, 1 all
        waves, 2 waves of at most 16 rows)."""
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
        # vector fields carry 9-entry node blocks: a larger span per wave keeps
        # more rows per wave (measured on 48^3 vector P2: 2.50 ms at 16 KiB,
        # 2.72 ms at 32 KiB, 2.97 ms at 12 KiB; scalar P2 is fastest at 12 KiB)
        budget = 16 * 1024 if C == 3 else 12 * 1024
        plan = Plan(d["node_conn"], basis.mesh.nelements, d["node_conn"].shape[1], d["nnodes"], C,
                    lds_budget=budget)
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
    import torch
    arr = np.ascontiguousarray(host)
    if not arr.flags.writeable:
        arr = arr.copy()
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


def to_per_qp(dev_tensor, Q):
    """(E,)+t -> (E,Q)+t on the device."""
    shape = dev_tensor.shape
    return dev_tensor.unsqueeze(1).expand((shape[0], Q) + tuple(shape[1:])).contiguous()


def assemble_bilinear_device(basis, plan, values_dev, *, value=None, gradient=None, symmetric_gradient=None,
                             divergence=None, tensor=None, per_qp=False, precision="fp64", symmetric_tensor=False):
    """Launch skfa_assemble_bilinear with device coefficient tensors."""
    d = basis.device_data()
    L = _native.lib()
    iso = not basis._affine
    tables = d["tables_iso"] if iso else d["tables"]
    prec = {"fp64": 0, "fp32": 1}[precision]
    if prec == 1 and (iso or per_qp):
        raise UnsupportedNativeForm("FP32 products are available for affine elements with element-constant "
                                    "coefficients")
    _native.check(L.skfa_assemble_bilinear(
        plan.handle, _native.ptr(d["xyz"]), d["xyz"].shape[0], _native.ptr(d["vconn"]),
        _native.ptr(d["gconn"]) if iso else None, d["gnodes"] if iso else 4,
        _native.ptr(tables), len(basis.W), _native.ptr(d["rtab"]), _native.ptr(d["rtab32"]), d["nqg"], d["nqm"],
        _native.ptr(value), _native.ptr(gradient), _native.ptr(symmetric_gradient), _native.ptr(divergence),
        _native.ptr(tensor), int(per_qp), prec, _native.FLAG_SYMMETRIC_TENSOR if symmetric_tensor else 0,
        _native.ptr(values_dev), _native.stream()))


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

    def assemble(self, *, value=None, gradient=None, symmetric_gradient=None, divergence=None, num_threads=0,
                 precision="fp64"):
        coefs = {"value": value, "gradient": gradient, "symmetric_gradient": symmetric_gradient,
                 "divergence": divergence}
        if all(v is None for v in coefs.values()):
            raise ValueError("at least one coefficient is required")
        dev, per_qp = _prepare_scalar_kinds(self.basis, coefs)
        self.assemble_device(per_qp=per_qp, precision=precision, **dev)
        if self._pinned is not None:
            to_host(self._values[:self.plan.nnz], self._pinned[:self.plan.nnz])
        else:
            np.copyto(self._matrix.data, to_host(self._values[:self.plan.nnz]))
        return self._matrix

    def assemble_device(self, *, value=None, gradient=None, symmetric_gradient=None, divergence=None,
                        tensor=None, per_qp=False, precision="fp64", out=None):
        """Assemble into a device tensor (default: this form's buffer) with
        coefficient tensors already on the device; returns the values."""
        vals = self._values if out is None else out
        assemble_bilinear_device(self.basis, self.plan, vals, value=value, gradient=gradient,
                                 symmetric_gradient=symmetric_gradient, divergence=divergence, tensor=tensor,
                                 per_qp=per_qp, precision=precision)
        return vals


class NativeCrossBilinearForm:
    """Gradient x gradient tensor assembly on one scalar basis
    (bilinear_form.py:102-178,336-367); returns a fresh CSR copy."""

    def __init__(self, test_basis, trial_basis, *, memory_limit_bytes=None, memory_safety_factor=1.25):
        if test_basis is not trial_basis:
            raise UnsupportedNativeForm("cross-basis assembly between different bases is outside the MI355X build")
        self.basis = test_basis
        self.memory_estimate = estimate_memory_upper(test_basis)
        enforce_memory_budget(self.memory_estimate, memory_limit_bytes, memory_safety_factor)
        self.plan = basis_plan(test_basis)
        import torch
        self._values = torch.zeros(max(self.plan.nnz, 1), dtype=torch.float64, device=_native.device())
        self.shape = (test_basis.N, trial_basis.N)
        self.coefficient_shape = test_basis.coefficient_shape
        self.test_components = test_basis.elem._dim
        self.trial_components = trial_basis.elem._dim
        self.dimension = 3

    def assemble_tensor(self, row_kind, column_kind, coefficient, *, num_threads=0, precision="fp64"):
        valid = {"value", "gradient"}
        if row_kind not in valid or column_kind not in valid:
            raise ValueError("tensor kind must be value or gradient")
        if (row_kind, column_kind) != ("gradient", "gradient") or self.test_components != 1:
            raise UnsupportedNativeForm("the MI355X build assembles gradient x gradient tensors of scalar fields")
        expected = self.coefficient_shape + (1, 3, 1, 3)
        coefficient = np.asarray(coefficient, dtype=np.float64)
        try:
            coefficient = np.broadcast_to(coefficient, expected)
        except ValueError as error:
            raise ValueError(f"tensor coefficient must broadcast to {expected}") from error
        arr, per_qp = field_layout(coefficient[:, :, 0, :, 0, :], self.coefficient_shape, "tensor", (3, 3))
        # A == A^T per element: the three off-diagonal pairs agree
        sym = all(np.array_equal(arr[..., i, j], arr[..., j, i]) for i, j in ((0, 1), (0, 2), (1, 2)))
        self.assemble_tensor_device(upload(arr), per_qp=per_qp, precision=precision, symmetric=sym)
        ip, ix = self.plan.host_csr_arrays()
        data = to_host(self._values[:self.plan.nnz])
        m = csr_matrix((data, ix.copy(), ip.copy()), shape=(self.plan.nrows, self.plan.nrows), copy=False)
        m.resize(self.shape)
        return m

    def assemble_tensor_device(self, tensor, *, per_qp=False, precision="fp64", out=None, symmetric=False):
        """Tensor (E,3,3) [(E,Q,3,3) when per_qp] already on the device;
        ``symmetric`` declares A == A^T per element (smaller payload)."""
        vals = self._values if out is None else out
        assemble_bilinear_device(self.basis, self.plan, vals, tensor=tensor, per_qp=per_qp, precision=precision,
                                 symmetric_tensor=symmetric)
        return vals


@dataclass(frozen=True)
class LinearFormDiagnostics:
    entity_count: int
    quadrature_point_count: int
    assembly_seconds: float


class NativeLinearForm:
    """f.v + G:grad(v) assembly (linear_form.py:17-107)."""

    def __init__(self, basis):
        self.basis = basis
        self.plan = basis_plan(basis)
        E, Q = basis.coefficient_shape
        C = basis.elem._dim
        self._shape = (E, Q, C)
        self._gradient_shape = (E, Q, C, 3)
        import torch
        self._out = torch.zeros(max(self.plan.nrows, 1), dtype=torch.float64, device=_native.device())

    @property
    def ndofs(self):
        return self.plan.nrows

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


def enforce_memory_budget(estimate, limit, safety_factor=1.25):
    if limit is None:
        return
    need = int(np.ceil(estimate.device_bytes * float(safety_factor)))
    if need > int(limit):
        raise AssemblyMemoryBudgetError(
            f"assembly needs about {need} bytes (safety factor {safety_factor}) but the limit is {int(limit)} bytes")
