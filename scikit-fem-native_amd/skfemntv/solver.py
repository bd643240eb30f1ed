"""Solving assembled systems on the GPU: device CSR, Dirichlet condensation,
batched FP64 preconditioned conjugate gradients (Jacobi or a factorised
approximate inverse M^{-1} = G^T G).

The reference leaves solving to the caller (python/skfemntv/capabilities.py:67);
its one solve pattern is the condensed direct solve ``solve_dirichlet`` of
examples/scikit-fem-gallery/comparison.py:80-87 (free = setdiff1d(arange(n),
boundary); solution[free] = spsolve(matrix[free][:, free], rhs[free])).
``solve_dirichlet`` below keeps that signature and result convention with an
iterative GPU solve.  Several independent systems are solved together as one
block-diagonal CSR; ``sys_ptr`` (nsys+1) gives each system's row range.
"""

from __future__ import annotations

import ctypes
import time
from dataclasses import dataclass, field

import numpy as np

from . import _native


def _torch():
    import torch
    return torch


class DeviceCSR:
    """CSR matrix in device memory: indptr int64 (n+1), indices int32, data float64."""

    def __init__(self, indptr, indices, data, shape):
        self.indptr, self.indices, self.data = indptr, indices, data
        self.shape = (int(shape[0]), int(shape[1]))

    @property
    def nnz(self):
        return int(self.data.shape[0])

    @classmethod
    def from_scipy(cls, m):
        torch = _torch()
        m = m.tocsr()
        m.sort_indices()
        dev = _native.device()
        return cls(torch.from_numpy(m.indptr.astype(np.int64)).to(dev),
                   torch.from_numpy(m.indices.astype(np.int32)).to(dev),
                   torch.from_numpy(np.ascontiguousarray(m.data, dtype=np.float64)).to(dev), m.shape)

    def to_scipy(self):
        from scipy.sparse import csr_matrix
        return csr_matrix((self.data.cpu().numpy(), self.indices.cpu().numpy(), self.indptr.cpu().numpy()),
                          shape=self.shape)

    def matvec(self, x, out=None, lanes_per_row=8):
        torch = _torch()
        y = torch.empty(self.shape[0], dtype=torch.float64, device=x.device) if out is None else out
        _native.check(_native.lib().skfa_spmv(_native.ptr(self.indptr), _native.ptr(self.indices),
                                              _native.ptr(self.data), _native.ptr(x), _native.ptr(y),
                                              self.shape[0], lanes_per_row, _native.stream()))
        return y

    def diagonal_inverse(self):
        torch = _torch()
        d = torch.empty(self.shape[0], dtype=torch.float64, device=self.data.device)
        _native.check(_native.lib().skfa_diag_inverse(_native.ptr(self.indptr), _native.ptr(self.indices),
                                                      _native.ptr(self.data), self.shape[0], _native.ptr(d),
                                                      _native.stream()))
        return d

    def condense(self, keep):
        """Rows and columns with keep != 0 (device uint8/bool, length n), in
        order.  Returns (DeviceCSR, newid) with newid[i] the new index of
        kept row i (n+1 entries, exclusive scan of keep)."""
        torch = _torch()
        L = _native.lib()
        n = self.shape[0]
        keep = keep.to(torch.uint8).contiguous()
        dev = self.data.device
        newid = torch.empty(n + 1, dtype=torch.int64, device=dev)
        ip = torch.empty(n + 1, dtype=torch.int64, device=dev)
        ix = torch.empty(max(self.nnz, 1), dtype=torch.int32, device=dev)
        vals = torch.empty(max(self.nnz, 1), dtype=torch.float64, device=dev)
        nr, nz = ctypes.c_int64(), ctypes.c_int64()
        _native.check(L.skfa_csr_condense(_native.ptr(self.indptr), _native.ptr(self.indices), _native.ptr(self.data),
                                          n, _native.ptr(keep), _native.ptr(newid), _native.ptr(ip), _native.ptr(ix),
                                          _native.ptr(vals), ctypes.byref(nr), ctypes.byref(nz), _native.stream()))
        m, z = nr.value, nz.value
        out = DeviceCSR(ip[:m + 1].clone(), ix[:z].clone(), vals[:z].clone(), (m, m))
        return out, newid

    def block(self, r0, r1):
        """Rows/columns [r0, r1) of a block-diagonal matrix as its own CSR."""
        ip = self.indptr[r0:r1 + 1]
        k0, k1 = int(ip[0]), int(ip[-1])
        return DeviceCSR((ip - k0).contiguous(), (self.indices[k0:k1] - r0).contiguous(),
                         self.data[k0:k1].contiguous(), (r1 - r0, r1 - r0))

    def lower(self):
        """Lower-triangular pattern (j <= i): (lptr, lidx, lsrc)."""
        torch = _torch()
        L = _native.lib()
        n = self.shape[0]
        dev = self.data.device
        lptr = torch.empty(n + 1, dtype=torch.int64, device=dev)
        nl = ctypes.c_int64()
        # the lower part has at most nnz entries: one pass with buffers of that size
        cap = max(self.nnz, 1)
        lidx = torch.empty(cap, dtype=torch.int32, device=dev)
        lsrc = torch.empty(cap, dtype=torch.int64, device=dev)
        _native.check(L.skfa_csr_lower(_native.ptr(self.indptr), _native.ptr(self.indices), n, _native.ptr(lptr),
                                       _native.ptr(lidx), _native.ptr(lsrc), ctypes.byref(nl), _native.stream()))
        return LowerPattern(lptr, lidx[:nl.value], lsrc[:nl.value], n, source=self)


@dataclass
class LowerPattern:
    """Lower-triangular pattern of a CSR with the source positions of its entries."""
    lptr: object
    lidx: object
    lsrc: object
    n: int
    source: object = field(default=None, repr=False)
    _transpose: tuple = field(default=None, repr=False)

    @property
    def nnz(self):
        return int(self.lidx.shape[0])

    def transpose(self):
        """(tptr, tidx, tsrc): CSR of the transpose and, per entry, its position in the lower pattern."""
        if self._transpose is None:
            torch = _torch()
            L = _native.lib()
            dev = self.lptr.device
            tptr = torch.empty(self.n + 1, dtype=torch.int64, device=dev)
            tidx = torch.empty(max(self.nnz, 1), dtype=torch.int32, device=dev)
            tsrc = torch.empty(max(self.nnz, 1), dtype=torch.int64, device=dev)
            P = _native.ptr
            done = False
            if self.source is not None:
                # structurally symmetric source: upper part of its pattern, no sort
                A = self.source
                rc = L.skfa_csr_transpose_lower(P(A.indptr), P(A.indices), self.n, P(self.lptr), P(self.lidx),
                                                P(tptr), P(tidx), P(tsrc), _native.stream())
                if rc == 0:
                    done = True
                elif rc != -1:
                    _native.check(rc)
            if not done:
                _native.check(L.skfa_csr_transpose(P(self.lptr), P(self.lidx), self.n, self.n, P(tptr), P(tidx),
                                                   P(tsrc), _native.stream()))
            self._transpose = (tptr, tidx[:self.nnz], tsrc[:self.nnz])
        return self._transpose


class JacobiPreconditioner:
    """M^{-1} = diag(A)^{-1}."""
    kind = 0

    def __init__(self, A):
        self.dinv = A.diagonal_inverse()

    def arrays(self):
        return (self.dinv, None, None, None, None, None, None)


class FactorPreconditioner:
    """M^{-1} = G^T G with G on the lower pattern of A (FP32 values)."""
    kind = 1

    def __init__(self, pattern, g_values):
        torch = _torch()
        self.pattern = pattern
        self.g = g_values.to(torch.float32).contiguous()
        tptr, tidx, tsrc = pattern.transpose()
        self.gt = torch.empty_like(self.g)
        _native.check(_native.lib().skfa_gather_values_f32(_native.ptr(self.g), _native.ptr(tsrc), pattern.nnz,
                                                           _native.ptr(self.gt), _native.stream()))
        self.tptr, self.tidx = tptr, tidx

    def arrays(self):
        p = self.pattern
        return (None, p.lptr, p.lidx, self.g, self.tptr, self.tidx, self.gt)

    def apply(self, r):
        """z = G^T (G r) (for tests)."""
        torch = _torch()
        p = self.pattern
        g64 = DeviceCSR(p.lptr, p.lidx, self.g.to(torch.float64), (p.n, p.n))
        gt64 = DeviceCSR(self.tptr, self.tidx, self.gt.to(torch.float64), (p.n, p.n))
        return gt64.matvec(g64.matvec(r))


@dataclass
class PCGResult:
    x: object                 # device solution (n,)
    iterations: np.ndarray    # per system
    batch_iterations: int
    seconds: float            # wall time of the solve (device synchronised)


def pcg(A, b, sys_ptr, preconditioner, *, tol=1e-8, max_iter=20000, check_every=8, spmv_lanes=8):
    """Batched FP64 PCG from x0 = 0 until ||r_s|| <= tol ||b_s|| per system."""
    torch = _torch()
    L = _native.lib()
    n = A.shape[0]
    dev = A.data.device
    if not isinstance(sys_ptr, torch.Tensor):
        sys_ptr = torch.as_tensor(np.asarray(sys_ptr, dtype=np.int64), device=dev)
    nsys = int(sys_ptr.shape[0]) - 1
    x = torch.empty(n, dtype=torch.float64, device=dev)
    work = torch.empty(7 * n + 1, dtype=torch.float64, device=dev)
    sys_work = torch.empty(6 * nsys, dtype=torch.float64, device=dev)
    int_work = torch.empty(nsys + 1, dtype=torch.int32, device=dev)
    iters = torch.empty(nsys, dtype=torch.int32, device=dev)
    total = ctypes.c_int32()
    dinv, gptr, gidx, gval, gtptr, gtidx, gtval = preconditioner.arrays()
    P = _native.ptr
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    _native.check(L.skfa_pcg_solve(P(A.indptr), P(A.indices), P(A.data), n, P(sys_ptr), nsys, P(b), P(x),
                                   preconditioner.kind, P(dinv), P(gptr), P(gidx), P(gval), P(gtptr), P(gtidx),
                                   P(gtval), float(tol), int(max_iter), int(check_every), int(spmv_lanes), P(work),
                                   P(sys_work), P(int_work), P(iters), ctypes.byref(total), _native.stream()))
    torch.cuda.synchronize()
    seconds = time.perf_counter() - t0
    return PCGResult(x, iters.cpu().numpy(), int(total.value), seconds)


def relative_residuals(A, b, x, sys_ptr):
    """Per-system true residual ||b - A x|| / ||b|| (FP64, on the device)."""
    torch = _torch()
    L = _native.lib()
    dev = A.data.device
    if not isinstance(sys_ptr, torch.Tensor):
        sys_ptr = torch.as_tensor(np.asarray(sys_ptr, dtype=np.int64), device=dev)
    nsys = int(sys_ptr.shape[0]) - 1
    r = b - A.matvec(x)
    rr = torch.empty(nsys, dtype=torch.float64, device=dev)
    bb = torch.empty(nsys, dtype=torch.float64, device=dev)
    _native.check(L.skfa_sys_dot(_native.ptr(r), _native.ptr(r), _native.ptr(sys_ptr), nsys, _native.ptr(rr), None,
                                 _native.stream()))
    _native.check(L.skfa_sys_dot(_native.ptr(b), _native.ptr(b), _native.ptr(sys_ptr), nsys, _native.ptr(bb), None,
                                 _native.stream()))
    return torch.sqrt(rr / bb).cpu().numpy()


def solve_dirichlet(matrix, rhs, boundary_dofs, *, preconditioner="jacobi", tol=1e-8, max_iter=20000):
    """Solution with zero Dirichlet values on ``boundary_dofs``
    (examples/scikit-fem-gallery/comparison.py:80-87), by FP64 PCG on the GPU.

    ``matrix`` is a scipy sparse matrix or a DeviceCSR; ``rhs`` a NumPy or
    device vector.  Returns a NumPy array (the reference's convention)."""
    torch = _torch()
    A = matrix if isinstance(matrix, DeviceCSR) else DeviceCSR.from_scipy(matrix)
    n = A.shape[0]
    dev = A.data.device
    keep = torch.ones(n, dtype=torch.uint8, device=dev)
    bd = torch.as_tensor(np.unique(np.asarray(boundary_dofs, dtype=np.int64)), device=dev)
    keep[bd] = 0
    Af, _ = A.condense(keep)
    b = rhs if isinstance(rhs, torch.Tensor) else torch.as_tensor(np.asarray(rhs, dtype=np.float64), device=dev)
    free = torch.nonzero(keep, as_tuple=True)[0]
    bf = b[free].contiguous()
    M = JacobiPreconditioner(Af) if preconditioner == "jacobi" else preconditioner(Af)
    res = pcg(Af, bf, [0, Af.shape[0]], M, tol=tol, max_iter=max_iter)
    solution = torch.zeros(n, dtype=torch.float64, device=dev)
    solution[free] = res.x
    return solution.cpu().numpy()
