"""Batched anisotropic-diffusion systems: many P1 tetrahedral problems on
uneven Kuhn-split boxes, assembled together and condensed on the GPU.

The systems are concatenated into one mesh whose components do not touch, so
one gather plan and one assembly launch pair (element payload pass + CSR
row-gather) produce every matrix at once as a block-diagonal CSR; the row
gather also sums the load vector and writes both through precomputed
full -> condensed position maps, so the Dirichlet condensation of the
boundary vertices costs no pass of its own (skfa_assemble_bilinear_load).  After one
eager call the whole sequence can be replayed as a HIP graph.

Per system s (seeded by ``np.random.default_rng([seed, s])``):
  * mesh: ``MeshTet.init_tensor`` of linspace(0,1,nx+1) x (ny+1) x (nz+1)
    (python/skfemntv/basis.py:599-616 numbering), interior vertices jittered
    per axis by U(-jitter h_axis, +jitter h_axis);
  * coefficient: per element A = Q diag(lambda) Q^T, Q a Haar rotation (QR of
    a Gaussian, sign-fixed, det +1), lambda_i = 10^U(-2, 2);
  * form dot(mul(A, grad u), grad v), load f = 1, default 4-point rule;
  * zero Dirichlet values on boundary vertices, condensed as
    examples/scikit-fem-gallery/comparison.py:80-87 does.
"""

from __future__ import annotations

import time
from dataclasses import dataclass

import numpy as np

from . import _native
from .basis import Basis
from .element import ElementTetP1, ElementVector
from .mesh import MeshTet


def random_rotations(E, rng):
    """(E,3,3) Haar-random rotations: the rotation of a uniformly random unit
    quaternion (a normalised 4-D Gaussian)."""
    q = rng.standard_normal((4, E))
    q /= np.sqrt((q * q).sum(axis=0))
    w, x, y, z = q
    R = np.empty((E, 3, 3))
    R[:, 0, 0] = 1. - 2. * (y * y + z * z)
    R[:, 0, 1] = 2. * (x * y - w * z)
    R[:, 0, 2] = 2. * (x * z + w * y)
    R[:, 1, 0] = 2. * (x * y + w * z)
    R[:, 1, 1] = 1. - 2. * (x * x + z * z)
    R[:, 1, 2] = 2. * (y * z - w * x)
    R[:, 2, 0] = 2. * (x * z - w * y)
    R[:, 2, 1] = 2. * (y * z + w * x)
    R[:, 2, 2] = 1. - 2. * (x * x + y * y)
    return R


def random_spd_tensors(E, rng, exponent_range=(-2., 2.)):
    """(E,3,3) SPD tensors Q diag(10^u) Q^T, Q Haar-random, u uniform in
    ``exponent_range`` (exactly symmetric)."""
    R = random_rotations(E, rng)
    lam = 10. ** rng.uniform(exponent_range[0], exponent_range[1], size=(3, E))
    A = np.empty((E, 3, 3))
    for i in range(3):
        for j in range(i, 3):
            v = R[:, i, 0] * R[:, j, 0] * lam[0] + R[:, i, 1] * R[:, j, 1] * lam[1] + R[:, i, 2] * R[:, j, 2] * lam[2]
            A[:, i, j] = v
            A[:, j, i] = v
    return A


def box_system(shape, rng, jitter=0.2):
    """(p, t, interior mask, A) of one jittered Kuhn box with per-element SPD A."""
    nx, ny, nz = (int(v) for v in shape)
    m = MeshTet.init_tensor(np.linspace(0., 1., nx + 1), np.linspace(0., 1., ny + 1), np.linspace(0., 1., nz + 1))
    p = m.p.copy()
    interior = np.all((p > 1e-12) & (p < 1. - 1e-12), axis=0)
    h = np.array([1. / nx, 1. / ny, 1. / nz])[:, None]
    k = int(interior.sum())
    p[:, interior] += rng.uniform(-jitter, jitter, size=(3, k)) * h
    A = random_spd_tensors(m.t.shape[1], rng)
    return p, m.t, interior, A


def batch_shapes(count, lo, hi, seed):
    """(count, 3) box sizes with i.i.d. per-axis cells in {lo..hi}."""
    rng = np.random.default_rng(seed)
    return rng.integers(lo, hi + 1, size=(count, 3))


@dataclass
class BatchTimings:
    mesh_s: float = 0.
    basis_s: float = 0.
    plan_s: float = 0.
    condense_s: float = 0.


class AnisotropicBatch:
    """Block-diagonal batch of condensed P1 anisotropic-diffusion systems."""

    def __init__(self, shapes, seed=0, jitter=0.2):
        import torch
        from ._assembly import NativeCrossBilinearForm, upload
        from .solver import DeviceCSR
        self.shapes = np.asarray(shapes, dtype=np.int64).reshape(-1, 3)
        self.seed, self.jitter = int(seed), float(jitter)
        self.timings = BatchTimings()
        t0 = time.perf_counter()
        ps, ts, keeps, As = [], [], [], []
        node_off = [0]
        elem_off = [0]
        for s, shape in enumerate(self.shapes):
            rng = np.random.default_rng([self.seed, s])
            p, t, interior, A = box_system(shape, rng, self.jitter)
            ts.append(t + node_off[-1])
            ps.append(p)
            keeps.append(interior)
            As.append(A)
            node_off.append(node_off[-1] + p.shape[1])
            elem_off.append(elem_off[-1] + t.shape[1])
        self.mesh = MeshTet(np.concatenate(ps, axis=1), np.concatenate(ts, axis=1))
        self.node_offsets = np.asarray(node_off, dtype=np.int64)
        self.element_offsets = np.asarray(elem_off, dtype=np.int64)
        keep = np.concatenate(keeps)
        tensors = np.concatenate(As, axis=0)
        self.timings.mesh_s = time.perf_counter() - t0
        t0 = time.perf_counter()
        self.basis = Basis(self.mesh, ElementVector(ElementTetP1(), dim=1))
        self.timings.basis_s = time.perf_counter() - t0
        t0 = time.perf_counter()
        self.form = NativeCrossBilinearForm(self.basis, self.basis)
        torch.cuda.synchronize()
        self.timings.plan_s = time.perf_counter() - t0
        dev = _native.device()
        self.tensors = upload(tensors)
        self._one = torch.ones(1, dtype=torch.float64, device=dev)
        t0 = time.perf_counter()
        # condensation structure: condense positions once to get the value map
        plan = self.form.plan
        full = DeviceCSR(plan.indptr, plan.indices[:plan.nnz],
                         torch.arange(plan.nnz, dtype=torch.float64, device=dev), (plan.nrows, plan.nrows))
        keep_dev = torch.from_numpy(keep.astype(np.uint8)).to(dev)
        cond, _ = full.condense(keep_dev)
        # output maps of the gather: full CSR entry -> condensed entry and
        # node row -> condensed row (-1: a boundary row or column, dropped)
        self.value_map = torch.full((plan.nnz,), -1, dtype=torch.int32, device=dev)
        self.value_map[cond.data.to(torch.int64)] = torch.arange(cond.nnz, dtype=torch.int32, device=dev)
        free = torch.nonzero(keep_dev, as_tuple=True)[0]
        self.row_map = torch.full((plan.nrows,), -1, dtype=torch.int32, device=dev)
        self.row_map[free] = torch.arange(free.numel(), dtype=torch.int32, device=dev)
        self.A = DeviceCSR(cond.indptr, cond.indices, torch.empty_like(cond.data), cond.shape)
        self.b = torch.empty(cond.shape[0], dtype=torch.float64, device=dev)
        kept_per_system = np.add.reduceat(keep.astype(np.int64), self.node_offsets[:-1])
        self.sys_ptr_host = np.concatenate(([0], np.cumsum(kept_per_system))).astype(np.int64)
        self.sys_ptr = torch.from_numpy(self.sys_ptr_host).to(dev)
        self.xyz = torch.from_numpy(np.ascontiguousarray(self.mesh.p[:, keep].T)).to(dev)
        torch.cuda.synchronize()
        self.timings.condense_s = time.perf_counter() - t0
        self._graph = None

    @property
    def nsystems(self):
        return len(self.shapes)

    @property
    def nelements(self):
        return int(self.element_offsets[-1])

    def _assemble(self):
        # condensed matrix and load vector straight from one gather
        # (skfa_assemble_bilinear_load with output maps)
        self.form.assemble_tensor_device(self.tensors, symmetric=True, out=self.A.data,
                                         load=(self._one, self.b, self.value_map, self.row_map))

    def assemble(self, graph=False):
        """Assemble every system (condensed A values and b, in place).  With
        ``graph=True`` the launch sequence is captured once into a HIP graph
        and replayed."""
        import torch
        if not graph:
            self._assemble()
            return self.A, self.b
        if self._graph is None:
            self._assemble()   # allocates the payload buffer outside the capture
            torch.cuda.synchronize()
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g):
                self._assemble()
            self._graph = g
        self._graph.replay()
        return self.A, self.b

    def system_csr(self, s):
        """System s as a host scipy CSR (tests)."""
        r0, r1 = int(self.sys_ptr_host[s]), int(self.sys_ptr_host[s + 1])
        A = self.A.to_scipy()
        return A[r0:r1, r0:r1].tocsr(), self.b[r0:r1].cpu().numpy()
