"""Learned factorised sparse approximate inverse (FSAI) preconditioner.

A 16-128-128-1 ReLU network (18,817 parameters) maps features of each
lower-triangular nonzero (i, j) of a batch of SPD systems to an entry of the
factor Gt of At = D^{-1/2} A D^{-1/2} (off-diagonal Gt_ij = o_ij, diagonal
Gt_ii = exp(o_ii)); the preconditioner is M^{-1} = G^T G with
G = Gt D^{-1/2}, SPD by construction.  A network whose output is zero is
exactly Jacobi.  Training minimises Kaporin's condition number of
Gt At Gt^T (log form, per system, averaged over a minibatch of systems):
  L = log(tr(Gt At Gt^T) / n) - (2/n) sum_i log Gt_ii.
Everything runs in the library's HIP kernels: features, the objective and
its gradient (FP64), forward/backward on BF16 MFMA with FP32 accumulation,
split-K weight gradients with a fixed reduction order, and Adam.

No counterpart in the reference (solver policy is external,
python/skfemntv/capabilities.py:67).
"""

from __future__ import annotations

import ctypes
import os
import time
from dataclasses import dataclass, field

import numpy as np

from . import _native
from .solver import DeviceCSR, FactorPreconditioner

NF = _native.FSAI_FEATURES
NPARAMS = _native.FSAI_PARAMS
HIDDEN = 128
_OFF = {"W1": (0, (NF, HIDDEN)), "b1": (NF * HIDDEN, (HIDDEN,)),
        "W2": (NF * HIDDEN + HIDDEN, (HIDDEN, HIDDEN)),
        "b2": (NF * HIDDEN + HIDDEN + HIDDEN * HIDDEN, (HIDDEN,)),
        "w3": (NF * HIDDEN + 2 * HIDDEN + HIDDEN * HIDDEN, (HIDDEN,)),
        "b3": (NF * HIDDEN + 3 * HIDDEN + HIDDEN * HIDDEN, (1,))}
WEIGHTS_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "fsai_mlp.npz")


def split_params(flat):
    """Flat parameter vector -> dict of named arrays (views)."""
    return {k: flat[o:o + int(np.prod(s))].reshape(s) for k, (o, s) in _OFF.items()}


def init_params(seed=0):
    """He-initialised hidden layers, zero output layer (the network starts as Jacobi)."""
    rng = np.random.default_rng(seed)
    flat = np.zeros(NPARAMS, dtype=np.float32)
    p = split_params(flat)
    p["W1"][...] = rng.standard_normal((NF, HIDDEN)) * np.sqrt(2. / NF)
    p["W2"][...] = rng.standard_normal((HIDDEN, HIDDEN)) * np.sqrt(2. / HIDDEN)
    return flat


class FSAINetwork:
    """Device parameters, BF16 fragments and Adam state of the MLP."""

    def __init__(self, params=None, seed=0):
        import torch
        dev = _native.device()
        flat = init_params(seed) if params is None else np.asarray(params, dtype=np.float32).reshape(NPARAMS)
        self.params = torch.from_numpy(flat.copy()).to(dev)
        self.frag = torch.empty(_native.FSAI_FRAG, dtype=torch.int16, device=dev)
        self.m = torch.zeros(NPARAMS, dtype=torch.float32, device=dev)
        self.v = torch.zeros(NPARAMS, dtype=torch.float32, device=dev)
        self.step = 0
        self.pack()

    @classmethod
    def pretrained(cls, path=WEIGHTS_PATH):
        with np.load(path, allow_pickle=False) as z:
            return cls(params=z["params"])

    def save(self, path, **meta):
        os.makedirs(os.path.dirname(path), exist_ok=True)
        extra = {k: np.asarray(v) for k, v in meta.items()}
        np.savez(path, params=self.params.cpu().numpy(), **extra)

    def pack(self):
        _native.check(_native.lib().skfa_mlp_pack(_native.ptr(self.params), _native.ptr(self.frag), _native.stream()))

    def forward(self, features, nnz, out=None):
        import torch
        o = torch.empty(max(nnz, 1), dtype=torch.float32, device=features.device) if out is None else out
        _native.check(_native.lib().skfa_mlp_forward(_native.ptr(self.params), _native.ptr(self.frag),
                                                     _native.ptr(features), nnz, _native.ptr(o), _native.stream()))
        return o

    def backward(self, features, grad_out, nnz, workspace=None):
        import torch
        L = _native.lib()
        need = int(L.skfa_mlp_workspace_bytes(nnz))
        if workspace is None or workspace.numel() < need:
            workspace = torch.empty(need, dtype=torch.uint8, device=features.device)
        grad = torch.empty(NPARAMS, dtype=torch.float32, device=features.device)
        _native.check(L.skfa_mlp_backward(_native.ptr(self.params), _native.ptr(self.frag), _native.ptr(features),
                                          _native.ptr(grad_out), nnz, _native.ptr(workspace), workspace.numel(),
                                          _native.ptr(grad), _native.stream()))
        return grad, workspace

    def adam(self, grad, lr, beta1=0.9, beta2=0.999, eps=1e-8, clip_scale=1.0):
        self.step += 1
        _native.check(_native.lib().skfa_mlp_adam(_native.ptr(self.params), _native.ptr(grad), _native.ptr(self.m),
                                                  _native.ptr(self.v), float(lr), float(beta1), float(beta2),
                                                  float(eps), self.step, float(clip_scale), _native.ptr(self.frag),
                                                  _native.stream()))


class FSAISystems:
    """Lower pattern, scaling and features of a block-diagonal batch of SPD
    systems (device).  ``xyz`` (n,3) holds the coordinates of each unknown."""

    def __init__(self, A: DeviceCSR, sys_ptr, xyz=None):
        import torch
        self.A = A
        dev = A.data.device
        self.sys_ptr = sys_ptr if isinstance(sys_ptr, torch.Tensor) else torch.as_tensor(
            np.asarray(sys_ptr, dtype=np.int64), device=dev)
        self.nsys = int(self.sys_ptr.shape[0]) - 1
        n = A.shape[0]
        self.n = n
        t0 = time.perf_counter()
        self.pattern = A.lower()
        nl = self.pattern.nnz
        self.nnz = nl
        self.rsqrt_diag = torch.empty(n, dtype=torch.float64, device=dev)
        self.row_stats = torch.empty(4 * n, dtype=torch.float64, device=dev)
        self.features = torch.empty((max(nl, 1), NF), dtype=torch.float32, device=dev)
        self.xyz = None if xyz is None else xyz.to(torch.float64).contiguous()
        _native.check(_native.lib().skfa_fsai_features(
            _native.ptr(A.indptr), _native.ptr(A.indices), _native.ptr(A.data), n, _native.ptr(self.xyz),
            _native.ptr(self.pattern.lptr), _native.ptr(self.pattern.lidx), _native.ptr(self.rsqrt_diag),
            _native.ptr(self.row_stats), _native.ptr(self.features), _native.stream()))
        self._work = None
        self._row_sys = None
        self.setup_seconds = time.perf_counter() - t0

    def kaporin(self, outputs, grad_scale=None, with_grad=True):
        """(loss per system (host), gradient wrt outputs (device) or None)."""
        import torch
        dev = self.A.data.device
        if self._work is None:
            self._work = torch.empty(3 * self.nnz + self.n + 3 * self.nsys + 1, dtype=torch.float64, device=dev)
            self._row_sys = torch.empty(max(self.n, 1), dtype=torch.int32, device=dev)
        loss = torch.empty(self.nsys, dtype=torch.float64, device=dev)
        go = torch.empty(max(self.nnz, 1), dtype=torch.float32, device=dev) if with_grad else None
        scale = 1.0 / self.nsys if grad_scale is None else grad_scale
        A, p = self.A, self.pattern
        _native.check(_native.lib().skfa_fsai_kaporin(
            _native.ptr(A.indptr), _native.ptr(A.indices), _native.ptr(A.data), self.n, _native.ptr(self.sys_ptr),
            self.nsys, _native.ptr(p.lptr), _native.ptr(p.lidx), _native.ptr(self.rsqrt_diag), _native.ptr(outputs),
            float(scale), _native.ptr(self._work), _native.ptr(self._row_sys), _native.ptr(loss), _native.ptr(go),
            _native.stream()))
        return loss, go

    def factor(self, outputs):
        """G = Gt D^{-1/2} (FP32, lower pattern) from network outputs."""
        import torch
        g = torch.empty(max(self.nnz, 1), dtype=torch.float32, device=self.A.data.device)
        _native.check(_native.lib().skfa_fsai_factor(_native.ptr(outputs), _native.ptr(self.pattern.lptr),
                                                     _native.ptr(self.pattern.lidx), _native.ptr(self.rsqrt_diag),
                                                     self.n, None, _native.ptr(g), None, _native.stream()))
        return g[:self.nnz]


class LearnedFSAI:
    """Preconditioner factory: network inference on a batch's features."""

    def __init__(self, network: FSAINetwork):
        self.network = network

    def __call__(self, systems: FSAISystems):
        out = self.network.forward(systems.features, systems.nnz)
        return FactorPreconditioner(systems.pattern, systems.factor(out))


@dataclass
class TrainingHistory:
    epochs: list = field(default_factory=list)
    train_loss: list = field(default_factory=list)
    valid_loss: list = field(default_factory=list)
    seconds: float = 0.
    steps: int = 0


def mean_loss(network, groups):
    total, count = 0., 0
    for g in groups:
        out = network.forward(g.features, g.nnz)
        loss, _ = g.kaporin(out, with_grad=False)
        total += float(loss.sum())
        count += g.nsys
    return total / max(count, 1)


def train(network, groups, *, epochs=20, lr=2e-3, lr_final=1e-4, clip=1.0, valid=(), seed=0, log=None):
    """Adam on the mean Kaporin objective; each group (an FSAISystems of a
    few systems) is one minibatch; groups are visited in a seeded random
    order every epoch, with a cosine learning-rate schedule."""
    import torch
    hist = TrainingHistory()
    rng = np.random.default_rng(seed)
    total_steps = epochs * len(groups)
    workspace = None
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    step = 0
    for epoch in range(epochs):
        order = rng.permutation(len(groups))
        acc = 0.
        for gi in order:
            g = groups[gi]
            out = network.forward(g.features, g.nnz)
            loss, go = g.kaporin(out)
            grad, workspace = network.backward(g.features, go, g.nnz, workspace)
            # global-norm clipping; the 75 KB gradient is read back for the norm
            gnorm = float(np.linalg.norm(grad.cpu().numpy().astype(np.float64)))
            scale = 1.0 if gnorm <= clip else clip / gnorm
            if not np.isfinite(gnorm):
                raise FloatingPointError("non-finite gradient in FSAI training")
            frac = step / max(total_steps - 1, 1)
            rate = lr_final + 0.5 * (lr - lr_final) * (1. + np.cos(np.pi * frac))
            network.adam(grad, rate, clip_scale=scale)
            acc += float(loss.mean())
            step += 1
        hist.epochs.append(epoch)
        hist.train_loss.append(acc / len(groups))
        if valid:
            hist.valid_loss.append(mean_loss(network, valid))
        if log:
            log(f"epoch {epoch}: train {hist.train_loss[-1]:.5f}"
                + (f" valid {hist.valid_loss[-1]:.5f}" if valid else ""))
    torch.cuda.synchronize()
    hist.seconds = time.perf_counter() - t0
    hist.steps = step
    return hist


def split_groups(batch_A, sys_ptr_host, xyz, group_size):
    """FSAISystems per consecutive group of systems of a block-diagonal batch."""
    groups = []
    S = len(sys_ptr_host) - 1
    for s0 in range(0, S, group_size):
        s1 = min(S, s0 + group_size)
        r0, r1 = int(sys_ptr_host[s0]), int(sys_ptr_host[s1])
        A = batch_A.block(r0, r1)
        sp = np.asarray(sys_ptr_host[s0:s1 + 1], dtype=np.int64) - r0
        groups.append(FSAISystems(A, sp, None if xyz is None else xyz[r0:r1]))
    return groups
