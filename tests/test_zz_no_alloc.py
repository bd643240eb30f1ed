"""The no-allocation bilinear entry point against the allocating one.

skfa_assemble_bilinear_noalloc keeps the element payload and element tables
in a fixed device array of the library; its results must be bit-identical
to skfa_assemble_bilinear's (same kernels, same payload values) on two calls
in a row, the second with a smaller mesh, and a plan whose payload does not
fit must be rejected with an error.  Both entry points are called through
the C ABI with the same arguments.
"""

import sys

import numpy as np
import pytest

from oracle import host

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()


def call(basis, plan, out, noalloc, **coef):
    import ctypes
    from skfemntv import _native
    d = basis.device_data()
    L = _native.lib()
    flags = _native.FLAG_SYMMETRIC_TENSOR if coef.pop("symmetric", False) else 0
    precision = coef.pop("precision", 0)
    entry = L.skfa_assemble_bilinear_noalloc if noalloc else L.skfa_assemble_bilinear
    get = lambda k: _native.ptr(coef[k]) if k in coef else None  # noqa: E731
    rc = entry(plan.handle, _native.ptr(d["xyz"]), d["xyz"].shape[0], _native.ptr(d["vconn"]), None, 4,
               _native.ptr(d["tables"]), len(basis.W), _native.ptr(d["rtab"]), _native.ptr(d["rtab32"]),
               d["nqg"], d["nqm"], get("value"), get("gradient"), get("symgrad"), get("divergence"),
               get("tensor"), 0, precision, flags, _native.ptr(out), _native.stream())
    return rc, (L.skfa_last_error().decode() if rc else "")


def p2_case(n, seed, components=1):
    import skfemntv as sk
    ax = np.linspace(0., 1., n + 1)
    p, t = host.kuhn_mesh(ax, ax, ax)
    p = host.jitter_interior(p, 1. / n, 0.2, seed)
    p2, t2 = host.p2_mesh(p, t)
    return sk.Basis(sk.MeshTet2(p2, t2), sk.ElementVector(sk.ElementTetP2(), dim=components),
                    quadrature=host.tet_rule_4())


def spd(E, seed):
    rng = np.random.default_rng(seed)
    g = rng.standard_normal((E, 3, 3))
    return np.einsum("eij,ekj->eik", g, g) + 3 * np.eye(3)


def compare(basis, plan, **coef):
    import torch
    from skfemntv import _native
    dev = _native.device()
    a = torch.full((max(plan.nnz, 1),), np.nan, dtype=torch.float64, device=dev)
    b = torch.full_like(a, np.nan)
    rc, msg = call(basis, plan, a, False, **dict(coef))
    assert rc == 0, msg
    rc, msg = call(basis, plan, b, True, **dict(coef))
    assert rc == 0, msg
    torch.cuda.synchronize()
    assert bool(torch.isfinite(b).all())
    assert bool((a == b).all()), float((a - b).abs().max())


@pytest.mark.parametrize("precision", [0, 1])
def test_noalloc_matches_allocating_entry_on_two_calls(precision):
    from skfemntv._assembly import basis_plan, upload
    for n, seed in ((6, 1), (3, 2)):   # the second call has a smaller mesh
        say(f"noalloc P2 symmetric tensor n={n} precision={precision}")
        basis = p2_case(n, seed)
        plan = basis_plan(basis)
        E = basis.mesh.nelements
        compare(basis, plan, tensor=upload(spd(E, seed)), symmetric=True, precision=precision)


def test_noalloc_other_payload_layouts():
    import skfemntv as sk
    from skfemntv._assembly import Plan, basis_plan, upload
    rng = np.random.default_rng(4)
    say("noalloc P2 nonsymmetric tensor + mass")
    basis = p2_case(4, 5)
    E = basis.mesh.nelements
    compare(basis, basis_plan(basis), tensor=upload(rng.standard_normal((E, 3, 3))),
            value=upload(rng.uniform(0.5, 2., E)))
    say("noalloc vector P2 elasticity")
    vb = p2_case(3, 6, components=3)
    E = vb.mesh.nelements
    compare(vb, basis_plan(vb), symgrad=upload(rng.uniform(0.5, 2., E)), divergence=upload(rng.uniform(0.5, 2., E)))
    say("noalloc P1 element-table gather (one row per lane)")
    ax = np.linspace(0., 1., 7)
    pb = sk.Basis(sk.MeshTet.init_tensor(ax, ax, ax), sk.ElementVector(sk.ElementTetP1(), dim=1))
    d = pb.device_data()
    E = pb.mesh.nelements
    plan = Plan(d["node_conn"], E, 4, d["nnodes"], 1, min_waves=0)
    compare(pb, plan, gradient=upload(rng.uniform(0.5, 2., E)), value=upload(rng.uniform(0.5, 2., E)))


def test_noalloc_rejects_a_payload_larger_than_its_scratch():
    import torch
    import skfemntv as sk
    from skfemntv import _native
    from skfemntv._assembly import basis_plan, upload
    say("noalloc rejects 60^3 P1 with a nonsymmetric tensor (80 B x 1.296 M elements)")
    ax = np.linspace(0., 1., 61)
    basis = sk.Basis(sk.MeshTet.init_tensor(ax, ax, ax), sk.ElementVector(sk.ElementTetP1(), dim=1))
    plan = basis_plan(basis)
    E = basis.mesh.nelements
    out = torch.full((plan.nnz,), 7., dtype=torch.float64, device=_native.device())
    rc, msg = call(basis, plan, out, True, tensor=upload(np.random.default_rng(7).standard_normal((E, 3, 3))))
    assert rc != 0 and "scratch of the no-allocation entry point" in msg, (rc, msg)
    torch.cuda.synchronize()
    assert bool((out == 7.).all())   # nothing was launched
