"""GPU parity of the standalone geometry tabulation and of the general
(per-point) linear-form kernel against the CPU oracle.

* ``skfa_tabulate_geometry`` replaces tabulate_basis_geometry
  (cpp/src/assembly/basis_geometry.cpp:49-209): all seven outputs (shape,
  grads, dx, x, det, tol, cond) are compared with the oracle's restatement,
  which is bit-exact against the reference's C++ (tests/test_oracle.py), on a
  jittered P2 mesh with the 11-point Keast rule and on a curved P2 mesh;
  rejections must name the same cell and quadrature point.
* ``k_assemble_linear`` replaces LinearFormAssembler::assemble with
  non-constant coefficients (cpp/src/assembly/linear_form_assembler.cpp:
  41-99): per-point f.v and G:grad(v) for P1, P2 and vector P2.
"""

import sys

import numpy as np
import pytest

import oracle
from oracle import host, pipeline

pytestmark = pytest.mark.gpu

GEO_TOL = 1e-14   # relative, per output array
FP64_TOL = 1e-12


def say(msg):
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()


def jittered_p2(n, amount, seed):
    ax = np.linspace(0., 1., n + 1)
    p, t = host.kuhn_mesh(ax, ax, ax)
    p = host.jitter_interior(p, 1. / n, amount, seed)
    return host.p2_mesh(p, t)


def rel(a, b):
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.))


def device_geometry(basis):
    import torch
    from skfemntv import _native
    E, Q = basis.mesh.nelements, len(basis.W)
    n = basis._ref_shape.shape[1]
    dev = _native.device()
    out = {k: torch.full(s, np.nan, dtype=torch.float64, device=dev) for k, s in (
        ("det", (E, Q)), ("tol", (E, Q)), ("cond", (E, Q)), ("dx", (E, Q)), ("x", (E, Q, 3)),
        ("shape", (E, Q, n)), ("grads", (E, Q, n, 3)))}
    basis._geometry_call(**out)
    return {k: v.cpu().numpy() for k, v in out.items()}


def oracle_geometry(p, t, X, W):
    shape, grad = host.tabulate_p2(X)
    gshape, ggrad = host.tabulate_p2(X) if t.shape[0] == 10 else host.tabulate_p1(X)
    (o_shape, o_grad, o_dx, o_x, o_det, o_tol, o_cond), bad = oracle.tabulate_geometry(
        p, t, shape, grad, gshape, ggrad, W)
    return {"shape": o_shape, "grads": o_grad, "dx": o_dx, "x": o_x, "det": o_det, "tol": o_tol,
            "cond": o_cond}, bad


@pytest.mark.parametrize("curved", [False, True])
def test_tabulate_geometry_all_outputs_match_oracle(curved):
    import skfemntv as sk
    say(f"tabulate_geometry p2 keast11 curved={curved}")
    p2, t2 = jittered_p2(4, 0.2, 11)
    if curved:
        V = int(t2[:4].max()) + 1
        rng = np.random.default_rng(12)
        p2 = p2.copy()
        p2[:, V:] += rng.uniform(-0.02, 0.02, size=(3, p2.shape[1] - V))
    basis = sk.Basis(sk.MeshTet2(p2, t2), sk.ElementVector(sk.ElementTetP2(), dim=1))
    assert len(basis.W) == 11 and basis._affine == (not curved)
    got = device_geometry(basis)
    X, W = host.tet_rule_keast11()
    ref, bad = oracle_geometry(p2, t2, X, W)
    assert bad[0] < 0
    for k in ("shape", "grads", "dx", "x", "det", "tol", "cond"):
        assert np.isfinite(got[k]).all(), k
        err = rel(got[k], ref[k])
        assert err <= GEO_TOL, (k, err)
    # the diagnostics the Basis derives from them
    diag = basis.geometry_diagnostics
    assert diag.maximum_condition_number == pytest.approx(ref["cond"].max(), rel=1e-14)
    # affine cells have one condition number at every point, equal up to
    # rounding, so the arg-max point may differ: check the value it names
    worst = ref["cond"][diag.worst_condition_element, diag.worst_condition_quadrature_point]
    assert worst == pytest.approx(ref["cond"].max(), rel=1e-14)
    assert diag.minimum_determinant == pytest.approx(ref["det"].min(), rel=1e-14)
    assert diag.maximum_determinant == pytest.approx(ref["det"].max(), rel=1e-14)
    # closed form: the volume of the unit cube (curved edge nodes on the
    # boundary change it)
    if not curved:
        assert abs(got["dx"].sum() - 1.) <= 1e-13


def test_geometry_rejection_names_the_same_cell_and_point():
    import skfemntv as sk
    say("geometry rejection cell/qp vs oracle")
    p2, t2 = jittered_p2(2, 0.1, 13)
    X, W = host.tet_rule_keast11()
    # fold one edge node of cell 5 through the opposite face: the Jacobian
    # changes sign inside the cell
    e = 5
    node = int(t2[4, e])
    p2 = p2.copy()
    verts = p2[:, t2[:4, e]]
    p2[:, node] = verts[:, 2] + 1.5 * (verts[:, 3] - verts[:, 2]) + 2.0 * (verts.mean(axis=1) - verts[:, 0])
    ref, (bc, bq, reason) = oracle_geometry(p2, t2, X, W)
    assert bc >= 0
    with pytest.raises(ValueError) as info:
        sk.Basis(sk.MeshTet2(p2, t2), sk.ElementVector(sk.ElementTetP2(), dim=1))
    msg = str(info.value)
    assert f"cell={bc}, quadrature_point={bq}," in msg, (msg, bc, bq, reason)
    assert ("reason=orientation_change" if reason == 2 else "reason=near_singular_or_non_finite") in msg
    det = float(msg.split("determinant=")[1].split(",")[0])
    assert det == pytest.approx(ref["det"][bc, bq], rel=1e-13)


@pytest.mark.parametrize("quadratic,C", [(False, 1), (True, 1), (True, 3)])
def test_general_linear_form_per_point_f_and_G_match_oracle(quadratic, C):
    import skfemntv as sk
    from skfemntv._assembly import NativeLinearForm
    say(f"general linear form quadratic={quadratic} C={C}")
    n = 4 if quadratic else 6
    ax = np.linspace(0., 1., n + 1)
    p, t = host.kuhn_mesh(ax, ax, ax)
    p = host.jitter_interior(p, 1. / n, 0.2, 21)
    if quadratic:
        p, t = host.p2_mesh(p, t)
        X, W = host.tet_rule_keast11()
        mesh, elem = sk.MeshTet2(p, t), sk.ElementTetP2()
    else:
        X, W = host.tet_rule_4()
        mesh, elem = sk.MeshTet(p, t), sk.ElementTetP1()
    basis = sk.Basis(mesh, sk.ElementVector(elem, dim=C))
    E, Q = t.shape[1], len(W)
    rng = np.random.default_rng(22)
    f = rng.standard_normal((E, Q, C))
    G = rng.standard_normal((E, Q, C, 3))
    form = NativeLinearForm(basis)
    (shape, grads, dx, *_), _ = pipeline.basis_geometry(p, t, X, W, quadratic)
    N, dofs = pipeline.dofs_of(t, p.shape[1], 10 if quadratic else 4, C)
    for kw in ({"value": f}, {"gradient": G}, {"value": f, "gradient": G}):
        b, diag = form.assemble(**kw)
        ref = oracle.assemble_linear(shape, grads, dx, dofs, N, C, value=kw.get("value"),
                                     gradient=kw.get("gradient"))
        assert b.shape == (N,)
        err = rel(b, ref)
        assert err <= FP64_TOL, (sorted(kw), err)
        assert diag.entity_count == E and diag.quadrature_point_count == Q
    # closed forms on the per-point path: the basis sums to one and its
    # gradients sum to zero, so f = 1 gives C |Omega| and a constant G gives 0
    b1, _ = form.assemble(value=np.ones((E, Q, C)))
    assert abs(b1.sum() - C) <= 1e-12
    Gc = np.ascontiguousarray(np.broadcast_to(rng.standard_normal((C, 3)), (E, Q, C, 3)))
    b0, _ = form.assemble(gradient=Gc)
    assert np.abs(b0.reshape(-1, C).sum(axis=0)).max() <= 1e-12 * np.abs(b0).max() * N


@pytest.mark.parametrize("quadratic,min_waves", [(False, 0), (False, 2048), (True, 0)])
def test_bilinear_with_fused_load_matches_separate_calls(quadratic, min_waves):
    """skfa_assemble_bilinear_load: the matrix equals skfa_assemble_bilinear's
    bit for bit and the load vector equals the linear-form kernel's to
    rounding, whether the load rides in the P1 element-table gather (min
    waves 0: one row per lane, the table gather) or runs after the gather
    (small-mesh split layout, P2); both against the oracle's load vector
    (linear_form_assembler.cpp:41-137 with a constant f)."""
    import torch
    import skfemntv as sk
    from skfemntv._assembly import NativeCrossBilinearForm, NativeLinearForm, Plan, upload
    say(f"fused load, quadratic={quadratic}, min_waves={min_waves}")
    n = 6
    ax = np.linspace(0., 1., n + 1)
    p, t = host.kuhn_mesh(ax, ax, ax)
    p = host.jitter_interior(p, 1. / n, 0.2, 5)
    X, W = host.tet_rule_4()
    if quadratic:
        p2, t2 = host.p2_mesh(p, t)
        mesh, el, nl = sk.MeshTet2(p2, t2), sk.ElementTetP2(), 10
    else:
        p2, t2 = p, t
        mesh, el, nl = sk.MeshTet(p, t), sk.ElementTetP1(), 4
    basis = sk.Basis(mesh, sk.ElementVector(el, dim=1), quadrature=(X, W))
    E = t.shape[1]
    rng = np.random.default_rng(6)
    g = rng.standard_normal((E, 3, 3))
    A = np.einsum("eij,ekj->eik", g, g) + np.eye(3)
    form = NativeCrossBilinearForm(basis, basis)
    d = basis.device_data()
    form.plan = Plan(d["node_conn"], E, nl, d["nnodes"], 1, min_waves=min_waves)
    form.plan._tuned = True
    A_dev = upload(A)
    f = torch.full((1,), 2.5, dtype=torch.float64, device=A_dev.device)
    ref_vals = form.assemble_tensor_device(A_dev, symmetric=True).clone()
    load = torch.full((basis.N,), np.nan, dtype=torch.float64, device=A_dev.device)
    vals = form.assemble_tensor_device(A_dev, symmetric=True, load=(f, load)).clone()
    assert torch.equal(vals, ref_vals)
    lin = NativeLinearForm(basis).assemble_device(f, False, None, False).clone()
    b = load.cpu().numpy()
    assert rel(b, lin.cpu().numpy()) <= 1e-14
    bo = 2.5 * pipeline.load_vector(p2, t2, X, W, quadratic, [1.])
    assert rel(b, bo) <= FP64_TOL, rel(b, bo)
    say(f"load rel err vs oracle {rel(b, bo):.1e}")


@pytest.mark.parametrize("quadratic,min_waves", [(False, 0), (False, 2048), (True, 0)])
def test_fused_load_with_output_maps_condenses(quadratic, min_waves):
    """skfa_assemble_bilinear_load with output maps: the values and load land
    at the condensed positions K[I][:, I], b[I] (the Dirichlet condensation
    of comparison.py:80-87) with dropped entries untouched, equal bit for bit
    to condensing the unmapped results, on the fused P1 table path (min
    waves 0) and through the plan's full-size buffers otherwise."""
    import torch
    import skfemntv as sk
    from skfemntv._assembly import NativeCrossBilinearForm, Plan, upload
    say(f"fused load + output maps, quadratic={quadratic}, min_waves={min_waves}")
    n = 6
    ax = np.linspace(0., 1., n + 1)
    p, t = host.kuhn_mesh(ax, ax, ax)
    p = host.jitter_interior(p, 1. / n, 0.2, 7)
    X, W = host.tet_rule_4()
    if quadratic:
        p2, t2 = host.p2_mesh(p, t)
        mesh, el, nl = sk.MeshTet2(p2, t2), sk.ElementTetP2(), 10
    else:
        mesh, el, nl = sk.MeshTet(p, t), sk.ElementTetP1(), 4
    basis = sk.Basis(mesh, sk.ElementVector(el, dim=1), quadrature=(X, W))
    E, N = t.shape[1], basis.N
    rng = np.random.default_rng(8)
    g = rng.standard_normal((E, 3, 3))
    A = upload(np.einsum("eij,ekj->eik", g, g) + np.eye(3))
    form = NativeCrossBilinearForm(basis, basis)
    d = basis.device_data()
    plan = Plan(d["node_conn"], E, nl, d["nnodes"], 1, min_waves=min_waves)
    plan._tuned = True
    form.plan = plan
    dev = A.device
    f = torch.full((1,), 1.5, dtype=torch.float64, device=dev)
    full_load = torch.empty(N, dtype=torch.float64, device=dev)
    full = form.assemble_tensor_device(A, symmetric=True, load=(f, full_load)).clone()
    ip, ix = plan.host_csr_arrays()
    keep = rng.random(N) < 0.7
    rows = np.repeat(np.arange(N), np.diff(ip))
    kept = keep[rows] & keep[ix]
    vmap = np.full(len(ix), -1, dtype=np.int32)
    vmap[kept] = np.arange(int(kept.sum()), dtype=np.int32)
    rmap = np.full(N, -1, dtype=np.int32)
    rmap[keep] = np.arange(int(keep.sum()), dtype=np.int32)
    vals = torch.full((int(kept.sum()),), np.nan, dtype=torch.float64, device=dev)
    load = torch.full((int(keep.sum()),), np.nan, dtype=torch.float64, device=dev)
    form.assemble_tensor_device(A, symmetric=True, out=vals,
                                load=(f, load, upload(vmap), upload(rmap)))
    np.testing.assert_array_equal(vals.cpu().numpy(), full.cpu().numpy()[:len(ix)][kept])
    np.testing.assert_array_equal(load.cpu().numpy(), full_load.cpu().numpy()[keep])
