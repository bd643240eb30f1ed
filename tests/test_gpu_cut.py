"""GPU parity of the cut-cell (ragged quadrature) assemblers: the pattern of
CutBilinearFormAssembler / CutCrossAssembler bit-exact against the oracle's
restatement of the reference, and values (load vector, square and cross
forms) within 1e-12 relative Frobenius.  Cases cover cells without points,
cells with many points, a cell that repeats a node, and the case where no
cell has points.  Each case prints its name first.
"""

import sys

import numpy as np
import pytest

from oracle import cut

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12


def say(msg):
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()


def rel(mine, ref):
    ref = np.asarray(ref)
    d = np.linalg.norm(np.asarray(mine) - ref)
    return d / max(np.linalg.norm(ref), 1e-300)


@pytest.mark.parametrize("quadratic,components", [(False, 1), (True, 1), (False, 3), (True, 3)])
def test_cut_bilinear_matches_oracle(quadratic, components):
    import skfemntv as sk
    say(f"cut bilinear quadratic={quadratic} C={components}")
    dofs, off, sh, gr, w = cut.random_case(quadratic, components, seed=10 + components)
    rng = np.random.default_rng(1)
    v, g = rng.uniform(0.5, 2., len(w)), rng.uniform(0.5, 2., len(w))
    A = sk.CutBilinearFormAssembler(dofs, off, sh, gr, w)
    values = A.values
    for value, gradient in ((v, None), (None, g), (v, g)):
        A.assemble(value, gradient)
        ip, ix, ref = cut.cut_bilinear(dofs, off, sh, gr, w, value, gradient)
        np.testing.assert_array_equal(A.indptr, ip)
        np.testing.assert_array_equal(A.indices, ix)
        assert A.values is values
        assert rel(A.values, ref) <= FP64_TOL
    assert A.ndofs == int(dofs.max()) + 1 and A.color_count == 0


@pytest.mark.parametrize("components", [1, 3])
def test_cut_linear_matches_oracle(components):
    import skfemntv as sk
    say(f"cut linear C={components}")
    dofs, off, sh, gr, w = cut.random_case(True, components, seed=20 + components)
    rng = np.random.default_rng(2)
    P = len(w)
    A = sk.CutLinearFormAssembler(dofs, off, sh, gr, w)
    assert A.cell_count == dofs.shape[0] and A.point_count == P
    cases = [(rng.standard_normal(components), None), (rng.standard_normal((P, components)), None),
             (None, rng.standard_normal((components, 3))), (None, rng.standard_normal((P, components, 3))),
             (rng.standard_normal((P, components)), rng.standard_normal((components, 3)))]
    for value, gradient in cases:
        out, seconds = A.assemble(value, gradient)
        ref = cut.cut_linear(dofs, off, sh, gr, w, value, gradient)
        assert out.shape == ref.shape and seconds >= 0.
        assert rel(out, ref) <= FP64_TOL


@pytest.mark.parametrize("rk,ck,Cr,Cc,tensor", [
    ("value", "value", 1, 1, False), ("gradient", "gradient", 3, 3, False), ("value", "value", 3, 3, True),
    ("value", "gradient", 1, 3, True), ("gradient", "value", 3, 1, True), ("gradient", "gradient", 1, 3, True)])
def test_cut_cross_matches_oracle(rk, ck, Cr, Cc, tensor):
    import skfemntv as sk
    say(f"cut cross {rk}x{ck} Cr={Cr} Cc={Cc} tensor={tensor}")
    rd, off, rs, rgr, w = cut.random_case(False, Cr, seed=30)
    cd, _, cs, cgr, _ = cut.random_case(True, Cc, seed=30)
    rng = np.random.default_rng(3)
    P = len(w)
    if tensor:
        shape = (P, Cr) + ((3,) if rk == "gradient" else ()) + (Cc,) + ((3,) if ck == "gradient" else ())
        coef = rng.standard_normal(shape)
    else:
        coef = rng.uniform(0.5, 2., P)
    A = sk.CutCrossAssembler(rd, cd, off, rs, cs, w, rgr, cgr)
    A.assemble(coef, rk, ck)
    ip, ix, ref = cut.cut_cross(rd, cd, off, rs, cs, w, rgr, cgr, coef, rk, ck)
    np.testing.assert_array_equal(A.indptr, ip)
    np.testing.assert_array_equal(A.indices, ix)
    assert rel(A.values, ref) <= FP64_TOL
    assert A.rows == len(ip) - 1 and A.columns == int(cd.max()) + 1
    if not tensor and rk == "value":
        A.assemble(None, rk, ck)
        _, _, ref1 = cut.cut_cross(rd, cd, off, rs, cs, w, rgr, cgr, None, rk, ck)
        assert rel(A.values, ref1) <= FP64_TOL


@pytest.mark.parametrize("quadratic,dim", [(True, 3), (False, 2)])
def test_cut_cell_matrices_ragged_point_counts(quadratic, dim):
    """Point counts 0..140 per cell cross the node-matrix pass's 8-point
    batches and 64-point chunks; the square form (one table for rows and
    columns) and the cross form (two tables) take different kernels."""
    import skfemntv as sk
    say(f"cut node matrices, ragged counts quadratic={quadratic} dim={dim}")
    dofs, off, sh, gr, w = cut.random_case(quadratic, 1, seed=50 + dim, n=2, dim=dim, max_points=140)
    counts = np.diff(off)
    special = [1, 7, 8, 9, 15, 16, 17, 63, 64, 65, 128, 129]
    counts[:len(special)] = special
    off = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
    rng = np.random.default_rng(5)
    P = int(off[-1])
    nodes = dofs.shape[1]
    sh, gr, w = rng.standard_normal((P, nodes)), rng.standard_normal((P, nodes, dim)), rng.uniform(0.1, 1., P)
    v, g = rng.uniform(0.5, 2., P), rng.uniform(0.5, 2., P)
    A = sk.CutBilinearFormAssembler(dofs, off, sh, gr, w)
    for value, gradient in ((v, None), (None, g), (v, g)):
        A.assemble(value, gradient)
        _, _, ref = cut.cut_bilinear(dofs, off, sh, gr, w, value, gradient)
        assert rel(A.values, ref) <= FP64_TOL
    # cross form: P1 rows against the same cells' P2 columns, two tables
    rd, _, _, _, _ = cut.random_case(False, 1, seed=60, n=2)
    cd, _, _, _, _ = cut.random_case(True, 1, seed=60, n=2)
    rs, rgr = rng.standard_normal((P, 4)), rng.standard_normal((P, 4, dim))
    cs, cgr = rng.standard_normal((P, 10)), rng.standard_normal((P, 10, dim))
    X = sk.CutCrossAssembler(rd, cd, off, rs, cs, w, rgr, cgr)
    for kind in ("value", "gradient"):
        X.assemble(v, kind, kind)
        _, _, ref = cut.cut_cross(rd, cd, off, rs, cs, w, rgr, cgr, v, kind, kind)
        assert rel(X.values, ref) <= FP64_TOL


def test_cut_edge_cases():
    import skfemntv as sk
    say("cut edge cases: many points, repeated node, no points")
    dofs, off, sh, gr, w = cut.random_case(True, 1, seed=40, n=2, max_points=3)
    # one cell with 300 points
    rng = np.random.default_rng(4)
    counts = np.diff(off)
    counts[5] = 300
    off = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
    P = int(off[-1])
    sh, gr, w = rng.standard_normal((P, 10)), rng.standard_normal((P, 10, 3)), rng.uniform(0.1, 1., P)
    v = rng.uniform(0.5, 2., P)
    A = sk.CutBilinearFormAssembler(dofs, off, sh, gr, w)
    A.assemble(v, v)
    _, _, ref = cut.cut_bilinear(dofs, off, sh, gr, w, v, v)
    assert rel(A.values, ref) <= FP64_TOL
    # a cell that repeats a node: its entries add twice, as the reference's scatter does
    rep = dofs.copy()
    rep[5, 1, 0] = rep[5, 0, 0]
    B = sk.CutBilinearFormAssembler(rep, off, sh, gr, w)
    B.assemble(v, v)
    ip, ix, ref = cut.cut_bilinear(rep, off, sh, gr, w, v, v)
    np.testing.assert_array_equal(B.indices, ix)
    assert rel(B.values, ref) <= FP64_TOL
    # no cell has points: empty pattern, zero load vector
    z = np.zeros(dofs.shape[0] + 1, dtype=np.int64)
    e = np.zeros((0, 10))
    C = sk.CutBilinearFormAssembler(dofs, z, e, np.zeros((0, 10, 3)), np.zeros(0))
    C.assemble(np.zeros(0))
    assert len(C.indices) == 0 and len(C.indptr) == int(dofs.max()) + 2
    Lf = sk.CutLinearFormAssembler(dofs, z, e, np.zeros((0, 10, 3)), np.zeros(0))
    out, _ = Lf.assemble(np.ones(1))
    assert out.shape == (int(dofs.max()) + 1,) and not out.any()


def _full_domain_cut(basis, keep=None):
    """Cut quadrature whose points in every kept cell are the basis's own
    rule, mapped affinely (the reference's full-domain cut-basis test)."""
    import skfemntv as sk
    m = basis.mesh
    E = m.nelements
    X, W = basis.X, basis.W
    Q = len(W)
    keep = np.ones(E, dtype=bool) if keep is None else keep
    counts = np.where(keep, Q, 0)
    off = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
    corners = m.p[:, m.t[:4]]
    det = np.abs(np.linalg.det(np.moveaxis(corners[:, 1:4] - corners[:, [0]], -1, 0)))
    Xp = np.tile(X.T, (int(keep.sum()), 1))
    w = (det[keep][:, None] * W[None, :]).ravel()
    return sk.CutCellBasis(basis, sk.CutCellQuadrature.from_reference_points(m, off, Xp, w))


def _mesh(quadratic, n=3, seed=1):
    import skfemntv as sk
    from oracle import host
    ax = np.linspace(0., 1., n + 1)
    p, t = host.kuhn_mesh(ax, ax, ax)
    p = host.jitter_interior(p, 1. / n, 0.2, seed)
    if quadratic:
        p2, t2 = host.p2_mesh(p, t)
        return sk.MeshTet2(p2, t2)
    return sk.MeshTet(p, t)


@pytest.mark.parametrize("quadratic,components", [(False, 1), (True, 3)])
def test_cut_basis_full_domain_asm_matches_regular_basis(quadratic, components):
    import skfemntv as sk
    from skfemntv.helpers import ddot, dot, grad
    say(f"cut basis through asm, full domain, quadratic={quadratic} C={components}")
    mesh = _mesh(quadratic)
    elem = sk.ElementVector(sk.ElementTetP2() if quadratic else sk.ElementTetP1(), dim=components)
    regular = sk.Basis(mesh, elem)
    cutb = _full_domain_cut(regular)
    assert cutb.nelems == mesh.nelements and cutb.npoints == mesh.nelements * len(regular.W)
    mass_stiffness = sk.BilinearForm(lambda u, v, w: (1. + w.x[0]) * dot(u, v) + ddot(grad(u), grad(v)))
    load = sk.LinearForm(lambda v, w: dot(w.force, v))
    force = np.arange(1., components + 1.)
    A_cut = sk.asm(mass_stiffness, cutb).toarray()
    A_reg = sk.asm(mass_stiffness, regular).toarray()
    assert rel(A_cut, A_reg) <= FP64_TOL
    f_cut = sk.asm(load, cutb, force=force)
    f_reg = sk.asm(load, regular, force=force)
    assert f_cut.shape == (regular.N,) and rel(f_cut, f_reg) <= FP64_TOL
    # interpolation reproduces the field's own values and gradients
    u = 2. + 3. * regular.doflocs[0] - regular.doflocs[1]
    field = cutb.interpolate(u)
    x = cutb.points
    want = 2. + 3. * x[:, 0] - x[:, 1]
    got = field.value if components == 1 else field.value[0]
    assert np.abs(got - want).max() < 1e-12
    with pytest.raises(ValueError, match="symmetric-gradient"):
        sk.NativeBilinearForm(cutb).assemble(symmetric_gradient=1.)


def test_cut_basis_partial_domain_and_cross_blocks():
    import skfemntv as sk
    from skfemntv.helpers import div, dot
    say("cut basis: half the cells, Taylor-Hood divergence block and P1 mass")
    mesh = _mesh(True, n=3, seed=3)
    from skfemntv.basis import tet_rule_keast11
    X, W = tet_rule_keast11()
    vp2 = sk.Basis(mesh, sk.ElementVector(sk.ElementTetP2(), dim=3), quadrature=(X, W))
    p1 = sk.Basis(mesh, sk.ElementVector(sk.ElementTetP1(), dim=1), quadrature=(X, W))
    centroid = mesh.p[:, mesh.t[:4]].mean(axis=1)
    keep = centroid[0] < 0.5
    cut_v, cut_p = _full_domain_cut(vp2, keep), _full_domain_cut(p1, keep)
    coldiv = sk.BilinearForm(lambda u, v, w: div(u) * v)
    B = sk.asm(coldiv, cut_v, cut_p).toarray()
    # the same block from regular bases, restricted to the kept cells by a
    # per-element coefficient that is 1 on kept cells and 0 elsewhere
    c = np.broadcast_to(keep[:, None].astype(float), (mesh.nelements, len(W)))
    coldiv_c = sk.BilinearForm(lambda u, v, w: w.c * (div(u) * v))
    B_reg = sk.asm(coldiv_c, vp2, p1, c=c).toarray()
    assert B.shape == (p1.N, vp2.N)
    assert rel(B, B_reg) <= FP64_TOL
    mass = sk.BilinearForm(lambda u, v, w: dot(u, v))
    mass_c = sk.BilinearForm(lambda u, v, w: w.c * dot(u, v))
    M = sk.asm(mass, cut_p).toarray()
    M_reg = sk.asm(mass_c, p1, c=c).toarray()
    assert rel(M, M_reg) <= FP64_TOL
    # the cut pattern holds only the kept cells' DOF pairs
    kept_nodes = np.unique(p1.element_dofs[:, keep])
    assert set(np.flatnonzero(np.abs(M).sum(axis=1) > 0)) <= set(kept_nodes.tolist())


def test_functional_on_volume_facet_and_cut_bases():
    """asm(Functional, basis) integrates a scalar of the geometry: the
    reference's cut-basis moment test, plus the unit cube's volume, first
    moments and boundary area as closed forms."""
    import skfemntv as sk
    say("Functional on volume, facet and cut bases")
    mesh = _mesh(False, n=3, seed=4)
    basis = sk.Basis(mesh, sk.ElementVector(sk.ElementTetP1(), dim=1))
    cutb = _full_domain_cut(basis)
    moment = sk.Functional(lambda w: w.x[0] + 2. * w.x[1])
    got = sk.asm(moment, cutb)
    want = cutb.weights @ (cutb.points[:, 0] + 2. * cutb.points[:, 1])
    assert abs(got - want) <= 1e-13 * abs(want)
    assert abs(got - 1.5) <= 1e-13
    assert abs(sk.asm(moment, basis) - 1.5) <= 1e-13
    unit = sk.Functional(lambda w: 1. + 0. * w.x[0])
    assert abs(sk.asm(unit, basis) - 1.) <= 1e-13
    fb = sk.FacetBasis(mesh, sk.ElementVector(sk.ElementTetP1(), dim=1))
    assert abs(sk.asm(unit, fb) - 6.) <= 1e-12
    flux = sk.Functional(lambda w: w.x[0] * w.n[0] + w.x[1] * w.n[1] + w.x[2] * w.n[2])
    assert abs(sk.asm(flux, fb) - 3.) <= 1e-12


def test_cut_band_full_size_matches_oracle():
    """The benchmark's cut band at full size: the cells of the jittered 64^3
    P2 cube whose centroid lies within one cell width of the sphere
    |x - (0.5, 0.5, 0.5)| = 0.3 (55,571 cells), 12 random points each.  The
    square value + gradient form (pattern bit-exact) and the load vector
    against the oracle's array forms."""
    import skfemntv as sk
    from skfemntv.basis import tabulate
    say("cut band 64^3 P2, 55,571 cells x 12 points, bilinear + linear")
    n = 64
    ax = np.linspace(0., 1., n + 1)
    base = sk.MeshTet.init_tensor(ax, ax, ax)
    p = base.p.copy()
    interior = np.all((p > 1e-12) & (p < 1 - 1e-12), axis=0)
    p[:, interior] += np.random.default_rng(0).uniform(-0.2 / n, 0.2 / n, size=(3, int(interior.sum())))
    mesh = sk.MeshTet2.from_mesh(sk.MeshTet(p, base.t))
    c = mesh.p[:, mesh.t[:4]].mean(axis=1)
    band = np.abs(np.linalg.norm(c - 0.5, axis=0) - 0.3) < 1. / n
    E = int(band.sum())
    assert E == 55571
    rng = np.random.default_rng(3)
    Pc = 12
    off = np.arange(E + 1, dtype=np.int64) * Pc
    P = E * Pc
    X = rng.dirichlet(np.ones(4), P)[:, 1:]
    shape, ref_grad = tabulate(X.T, True)
    corners = mesh.p[:, mesh.t[:4][:, band]]
    Jinv = np.linalg.inv(np.moveaxis(corners[:, 1:4] - corners[:, [0]], -1, 0))
    grads = np.einsum("pnk,pki->pni", ref_grad, np.repeat(Jinv, Pc, axis=0))
    w = rng.uniform(0.1, 1., P) / (n ** 3 * 6 * Pc)
    dofs = mesh.t[:10, band].T[:, :, None].astype(np.int64)
    v, g = rng.uniform(0.5, 2., P), rng.uniform(0.5, 2., P)
    A = sk.CutBilinearFormAssembler(dofs, off, shape, grads, w)
    A.assemble(v, g)
    ip, ix, ref = cut.cut_bilinear_vectorized(dofs, off, shape, grads, w, v, g)
    np.testing.assert_array_equal(A.indptr, ip)
    np.testing.assert_array_equal(A.indices, ix)
    assert rel(A.values, ref) <= FP64_TOL
    f, G = rng.standard_normal((P, 1)), rng.standard_normal((P, 1, 3))
    L = sk.CutLinearFormAssembler(dofs, off, shape, grads, w)
    out, _ = L.assemble(f, G)
    assert rel(out, cut.cut_linear_vectorized(dofs, off, shape, grads, w, f, G)) <= FP64_TOL
