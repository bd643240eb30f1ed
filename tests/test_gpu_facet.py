"""GPU parity of the facet (boundary) path: FacetBasis geometry against the
oracle's per-facet restatement of the reference, and the facet load vector
from skfa_assemble_facet_linear against the oracle's LinearFormAssembler
restatement on the same facet tabulation.  Each case prints its name first.
"""

import sys

import numpy as np
import pytest

import oracle
from oracle import host

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12


def say(msg):
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()


def p2_kuhn(n, jitter, seed):
    ax = np.linspace(0., 1., n + 1)
    p, t = host.kuhn_mesh(ax, ax, ax)
    if jitter:
        p = host.jitter_interior(p, 1. / n, jitter, seed)
    return p, t, host.p2_mesh(p, t)


@pytest.mark.parametrize("quadratic,components", [(False, 1), (True, 1), (True, 3)])
def test_facet_basis_and_load_match_oracle(quadratic, components):
    import skfemntv as sk
    from skfemntv.helpers import dot
    say(f"facet basis quadratic={quadratic} C={components}")
    p, t, (p2, t2) = p2_kuhn(3, 0.2, seed=7)
    mesh = sk.MeshTet2(p2, t2) if quadratic else sk.MeshTet(p, t)
    elem = sk.ElementVector(sk.ElementTetP2() if quadratic else sk.ElementTetP1(), dim=components)
    fb = sk.FacetBasis(mesh, elem)
    ref = host.facet_basis(p2 if quadratic else p, t2 if quadratic else t, quadratic, components)
    np.testing.assert_array_equal(fb.facet_ids, ref["facets"])
    np.testing.assert_array_equal(fb.parent_elements, ref["parents"])
    F = fb.facet_ids.size
    np.testing.assert_array_equal(fb.element_dofs.T.reshape(F, -1, components), ref["dofs"])
    for name, mine, theirs in (("dx", fb.dx, ref["dx"]), ("x", fb.global_coordinates, ref["x"]),
                               ("normals", fb.normals, ref["normals"]), ("shape", fb.tabulated_shape, ref["shape"])):
        np.testing.assert_allclose(mine, theirs, rtol=0, atol=1e-14, err_msg=name)
    Q = fb.dx.shape[1]
    n = ref["shape"].shape[2]
    grads = np.zeros((F, Q, n, 3))
    say(f"facet load constant quadratic={quadratic} C={components}")
    f = np.arange(1., components + 1.)
    b = sk.asm(sk.LinearForm(lambda v, w: dot(w.f, v)), fb, f=f)
    expect = oracle.assemble_linear(ref["shape"], grads, ref["dx"], ref["dofs"], ref["N"], components, value=f)
    assert b.shape == (fb.N,)
    err = np.linalg.norm(b - expect) / np.linalg.norm(expect)
    assert err <= FP64_TOL, err
    if components == 1:
        say(f"facet load x.n quadratic={quadratic}")
        # g = x . n on the boundary: sum of the load = int x.n ds = 3 |Omega|
        g = sk.LinearForm(lambda v, w: dot(w.x[0] * w.n[0] + w.x[1] * w.n[1] + w.x[2] * w.n[2], v))
        b = sk.asm(g, fb)
        gv = np.einsum("fqi,fqi->fq", ref["x"], ref["normals"])[..., None]
        expect = oracle.assemble_linear(ref["shape"], grads, ref["dx"], ref["dofs"], ref["N"], 1, value=gv)
        err = np.linalg.norm(b - expect) / np.linalg.norm(expect)
        assert err <= FP64_TOL, err
        assert abs(b.sum() - 3.) <= 1e-12


@pytest.mark.parametrize("quadratic,components", [(False, 1), (True, 1), (True, 3)])
def test_facet_mass_matches_oracle(quadratic, components):
    """Boundary mass k u.v on the facet DOFs: the reference's facet pattern
    (every DOF pair of each facet's parent element) bit-exact, values
    against the oracle's BilinearFormAssembler restatement; 1^T M 1 = C x
    boundary area for k = 1."""
    import skfemntv as sk
    from skfemntv.helpers import dot
    say(f"facet mass quadratic={quadratic} C={components}")
    p, t, (p2, t2) = p2_kuhn(3, 0.2, seed=5)
    mesh = sk.MeshTet2(p2, t2) if quadratic else sk.MeshTet(p, t)
    elem = sk.ElementVector(sk.ElementTetP2() if quadratic else sk.ElementTetP1(), dim=components)
    fb = sk.FacetBasis(mesh, elem)
    ref = host.facet_basis(p2 if quadratic else p, t2 if quadratic else t, quadratic, components)
    F, Q, n = ref["shape"].shape
    k = np.random.default_rng(components).uniform(0.5, 2., (F, 1))
    M = sk.asm(sk.BilinearForm(lambda u, v, w: w.k * dot(u, v)), fb, k=k)
    rows = ref["dofs"].reshape(F, -1)
    ip, ix, sc = oracle.pattern(rows, nrows=ref["N"])
    vals = oracle.assemble_bilinear(ref["shape"], np.zeros((F, Q, n, 3)), ref["dx"], sc, len(ix), components,
                                    value=np.ascontiguousarray(np.broadcast_to(k, (F, Q))))
    M = M.tocsr()
    assert M.shape == (fb.N, fb.N)
    np.testing.assert_array_equal(np.asarray(M.indptr[:len(ip)], dtype=np.int64), ip)
    np.testing.assert_array_equal(np.asarray(M.indices, dtype=np.int64), ix)
    err = np.linalg.norm(M.data - vals) / np.linalg.norm(vals)
    assert err <= FP64_TOL, err
    M1 = sk.asm(sk.BilinearForm(lambda u, v, w: dot(u, v)), fb)
    one = np.ones(fb.N)
    assert abs(one @ (M1 @ one) - 6. * components) <= 1e-12
    if components == 1:
        with pytest.raises(ValueError, match="3-component"):
            sk.NativeFacetBilinearForm(fb).assemble(symmetric_gradient=1.)
    else:
        say("facet symmetric-gradient + divergence + gradient + value kinds")
        rng = np.random.default_rng(9)
        coefs = {k: rng.uniform(0.5, 2., (F, Q)) for k in ("value", "gradient", "symgrad", "divergence")}
        M = sk.NativeFacetBilinearForm(fb).assemble(value=coefs["value"], gradient=coefs["gradient"],
                                                    symmetric_gradient=coefs["symgrad"],
                                                    divergence=coefs["divergence"])
        vals = oracle.assemble_bilinear(ref["shape"], np.ascontiguousarray(fb.tabulated_gradients), ref["dx"], sc,
                                        len(ix), components, **coefs)
        np.testing.assert_array_equal(np.asarray(M.indices, dtype=np.int64), ix)
        err = np.linalg.norm(M.data - vals) / np.linalg.norm(vals)
        assert err <= FP64_TOL, err


@pytest.mark.parametrize("quadratic,components", [(False, 1), (True, 1), (True, 3)])
def test_facet_gradient_and_normal_gradient_terms(quadratic, components):
    """Gradient terms on a FacetBasis run the ragged-quadrature kernels with
    the facets as cells of Q points: the physical gradients reproduce an
    affine field exactly; grad:grad + mass and G:grad(v) match the oracle's
    restatement of the reference's cut assemblers on the same tables; the
    normal-gradient kinds satisfy K^T x = int n_x psi and K x = int phi n_x
    for the coordinate field x."""
    import skfemntv as sk
    from skfemntv.helpers import ddot, dot, grad
    from oracle import cut
    say(f"facet gradient terms quadratic={quadratic} C={components}")
    p, t, (p2, t2) = p2_kuhn(3, 0.2, seed=11)
    mesh = sk.MeshTet2(p2, t2) if quadratic else sk.MeshTet(p, t)
    elem = sk.ElementVector(sk.ElementTetP2() if quadratic else sk.ElementTetP1(), dim=components)
    fb = sk.FacetBasis(mesh, elem)
    F, Q = fb.dx.shape
    r = fb.ragged()
    dofs = r["dofs"]
    # affine field u = 2 + 3x - y + z/2 in component 0: its discrete gradient is exact
    u = 2. + 3. * fb.doflocs[0] - fb.doflocs[1] + 0.5 * fb.doflocs[2]
    gh = np.einsum("fqnk,fn->fqk", fb.tabulated_gradients, u[dofs[:, :, 0]])
    assert np.abs(gh - np.array([3., -1., 0.5])).max() < 1e-12
    K = sk.asm(sk.BilinearForm(lambda u, v, w: ddot(grad(u), grad(v)) + 2. * dot(u, v)), fb)
    P = F * Q
    ip, ix, ref = cut.cut_bilinear(dofs, r["offsets"], r["shape"], r["gradients"], r["weights"],
                                   value=np.full(P, 2.), gradient=np.ones(P))
    n0 = len(ip) - 1
    dense = np.zeros((n0, n0))
    for row in range(n0):
        dense[row, ix[ip[row]:ip[row + 1]]] = ref[ip[row]:ip[row + 1]]
    Kd = K.toarray()
    assert K.shape == (fb.N, fb.N) and not Kd[n0:].any()
    assert np.linalg.norm(Kd[:n0, :n0] - dense) <= FP64_TOL * np.linalg.norm(dense)
    rng = np.random.default_rng(components)
    G = rng.standard_normal((F, Q, components, 3))
    out, diag = sk.NativeFacetLinearForm(fb).assemble(gradient=G)
    want = cut.cut_linear(dofs, r["offsets"], r["shape"], r["gradients"], r["weights"],
                          gradient=G.reshape(P, components, 3))
    assert out.shape == (fb.N,) and diag.entity_count == F
    assert np.linalg.norm(out[:len(want)] - want) <= FP64_TOL * np.linalg.norm(want)
    if components == 1:
        X = sk.NativeCrossBilinearForm(fb, fb)
        x = fb.doflocs[0]
        load_nx = cut.cut_linear(dofs, r["offsets"], r["shape"], r["gradients"], r["weights"],
                                 value=r["normals"][:, :1])
        Kr = X.assemble_kinds("normal_gradient", "value", 1.)
        lhs = (Kr.T @ x)[:len(load_nx)]
        assert np.linalg.norm(lhs - load_nx) <= 1e-12 * np.linalg.norm(load_nx)
        Kc = X.assemble_kinds("value", "normal_gradient", 1.)
        lhs = (Kc @ x)[:len(load_nx)]
        assert np.linalg.norm(lhs - load_nx) <= 1e-12 * np.linalg.norm(load_nx)
        # gradient x gradient with a normal on one side: (n . grad u)(sum_j d_j v)
        Kg = X.assemble_kinds("normal_gradient", "normal_gradient", 1.)
        nn = X.assemble_kinds("gradient", "gradient", 1.)
        assert Kg.shape == nn.shape == (fb.N, fb.N)
        assert abs((Kg @ x) @ x - cut.cut_linear(dofs, r["offsets"], r["shape"], r["gradients"], r["weights"],
                                                 value=r["normals"][:, :1] ** 2).sum()) < 1e-12


def _permuted_mesh(n, seed, quadratic):
    """Jittered Kuhn cube whose elements list their vertices in a random even
    permutation (orientation kept), so facet corner orders differ from the
    parents' local face orders."""
    ax = np.linspace(0., 1., n + 1)
    p, t = host.kuhn_mesh(ax, ax, ax)
    p = host.jitter_interior(p, 1. / n, 0.2, seed)
    rng = np.random.default_rng(seed)
    even = [pm for pm in __import__("itertools").permutations(range(4))
            if sum(pm[i] > pm[j] for i in range(4) for j in range(i + 1, 4)) % 2 == 0]
    choice = rng.integers(0, len(even), t.shape[1])
    t = np.stack([t[list(even[c]), e] for e, c in enumerate(choice)], axis=1)
    if quadratic:
        return host.p2_mesh(p, t)
    return p, t


@pytest.mark.parametrize("quadratic", [False, True])
def test_facet_basis_on_permuted_connectivity(quadratic):
    """Facet rule points follow the facet's own corner order
    (basis.py:2634-2645), not the parent's local face order."""
    import skfemntv as sk
    from skfemntv.helpers import dot
    say(f"facet basis on permuted connectivity quadratic={quadratic}")
    p, t = _permuted_mesh(3, 13, quadratic)
    mesh = sk.MeshTet2(p, t) if quadratic else sk.MeshTet(p, t)
    elem = sk.ElementVector(sk.ElementTetP2() if quadratic else sk.ElementTetP1(), dim=1)
    fb = sk.FacetBasis(mesh, elem)
    ref = host.facet_basis(p, t, quadratic, 1)
    assert fb._face_shape.shape[0] > 4
    for name, mine, theirs in (("dx", fb.dx, ref["dx"]), ("x", fb.global_coordinates, ref["x"]),
                               ("normals", fb.normals, ref["normals"]), ("shape", fb.tabulated_shape, ref["shape"])):
        np.testing.assert_allclose(mine, theirs, rtol=0, atol=1e-14, err_msg=name)
    F, Q = fb.dx.shape
    n = ref["shape"].shape[2]
    b = sk.asm(sk.LinearForm(lambda v, w: dot(1. + w.x[0], v)), fb)
    fx = (1. + ref["x"][..., 0])[..., None]
    expect = oracle.assemble_linear(ref["shape"], np.zeros((F, Q, n, 3)), ref["dx"], ref["dofs"], ref["N"], 1,
                                    value=np.ascontiguousarray(fx))
    assert np.linalg.norm(b - expect) <= FP64_TOL * np.linalg.norm(expect)


@pytest.mark.parametrize("quadratic", [False, True])
def test_interior_facets_share_points_and_jump_penalty(quadratic):
    """InteriorFacetBasis sides share points and normals (the reference's
    test_3d_interior_sides_share_points_normals_and_continuous_trace), and
    asm(form, [sides], [sides]) sums the (trial side, test side) cross blocks
    with w.idx signs: against the oracle's restatement, symmetric, and zero
    on continuous fields (their jumps vanish)."""
    import skfemntv as sk
    from skfemntv.helpers import dot, jump
    from oracle import cut
    say(f"interior facets quadratic={quadratic}")
    p, t = _permuted_mesh(2, 17, quadratic)
    mesh = sk.MeshTet2(p, t) if quadratic else sk.MeshTet(p, t)
    elem = sk.ElementVector(sk.ElementTetP2() if quadratic else sk.ElementTetP1(), dim=1)
    bases = [sk.InteriorFacetBasis(mesh, elem, side=s) for s in (0, 1)]
    refs = [host.facet_basis(p, t, quadratic, 1, side=s, interior=True) for s in (0, 1)]
    for b, r in zip(bases, refs):
        np.testing.assert_array_equal(b.parent_elements, r["parents"])
        for name in ("dx", "x", "normals", "shape"):
            mine = {"dx": b.dx, "x": b.global_coordinates, "normals": b.normals, "shape": b.tabulated_shape}[name]
            np.testing.assert_allclose(mine, r[name], rtol=0, atol=1e-14, err_msg=name)
    np.testing.assert_allclose(bases[0].global_coordinates, bases[1].global_coordinates, rtol=0, atol=3e-14)
    np.testing.assert_allclose(bases[0].normals, bases[1].normals, rtol=0, atol=3e-14)
    u = bases[0].doflocs[0] + 2. * bases[0].doflocs[1] - bases[0].doflocs[2]
    r0, r1 = bases[0].ragged(), bases[1].ragged()
    tr = [np.einsum("pn,pn->p", r["shape"], u[r["dofs"][:, :, 0].repeat(len(bases[0].W), axis=0)])
          for r in (r0, r1)]
    assert np.abs(tr[0] - tr[1]).max() < 1e-13

    # the four (trial side, test side) blocks of (1 + z) u v, summed
    A = sk.asm(sk.BilinearForm(lambda u, v, w: (1. + w.x[2]) * dot(u, v)), bases, bases).toarray()
    coef = (1. + bases[0].global_coordinates[..., 2]).ravel()
    N = bases[0].N
    want = np.zeros((N, N))
    rag = [r0, r1]
    for i in (0, 1):          # trial side
        for j in (0, 1):      # test side
            ip, ix, vals = cut.cut_cross(rag[j]["dofs"], rag[i]["dofs"], r0["offsets"], rag[j]["shape"],
                                         rag[i]["shape"], r0["weights"], rag[j]["gradients"], rag[i]["gradients"],
                                         coef, "value", "value")
            for row in range(len(ip) - 1):
                np.add.at(want[row], ix[ip[row]:ip[row + 1]], vals[ip[row]:ip[row + 1]])
    assert np.linalg.norm(A - want) <= FP64_TOL * np.linalg.norm(want)
    assert np.abs(A - A.T).max() <= 1e-14 * np.abs(A).max()

    # with w.idx signs the blocks form a jump penalty: continuous basis
    # functions have no jump across interior facets, so it vanishes
    def penalty(u, v, w):
        ju, jv = jump(w, u, v)
        return (1. + w.x[2]) * dot(ju, jv)
    J = sk.asm(sk.BilinearForm(penalty), bases, bases).toarray()
    assert np.abs(J).max() <= 1e-13 * np.abs(A).max()


def test_facet_elasticity_kinds_through_asm_annihilate_rigid_modes():
    """asm of 2 eps(u):eps(v) + div u div v on a vector-P2 FacetBasis (the
    tracer's symmetric-gradient and divergence kinds): rigid motions have
    zero strain and divergence, so K r = 0 for the six rigid modes, while K
    itself is nonzero and symmetric."""
    import skfemntv as sk
    from skfemntv.helpers import ddot, div, sym_grad
    say("facet elasticity kinds through asm")
    p, t, (p2, t2) = p2_kuhn(2, 0.2, seed=21)
    fb = sk.FacetBasis(sk.MeshTet2(p2, t2), sk.ElementVector(sk.ElementTetP2(), dim=3))
    K = sk.asm(sk.BilinearForm(lambda u, v, w: 2. * ddot(sym_grad(u), sym_grad(v)) + div(u) * div(v)), fb)
    K = K.toarray()
    assert np.abs(K).max() > 0. and np.abs(K - K.T).max() <= 1e-13 * np.abs(K).max()
    x = fb.doflocs
    comp = np.arange(fb.N) % 3
    modes = [(comp == k).astype(float) for k in range(3)]
    for a, b in ((0, 1), (1, 2), (2, 0)):
        r = np.zeros(fb.N)
        r[comp == a] = -x[b, comp == a]
        r[comp == b] = x[a, comp == b]
        modes.append(r)
    for r in modes:
        assert np.abs(K @ r).max() <= 1e-12 * np.abs(K).max() * max(1., np.abs(r).max())
