"""Cut-cell (ragged quadrature) assembly on the CPU: the oracle's restatement
pinned against the reference's own compiled CutLinearFormAssembler,
CutBilinearFormAssembler and CutCrossAssembler (oracle/_ref) on the same
inputs and against closed forms (partition of unity, zero row sums of
stiffness), plus the host-side validation of the GPU assemblers, which
raises before any device call.
"""

import numpy as np
import pytest

import oracle
from oracle import cut

TOL = 1e-13


@pytest.fixture(scope="module")
def ref():
    m = oracle.ref_module()
    if m is None:
        pytest.skip("oracle/_ref (reference C++ build) not present")
    return m


def _close(mine, theirs):
    theirs = np.asarray(theirs)
    scale = max(np.abs(theirs).max(), 1.)
    assert np.abs(np.asarray(mine) - theirs).max() <= TOL * scale


@pytest.mark.parametrize("quadratic,components", [(False, 1), (True, 1), (False, 3), (True, 3)])
def test_cut_bilinear_oracle_matches_reference_cpp(ref, quadratic, components):
    dofs, off, sh, gr, w = cut.random_case(quadratic, components, seed=components + 2 * quadratic)
    rng = np.random.default_rng(5)
    v, g = rng.uniform(0.5, 2., len(w)), rng.uniform(0.5, 2., len(w))
    A = ref.CutBilinearFormAssembler(dofs, off, sh, gr, w)
    for value, gradient in ((v, None), (None, g), (v, g)):
        A.assemble(value, gradient, 1)
        ip, ix, vals = cut.cut_bilinear(dofs, off, sh, gr, w, value, gradient)
        np.testing.assert_array_equal(ip, A.indptr)
        np.testing.assert_array_equal(ix, A.indices)
        _close(vals, A.values)


@pytest.mark.parametrize("components", [1, 3])
def test_cut_linear_oracle_matches_reference_cpp(ref, components):
    dofs, off, sh, gr, w = cut.random_case(True, components, seed=11)
    rng = np.random.default_rng(6)
    P = len(w)
    A = ref.CutLinearFormAssembler(dofs, off, sh, gr, w)
    cases = [(rng.standard_normal(components), None), (rng.standard_normal((P, components)), None),
             (None, rng.standard_normal((components, 3))), (None, rng.standard_normal((P, components, 3))),
             (rng.standard_normal((P, components)), rng.standard_normal((P, components, 3)))]
    for value, gradient in cases:
        theirs, _ = A.assemble(value, gradient, 1)
        _close(cut.cut_linear(dofs, off, sh, gr, w, value, gradient), theirs)


@pytest.mark.parametrize("rk,ck,Cr,Cc,tensor", [
    ("value", "value", 1, 1, False), ("gradient", "gradient", 3, 3, False), ("value", "gradient", 1, 3, True),
    ("gradient", "value", 3, 1, True), ("gradient", "gradient", 1, 3, True)])
def test_cut_cross_oracle_matches_reference_cpp(ref, rk, ck, Cr, Cc, tensor):
    rd, off, rs, rgr, w = cut.random_case(False, Cr, seed=3)
    cd, off2, cs, cgr, _ = cut.random_case(True, Cc, seed=3)
    assert np.array_equal(off, off2)
    rng = np.random.default_rng(8)
    P = len(w)
    if tensor:
        shape = (P, Cr) + ((3,) if rk == "gradient" else ()) + (Cc,) + ((3,) if ck == "gradient" else ())
        coef = rng.standard_normal(shape)
    else:
        coef = rng.uniform(0.5, 2., P)
    A = ref.CutCrossAssembler(rd, cd, off, rs, cs, w, rgr, cgr)
    A.assemble(coef, rk, ck, 1)
    ip, ix, vals = cut.cut_cross(rd, cd, off, rs, cs, w, rgr, cgr, coef, rk, ck)
    np.testing.assert_array_equal(ip, A.indptr)
    np.testing.assert_array_equal(ix, A.indices)
    _close(vals, A.values)
    assert A.rows == len(ip) - 1


def _p1_cut_case(seed):
    """Real P1 tabulation: barycentric shape values at random points of each
    cut cell and the cell's constant physical gradients."""
    from oracle import host
    from oracle.pipeline import dofs_of
    ax = np.linspace(0., 1., 3)
    p, t = host.kuhn_mesh(ax, ax, ax)
    _, dofs = dofs_of(t, p.shape[1], 4, 1)
    rng = np.random.default_rng(seed)
    E = t.shape[1]
    counts = rng.integers(0, 5, E)
    off = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
    sh, gr = [], []
    for e in range(E):
        J = p[:, t[1:4, e]] - p[:, [t[0, e]]]
        G = np.vstack([-np.ones(3), np.eye(3)]) @ np.linalg.inv(J)
        for _ in range(counts[e]):
            x = rng.dirichlet(np.ones(4))
            sh.append(x)
            gr.append(G)
    return dofs, off, np.array(sh), np.array(gr), rng.uniform(0.1, 1., off[-1])


def test_cut_oracle_closed_forms():
    dofs, off, sh, gr, w = _p1_cut_case(2)
    # partition of unity: the mass matrix sums to the total weight and the
    # unit load vector does too; grad of a constant vanishes, so stiffness
    # rows sum to zero
    _, _, m = cut.cut_bilinear(dofs, off, sh, gr, w, value=np.ones(len(w)))
    assert abs(m.sum() - w.sum()) < 1e-13 * w.sum()
    ip, ix, k = cut.cut_bilinear(dofs, off, sh, gr, w, gradient=np.ones(len(w)))
    rows = np.add.reduceat(k, ip[:-1][np.diff(ip) > 0])
    assert np.abs(rows).max() < 1e-12 * np.abs(k).max()
    f = cut.cut_linear(dofs, off, sh, gr, w, value=np.ones(1))
    assert abs(f.sum() - w.sum()) < 1e-13 * w.sum()
    # a linear field u = x: K u = int grad phi . grad u and the load of G = e_x agree
    from oracle import host
    ax = np.linspace(0., 1., 3)
    p, _ = host.kuhn_mesh(ax, ax, ax)
    Ku = np.zeros(len(ip) - 1)
    for r in range(len(ip) - 1):
        Ku[r] = k[ip[r]:ip[r + 1]] @ p[0, ix[ip[r]:ip[r + 1]]]
    fx = cut.cut_linear(dofs, off, sh, gr, w, gradient=np.array([[1., 0., 0.]]))
    np.testing.assert_allclose(Ku, fx[:len(Ku)], atol=1e-13)


def test_cut_host_validation_messages():
    import skfemntv as sk
    dofs, off, sh, gr, w = cut.random_case(False, 1, seed=0)
    with pytest.raises(ValueError, match="cell dofs must have shape"):
        sk.CutBilinearFormAssembler(dofs[:, :, 0], off, sh, gr, w)
    with pytest.raises(ValueError, match="cell offsets must have shape"):
        sk.CutLinearFormAssembler(dofs, off[:-1], sh, gr, w)
    with pytest.raises(ValueError, match="cut shape values"):
        sk.CutBilinearFormAssembler(dofs, off, sh[:, :3], gr, w)
    with pytest.raises(ValueError, match="cut shape gradients"):
        sk.CutBilinearFormAssembler(dofs, off, sh, gr[:-1], w)
    with pytest.raises(ValueError, match="cut weights"):
        sk.CutBilinearFormAssembler(dofs, off, sh, gr, w[:-1])
    bad = off.copy()
    bad[-1] += 1
    with pytest.raises(ValueError, match="do not span"):
        sk.CutBilinearFormAssembler(dofs, bad, sh, gr, w)
    bad = off.copy()
    bad[1], bad[2] = 3, 1
    with pytest.raises(ValueError, match="nondecreasing"):
        sk.CutBilinearFormAssembler(dofs, bad, sh, gr, w)
    neg = dofs.copy()
    neg[0, 0, 0] = -1
    with pytest.raises(ValueError, match="negative dof"):
        sk.CutBilinearFormAssembler(neg, off, sh, gr, w)
    d3, o3, s3, g3, w3 = cut.random_case(False, 3, seed=0)
    swapped = d3[:, :, ::-1].copy()
    with pytest.raises(ValueError, match="consecutively"):
        sk.CutBilinearFormAssembler(swapped, o3, s3, g3, w3)
    with pytest.raises(ValueError, match="cut cross cell dofs"):
        sk.CutCrossAssembler(dofs, dofs[:-1], off, sh, sh, w, gr, gr)
    with pytest.raises(ValueError, match="cut cross offsets do not span"):
        sk.CutCrossAssembler(dofs, dofs, off - 1, sh, sh, w, gr, gr)


def test_functional_validation_on_host():
    import skfemntv as sk
    from skfemntv._assembly import integrate_functional
    with pytest.raises(ValueError, match="matching"):
        integrate_functional(np.ones((2, 3)), np.ones((3, 2)))
    assert integrate_functional(np.zeros((0, 4)), np.zeros((0, 4))) == 0.
    with pytest.raises(sk.UnsupportedNativeForm, match="one Basis"):
        sk.asm(sk.Functional(lambda w: 1.))


@pytest.mark.parametrize("quadratic,components,dim", [(False, 1, 3), (True, 1, 3), (True, 3, 3), (False, 1, 2)])
def test_cut_vectorized_oracle_matches_loops(quadratic, components, dim):
    """The array forms used for full-size GPU checks give the loop forms'
    results (cells without points, 1..12 points per cell)."""
    dofs, off, sh, gr, w = cut.random_case(quadratic, components, seed=70 + components, dim=dim, max_points=12)
    rng = np.random.default_rng(7)
    P = len(w)
    v, g = rng.uniform(0.5, 2., P), rng.uniform(0.5, 2., P)
    for value, gradient in ((v, None), (None, g), (v, g)):
        ip, ix, ref = cut.cut_bilinear(dofs, off, sh, gr, w, value, gradient)
        ip2, ix2, mine = cut.cut_bilinear_vectorized(dofs, off, sh, gr, w, value, gradient)
        np.testing.assert_array_equal(ip2, ip)
        np.testing.assert_array_equal(ix2, ix)
        _close(mine, ref)
    cases = [(rng.standard_normal(components), None), (None, rng.standard_normal((P, components, dim))),
             (rng.standard_normal((P, components)), rng.standard_normal((components, dim)))]
    for value, gradient in cases:
        _close(cut.cut_linear_vectorized(dofs, off, sh, gr, w, value, gradient),
               cut.cut_linear(dofs, off, sh, gr, w, value, gradient))
