"""CPU tests of the host side: the C-ABI library exports what the header
declares, mesh/DOF numbering against the oracle, quadrature reduction,
coefficient layouts and the learned-FSAI parameter layout.  No GPU."""

import os
import re

import numpy as np
import pytest

import skfemntv as sk
from skfemntv import _native
from skfemntv._assembly import field_layout
from skfemntv.basis import reduced_tables, tabulate, tet_rule_4, tet_rule_keast11
from oracle import fsai as of
from oracle import host

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    text = open(os.path.join(ROOT, "include", "skfem_amd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return set(re.findall(r"\b(skfa_\w+)\s*\(", text))


def test_library_exports_every_header_symbol():
    import ctypes
    names = header_functions()
    assert len(names) >= 35
    lib = ctypes.CDLL(_native.LIB_PATH)
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing
    assert set(_native.SYMBOLS) == names
    L = _native.load_library()
    assert L.skfa_abi_version() == 1


def test_header_constants_match_host():
    text = open(os.path.join(ROOT, "include", "skfem_amd.h")).read()
    consts = dict(re.findall(r"#define (SKFA_\w+) (\d+)", text))
    assert int(consts["SKFA_FSAI_PARAMS"]) == _native.FSAI_PARAMS == of.NPARAMS
    assert int(consts["SKFA_FSAI_FEATURES"]) == _native.FSAI_FEATURES == of.NF
    assert int(consts["SKFA_FSAI_FRAG"]) == _native.FSAI_FRAG == 8 * 64 * 4 + 2 * 8 * 4 * 64 * 8
    assert int(consts["SKFA_FLAG_SYMMETRIC_TENSOR"]) == _native.FLAG_SYMMETRIC_TENSOR
    assert int(consts["SKFA_FLAG_VECTOR3_DIAGONAL"]) == _native.FLAG_VECTOR3_DIAGONAL
    from skfemntv._assembly import _CROSS_KIND, _CROSS_MODE
    assert int(consts["SKFA_CROSS_VALUE"]) == _CROSS_KIND["value"]
    assert int(consts["SKFA_CROSS_GRADIENT"]) == _CROSS_KIND["gradient"]
    for name in ("scalar", "row_divergence", "column_divergence", "tensor"):
        assert int(consts[f"SKFA_CROSS_{name.upper()}"]) == _CROSS_MODE[name]
    from skfemntv._assembly import _TUNE_KEYS
    for key in ("XCD_CHUNK", "LDS_ATOMIC", "PLACEMENT", "STREAMING", "ELEMENT_TABLE", "ROW_SPLIT", "MFMA"):
        assert int(consts[f"SKFA_TUNE_{key}"]) == getattr(_native, f"TUNE_{key}") == _TUNE_KEYS[key.lower()]


def test_device_calls_fail_loudly_without_gpu():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a device is present")
    _native._device_checked = False
    with pytest.raises(_native.NativeUnavailable):
        _native.lib()


@pytest.mark.parametrize("shape", [(2, 3, 4), (5, 1, 2), (3, 3, 3)])
def test_meshes_match_oracle_numbering(shape):
    axes = [np.linspace(0., 1., k + 1) ** 1.3 for k in shape]
    m = sk.MeshTet.init_tensor(*axes)
    p, t = host.kuhn_mesh(*axes)
    np.testing.assert_array_equal(m.p, p)
    np.testing.assert_array_equal(m.t, t)
    m2 = sk.MeshTet2.from_mesh(m)
    p2, t2 = host.p2_mesh(p, t)
    np.testing.assert_array_equal(m2.p, p2)
    np.testing.assert_array_equal(m2.t, t2)


@pytest.mark.parametrize("quadratic", [False, True])
def test_reduced_tables_reproduce_rule_sums(quadratic):
    for X, W in (tet_rule_4(), tet_rule_keast11()):
        tab, qg, qm = reduced_tables(X, W, quadratic)
        s, g = tabulate(X, quadratic)
        n = s.shape[1]
        wg = tab[:qg]
        dg = tab[qg:qg + qg * n * 3].reshape(qg, n, 3)
        o = qg + qg * n * 3
        wm = tab[o:o + qm]
        pm = tab[o + qm:o + qm + qm * n].reshape(qm, n)
        if quadratic:
            # barycentric coefficients of the gradient rule follow: they
            # rebuild the tabulated gradients from grad L_m
            o2 = o + qm + qm * n
            c = tab[o2:o2 + 16 * qg].reshape(16, qg)
            DL = np.array([[-1., -1., -1.], [1., 0., 0.], [0., 1., 0.], [0., 0., 1.]])
            pairs = [(k, k) for k in range(4)]
            for e, (i, j) in enumerate(((0, 1), (1, 2), (2, 0), (0, 3), (1, 3), (2, 3))):
                pairs += [(4 + e, i), (4 + e, j)]
            rec = np.zeros_like(dg)
            for p_, (b, m) in enumerate(pairs):
                rec[:, b] += c[p_][:, None] * DL[m]
            np.testing.assert_allclose(rec, dg, rtol=0., atol=1e-14)
            # then the per-pair factors of the gradient sums and the mass
            # sums: sum_i grad L_m(a,i) q_(ab,i)^T must equal the basis
            # rule's gradient sum, [6] its mass sum
            po = o2 + 16 * qg
            po += -po % 4   # 16-byte aligned in the FP32 copy
            assert not tab[o2 + 16 * qg:po].any()
            pt = tab[po:].reshape(n, n, 8)
            assert pt.shape == (10, 10, 8) and not pt[:, :, 7].any()
            Gb = np.einsum("q,qai,qbj->abij", W, g, g)
            terms = {k: [(k, 0)] for k in range(4)}
            for e, (i, j) in enumerate(((0, 1), (1, 2), (2, 0), (0, 3), (1, 3), (2, 3))):
                terms[4 + e] = [(i, 0), (j, 1)]
            for a in range(n):
                rebuilt = sum(np.einsum("i,bj->bij", DL[m], pt[a, :, 3 * k:3 * k + 3]) for m, k in terms[a])
                np.testing.assert_allclose(rebuilt, Gb[a], rtol=0., atol=1e-14)
                if a < 4:
                    assert not pt[a, :, 3:6].any()
            np.testing.assert_allclose(pt[:, :, 6], np.einsum("q,qa,qb->ab", W, s, s), rtol=1e-13, atol=1e-16)
        else:
            assert len(tab) == o + qm + qm * n
        G = np.einsum("q,qai,qbj->aibj", W, g, g)
        M = np.einsum("q,qa,qb->ab", W, s, s)
        np.testing.assert_allclose(np.einsum("q,qai,qbj->aibj", wg, dg, dg), G, rtol=1e-13, atol=1e-15)
        np.testing.assert_allclose(np.einsum("q,qa,qb->ab", wm, pm, pm), M, rtol=1e-13, atol=1e-16)
        if not quadratic:
            assert qg == 1   # P1 gradients are constant


def test_field_layout_detects_element_constant_fields():
    E, Q = 5, 4
    const = np.arange(E, dtype=float)[:, None] * np.ones((1, Q))
    arr, per_qp = field_layout(const, (E, Q), "value")
    assert not per_qp and arr.shape == (E,)
    varying = np.random.default_rng(0).random((E, Q))
    arr, per_qp = field_layout(varying, (E, Q), "value")
    assert per_qp and arr.shape == (E, Q)
    t = np.broadcast_to(np.eye(3), (E, Q, 3, 3))
    arr, per_qp = field_layout(t, (E, Q), "tensor", (3, 3))
    assert not per_qp and arr.shape == (E, 3, 3)
    with pytest.raises(ValueError):
        field_layout(np.ones((E + 1, Q)), (E, Q), "value")


def test_fsai_parameter_layout_matches_oracle():
    from skfemntv.fsai import init_params, split_params
    flat = init_params(0)
    assert flat.shape == (of.NPARAMS,)
    a = split_params(flat)
    b = of.unflatten(flat)
    for k in of.ORDER:
        np.testing.assert_array_equal(a[k], b[k])
    # zero output layer: the untrained network is exactly Jacobi
    assert not a["w3"].any() and not a["b3"].any()
    o, _ = of.mlp_forward(flat, np.random.default_rng(1).standard_normal((7, 16)))
    np.testing.assert_array_equal(o, 0.)


def test_batch_generator_is_seeded_and_bounded():
    from skfemntv.batch import batch_shapes, box_system
    s1 = batch_shapes(256, 16, 48, 5)
    assert s1.shape == (256, 3) and s1.min() >= 16 and s1.max() <= 48
    np.testing.assert_array_equal(s1, batch_shapes(256, 16, 48, 5))
    p, t, interior, A = box_system((3, 4, 2), np.random.default_rng([1, 2]), 0.2)
    p0, t0 = host.kuhn_mesh(np.linspace(0, 1, 4), np.linspace(0, 1, 5), np.linspace(0, 1, 3))
    np.testing.assert_array_equal(t, t0)
    np.testing.assert_array_equal(p[:, ~interior], p0[:, ~interior])
    h = np.array([1 / 3, 1 / 4, 1 / 2])[:, None]
    assert np.all(np.abs(p - p0) <= 0.2 * h + 1e-15)
    assert interior.sum() == 2 * 3 * 1
    w = np.linalg.eigvalsh(A)
    assert np.all(w > 0) and w.min() >= 1e-2 * (1 - 1e-9) and w.max() <= 1e2 * (1 + 1e-9)


def test_composite_relabelling_keeps_csr_order():
    """A field-local block relabelled by increasing row/column maps equals
    the dense block scattered into the composite numbering."""
    from scipy.sparse import random as sprandom
    from skfemntv._assembly import _composite_csr
    rng = np.random.default_rng(0)
    local = sprandom(7, 5, density=0.5, random_state=1, format="csr")
    local.sort_indices()
    N = 20
    row_map = np.sort(rng.choice(N, 7, replace=False))
    col_map = np.sort(rng.choice(N, 5, replace=False))
    K = _composite_csr(local.indptr, local.indices, local.data, row_map, col_map, N)
    dense = np.zeros((N, N))
    dense[np.ix_(row_map, col_map)] = local.toarray()
    np.testing.assert_array_equal(K.toarray(), dense)
    assert K.has_sorted_indices and K.shape == (N, N)
