"""GPU parity of the batched assembly, condensation, PCG and learned-FSAI
kernels against the CPU oracle (oracle/fsai.py, oracle/fem_oracle.c)."""

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from oracle import fsai as of
from oracle import host, pipeline
import oracle

pytestmark = pytest.mark.gpu

SHAPES = [(3, 4, 5), (5, 3, 4), (4, 4, 4)]


def oracle_system(shape, seed, s):
    """System s of a batch, assembled and condensed by the oracle."""
    from skfemntv.batch import box_system
    p, t, interior, A = box_system(shape, np.random.default_rng([seed, s]))
    E = t.shape[1]
    X, W = host.tet_rule_4()
    (shape_, grads, dx, x, det, tol, cond), _ = pipeline.basis_geometry(p, t, X, W, False)
    N, dofs = pipeline.dofs_of(t, p.shape[1], 4, 1)
    ip, ix, sc = oracle.pattern(dofs.reshape(E, -1), nrows=N)
    A_eq = np.ascontiguousarray(np.broadcast_to(A[:, None], (E, 4, 3, 3)))
    vals = oracle.assemble_tensor(grads, dx, A_eq, sc, len(ix))
    K = sp.csr_matrix((vals, ix, ip), shape=(N, N))
    b = np.zeros(N)
    np.add.at(b, t.T.ravel(), np.repeat(dx.sum(1) / 4, 4))
    Af, bf, free = of.condense(K, b, interior)
    return Af, bf, p[:, free].T.copy()


@pytest.fixture(scope="module")
def batch():
    from skfemntv.batch import AnisotropicBatch
    b = AnisotropicBatch(SHAPES, seed=11)
    b.assemble()
    return b


@pytest.fixture(scope="module")
def oracle_blocks():
    return [oracle_system(s, 11, i) for i, s in enumerate(SHAPES)]


def rel_frob(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def test_batch_assembly_matches_oracle(batch, oracle_blocks):
    print("case: batch assembly of", SHAPES, flush=True)
    for s, (Ao, bo, xyz) in enumerate(oracle_blocks):
        A, b = batch.system_csr(s)
        A.sort_indices()
        Ao = Ao.tocsr()
        Ao.sort_indices()
        np.testing.assert_array_equal(A.indptr, Ao.indptr)
        np.testing.assert_array_equal(A.indices, Ao.indices)
        assert rel_frob(A.data, Ao.data) <= 1e-12
        assert rel_frob(b, bo) <= 1e-13
        r0, r1 = batch.sys_ptr_host[s], batch.sys_ptr_host[s + 1]
        np.testing.assert_array_equal(batch.xyz[r0:r1].cpu().numpy(), xyz)


def test_batch_graph_replay_is_identical(batch):
    import torch
    print("case: hip graph replay of batched assembly", flush=True)
    A, b = batch.assemble()
    ref_a, ref_b = A.data.clone(), b.clone()
    for k in range(3):
        A.data.zero_()
        b.zero_()
        torch.cuda.synchronize()
        batch.assemble(graph=True)
        torch.cuda.synchronize()
        assert torch.equal(A.data, ref_a), k
        assert torch.equal(b, ref_b), k
        _ = (ref_a * 2).sum().item()   # eager work between replays


def test_device_csr_ops_match_scipy(oracle_blocks):
    import torch
    from skfemntv.solver import DeviceCSR
    print("case: condense / lower / transpose / spmv", flush=True)
    Ao = oracle_blocks[0][0].tocsr()
    A = DeviceCSR.from_scipy(Ao)
    rng = np.random.default_rng(0)
    x = rng.standard_normal(Ao.shape[0])
    y = A.matvec(torch.from_numpy(x).cuda()).cpu().numpy()
    assert rel_frob(y, Ao @ x) <= 1e-14
    keep = rng.random(Ao.shape[0]) > 0.3
    Ac, newid = A.condense(torch.from_numpy(keep.astype(np.uint8)).cuda())
    Ar, _, _ = of.condense(Ao, np.zeros(Ao.shape[0]), keep)
    Ar.sort_indices()
    Ag = Ac.to_scipy()
    np.testing.assert_array_equal(Ag.indptr, Ar.indptr)
    np.testing.assert_array_equal(Ag.indices, Ar.indices)
    np.testing.assert_array_equal(Ag.data, Ar.data)
    np.testing.assert_array_equal(newid.cpu().numpy()[:-1][keep], np.arange(keep.sum()))
    pat = A.lower()
    lptr, lidx = of.lower_pattern(Ao)
    np.testing.assert_array_equal(pat.lptr.cpu().numpy(), lptr)
    np.testing.assert_array_equal(pat.lidx.cpu().numpy(), lidx)
    L = sp.tril(Ao, format="csr")
    np.testing.assert_array_equal(Ao.data[pat.lsrc.cpu().numpy()], L.data)
    tptr, tidx, tsrc = pat.transpose()
    U = sp.csr_matrix(L.T)
    U.sort_indices()
    np.testing.assert_array_equal(tptr.cpu().numpy(), U.indptr)
    np.testing.assert_array_equal(tidx.cpu().numpy(), U.indices)
    np.testing.assert_array_equal(L.data[tsrc.cpu().numpy()], U.data)


def _block_diag(blocks):
    A = sp.block_diag([b[0] for b in blocks]).tocsr()
    A.sort_indices()
    b = np.concatenate([b[1] for b in blocks])
    sizes = [b[0].shape[0] for b in blocks]
    return A, b, np.concatenate(([0], np.cumsum(sizes))).astype(np.int64)


def test_pcg_jacobi_matches_oracle(batch, oracle_blocks):
    from skfemntv.solver import JacobiPreconditioner, pcg, relative_residuals
    print("case: batched Jacobi PCG", flush=True)
    A, b = batch.assemble()
    res = pcg(A, b, batch.sys_ptr, JacobiPreconditioner(A), tol=1e-8)
    x = res.x.cpu().numpy()
    for s, (Ao, bo, _) in enumerate(oracle_blocks):
        d = Ao.diagonal()
        _, it = of.pcg(Ao, bo, lambda r: r / d, tol=1e-8)
        assert abs(int(res.iterations[s]) - it) <= 2, (s, res.iterations[s], it)
        r0, r1 = batch.sys_ptr_host[s], batch.sys_ptr_host[s + 1]
        exact = spla.spsolve(Ao.tocsc(), bo)
        assert rel_frob(x[r0:r1], exact) <= 1e-6
    rr = relative_residuals(A, b, res.x, batch.sys_ptr)
    assert np.all(rr <= 2e-8), rr


def test_pcg_factor_with_exact_fsai_matches_oracle(oracle_blocks):
    import torch
    from skfemntv.solver import DeviceCSR, FactorPreconditioner, pcg
    print("case: FSAI-preconditioned PCG with the classical factor", flush=True)
    A, b, sys_ptr = _block_diag(oracle_blocks)
    G = of.exact_fsai(A)
    G.sort_indices()
    Ad = DeviceCSR.from_scipy(A)
    pat = Ad.lower()
    np.testing.assert_array_equal(pat.lptr.cpu().numpy(), G.indptr)
    M = FactorPreconditioner(pat, torch.from_numpy(G.data.astype(np.float32)).cuda())
    r = np.random.default_rng(2).standard_normal(A.shape[0])
    G32 = sp.csr_matrix((G.data.astype(np.float32).astype(np.float64), G.indices, G.indptr), shape=G.shape)
    z = M.apply(torch.from_numpy(r).cuda()).cpu().numpy()
    assert rel_frob(z, G32.T @ (G32 @ r)) <= 1e-14
    res = pcg(Ad, torch.from_numpy(b).cuda(), sys_ptr, M, tol=1e-8)
    for s in range(len(oracle_blocks)):
        r0, r1 = sys_ptr[s], sys_ptr[s + 1]
        Gs = G32[r0:r1, r0:r1]
        _, it = of.pcg(A[r0:r1, r0:r1], b[r0:r1], lambda v: Gs.T @ (Gs @ v), tol=1e-8)
        assert abs(int(res.iterations[s]) - it) <= 2


def test_features_and_kaporin_match_oracle(oracle_blocks):
    import torch
    from skfemntv.fsai import FSAISystems
    from skfemntv.solver import DeviceCSR
    print("case: FSAI features and Kaporin objective", flush=True)
    A, b, sys_ptr = _block_diag(oracle_blocks)
    xyz = np.concatenate([blk[2] for blk in oracle_blocks])
    fs = FSAISystems(DeviceCSR.from_scipy(A), sys_ptr, torch.from_numpy(xyz).cuda())
    lptr, lidx, F = of.features(A, xyz)
    Fg = fs.features[:fs.nnz].cpu().numpy().astype(np.float64)
    assert Fg.shape == F.shape
    # FP32 features of FP64 quantities
    np.testing.assert_allclose(Fg, F, rtol=2e-6, atol=2e-7)
    rng = np.random.default_rng(4)
    o = (0.3 * rng.standard_normal(len(lidx))).astype(np.float32)
    loss, go = fs.kaporin(torch.from_numpy(o).cuda(), grad_scale=1.0)
    lo, gr = of.kaporin(A, sys_ptr, lptr, lidx, o.astype(np.float64))
    np.testing.assert_allclose(loss.cpu().numpy(), lo, rtol=1e-12)
    # gradient is returned in FP32
    assert rel_frob(go[:fs.nnz].cpu().numpy().astype(np.float64), gr) <= 1e-6
    g = fs.factor(torch.from_numpy(o).cuda()).cpu().numpy()
    np.testing.assert_allclose(g, of.factor_values(o.astype(np.float64), lptr, lidx, A), rtol=1e-6)


def test_features_and_factor_transpose_on_dense_rows():
    """A dense 120 x 120 SPD matrix: rows of 120 entries take several chunks
    of the 16-lane pattern kernels, and the first feature blocks span more A
    entries than the kernel stages in LDS, so their rows i are read from
    global memory (the other path)."""
    import torch
    from skfemntv.fsai import FSAISystems
    from skfemntv.solver import DeviceCSR, FactorPreconditioner
    print("case: FSAI features and factor transpose, dense 120 x 120", flush=True)
    rng = np.random.default_rng(21)
    n = 120
    M = rng.standard_normal((n, n))
    A = sp.csr_matrix(M @ M.T + n * np.eye(n))
    xyz = rng.random((n, 3))
    fs = FSAISystems(DeviceCSR.from_scipy(A), [0, n], torch.from_numpy(xyz).cuda())
    lptr, lidx, F = of.features(A, xyz)
    np.testing.assert_array_equal(fs.pattern.lptr.cpu().numpy(), lptr)
    np.testing.assert_array_equal(fs.pattern.lidx.cpu().numpy(), lidx)
    Fg = fs.features[:fs.nnz].cpu().numpy().astype(np.float64)
    np.testing.assert_allclose(Fg, F, rtol=2e-6, atol=2e-7)
    o = (0.3 * rng.standard_normal(len(lidx))).astype(np.float32)
    g = fs.factor(torch.from_numpy(o).cuda())
    P = FactorPreconditioner(fs.pattern, g)
    G = sp.csr_matrix((g.cpu().numpy().astype(np.float64), lidx, lptr), shape=(n, n)).toarray()
    r = rng.standard_normal(n)
    z = P.apply(torch.from_numpy(r).cuda()).cpu().numpy()
    np.testing.assert_allclose(z, G.T @ (G @ r), rtol=1e-12, atol=1e-12 * np.abs(G.T @ (G @ r)).max())


def _random_params(seed):
    from skfemntv.fsai import init_params
    flat = init_params(seed).astype(np.float64)
    rng = np.random.default_rng(seed + 100)
    flat[18688:18816] = 0.1 * rng.standard_normal(128)   # w3
    flat[2048:2176] = 0.05 * rng.standard_normal(128)    # b1
    flat[18560:18688] = 0.05 * rng.standard_normal(128)  # b2
    flat[18816] = 0.02
    return flat.astype(np.float32)


def test_mlp_forward_backward_match_oracle():
    import torch
    from skfemntv.fsai import FSAINetwork
    print("case: BF16 MFMA MLP forward/backward", flush=True)
    flat = _random_params(3)
    net = FSAINetwork(params=flat)
    for nnz in (1, 37, 1000, 4131):
        rng = np.random.default_rng(nnz)
        X = rng.standard_normal((nnz, 16)).astype(np.float32)
        go = rng.standard_normal(nnz).astype(np.float32)
        Xd = torch.from_numpy(X).cuda()
        out = net.forward(Xd, nnz).cpu().numpy()[:nnz].astype(np.float64)
        ref, _ = of.mlp_forward(flat.astype(np.float64), X.astype(np.float64))
        # FP32 accumulation order inside MFMA vs FP64 in the oracle, and rare
        # BF16 rounding-tie flips of hidden activations
        assert rel_frob(out, ref) <= 1e-3, nnz
        grad, _ = net.backward(Xd, torch.from_numpy(go).cuda(), nnz)
        g = grad.cpu().numpy().astype(np.float64)
        gref = of.mlp_backward(flat.astype(np.float64), X.astype(np.float64), go.astype(np.float64))
        for name, a, b in (("W1", 0, 2048), ("b1", 2048, 2176), ("W2", 2176, 18560), ("b2", 18560, 18688),
                           ("w3", 18688, 18816), ("b3", 18816, 18817)):
            assert rel_frob(g[a:b], gref[a:b]) <= 5e-3, (nnz, name, rel_frob(g[a:b], gref[a:b]))


def test_mlp_large_batch_matches_oracle():
    """300,001 rows: past the inference grid's 2,048 blocks (262,144 rows per
    pass, so the grid-stride loop runs twice) and the backward kernels' block
    caps (512 activation blocks = 65,536 rows, 512 weight-gradient blocks of
    eight chunks), with a ragged last chunk."""
    import torch
    from skfemntv.fsai import FSAINetwork
    print("case: BF16 MFMA MLP forward/backward, 300001 rows", flush=True)
    flat = _random_params(7)
    net = FSAINetwork(params=flat)
    nnz = 300_001
    rng = np.random.default_rng(11)
    X = rng.standard_normal((nnz, 16)).astype(np.float32)
    go = (rng.standard_normal(nnz) / nnz).astype(np.float32)
    Xd = torch.from_numpy(X).cuda()
    out = net.forward(Xd, nnz).cpu().numpy()[:nnz].astype(np.float64)
    ref, _ = of.mlp_forward(flat.astype(np.float64), X.astype(np.float64))
    assert rel_frob(out, ref) <= 1e-3
    grad, _ = net.backward(Xd, torch.from_numpy(go).cuda(), nnz)
    g = grad.cpu().numpy().astype(np.float64)
    gref = of.mlp_backward(flat.astype(np.float64), X.astype(np.float64), go.astype(np.float64))
    for name, a, b in (("W1", 0, 2048), ("b1", 2048, 2176), ("W2", 2176, 18560), ("b2", 18560, 18688),
                       ("w3", 18688, 18816), ("b3", 18816, 18817)):
        assert rel_frob(g[a:b], gref[a:b]) <= 5e-3, (name, rel_frob(g[a:b], gref[a:b]))


def test_mlp_backward_deterministic():
    import torch
    from skfemntv.fsai import FSAINetwork
    print("case: MLP backward determinism", flush=True)
    net = FSAINetwork(params=_random_params(5))
    rng = np.random.default_rng(9)
    X = torch.from_numpy(rng.standard_normal((20000, 16)).astype(np.float32)).cuda()
    go = torch.from_numpy(rng.standard_normal(20000).astype(np.float32)).cuda()
    g1, ws = net.backward(X, go, 20000)
    g2, _ = net.backward(X, go, 20000, ws)
    assert torch.equal(g1, g2)


def test_training_decreases_loss_and_zero_net_is_jacobi(batch):
    import torch
    from skfemntv.fsai import FSAINetwork, LearnedFSAI, FSAISystems, mean_loss, split_groups, train
    from skfemntv.solver import JacobiPreconditioner, pcg
    print("case: FSAI training steps", flush=True)
    A, b = batch.assemble()
    groups = split_groups(A, batch.sys_ptr_host, batch.xyz, 1)
    net = FSAINetwork(seed=0)
    before = mean_loss(net, groups)
    assert abs(before) < 1e-12   # zero output layer: Gt = I
    fs = FSAISystems(A, batch.sys_ptr, batch.xyz)
    res_zero = pcg(A, b, batch.sys_ptr, LearnedFSAI(net)(fs), tol=1e-8)
    res_jac = pcg(A, b, batch.sys_ptr, JacobiPreconditioner(A), tol=1e-8)
    assert np.all(np.abs(res_zero.iterations - res_jac.iterations) <= 1)
    hist = train(net, groups, epochs=6, lr=3e-3, seed=0)
    after = mean_loss(net, groups)
    assert after < before - 0.05, (before, after, hist.train_loss)
    net2 = FSAINetwork(seed=0)
    train(net2, groups, epochs=6, lr=3e-3, seed=0)
    assert torch.equal(net.params, net2.params)


def test_pretrained_network_gain_over_jacobi():
    """The shipped network (trained on 2,000 assembled systems with cells
    per axis in {8..16}) against the same preconditioner without it (zero
    network output = Jacobi) on held-out, larger systems."""
    import time
    import torch
    from skfemntv.batch import AnisotropicBatch, batch_shapes
    from skfemntv.fsai import FSAINetwork, FSAISystems, LearnedFSAI
    from skfemntv.solver import JacobiPreconditioner, pcg, relative_residuals
    print("case: pretrained FSAI vs Jacobi, 8 held-out systems 14..22 cells/axis", flush=True)
    b = AnisotropicBatch(batch_shapes(8, 14, 22, 123), seed=4242)
    A, rhs = b.assemble()
    net = FSAINetwork.pretrained()
    rj = pcg(A, rhs, b.sys_ptr, JacobiPreconditioner(A), tol=1e-8)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    M = LearnedFSAI(net)(FSAISystems(A, b.sys_ptr, b.xyz))
    rl = pcg(A, rhs, b.sys_ptr, M, tol=1e-8)
    learned_s = time.perf_counter() - t0
    gain = rj.iterations.astype(float) / rl.iterations
    print(f"jacobi iterations {rj.iterations.tolist()} ({rj.seconds * 1e3:.2f} ms), learned "
          f"{rl.iterations.tolist()} ({learned_s * 1e3:.2f} ms incl. features + inference); "
          f"gain {gain.min():.2f}..{gain.max():.2f}", flush=True)
    assert np.all(gain >= 1.8), gain
    for res in (rj, rl):
        assert np.all(relative_residuals(A, rhs, res.x, b.sys_ptr) <= 1.5e-8)
    # both solve the same systems
    assert rel_frob(rl.x.cpu().numpy(), rj.x.cpu().numpy()) <= 1e-6
