"""Benchmark: FP64 P2-tetrahedral anisotropic-diffusion assembly on one MI355X.

Workload (one "step" = one full assembly of the CSR values):
  64^3 Kuhn-split unit cube (1,572,864 tets, 2,146,689 P2 DOFs), interior
  vertices jittered uniformly by +-0.2h (seed 0), per-element SPD tensor
  A = Q diag(lambda) Q^T with Q a Haar rotation and lambda log-uniform in
  [1e-2, 1e2] (seed 1), 4-point quadrature, form dot(mul(A, grad u), grad v).
Inputs (vertex coordinates, connectivity, A, gather plan) are resident in
HBM before timing; the timed region is K back-to-back assemblies, after an
untimed clock ramp (--ramp seconds) and W warm-up steps.

Prints one JSON line.  `value` = assembled elements/s of the FP64 path.
`roofline` uses the algorithmic bytes of one assembly (825.7 B per element:
vertices, int32 connectivity, 72 B coefficient, int32 element->CSR map,
FP64 CSR values) over the kernel's mean duration measured with HIP events on
the launch stream.  `configs` holds the C1 (P1 Laplace 32^3), C3 (vector-P2
elasticity 48^3) and C4 (N1 96^3) assemblies measured the same way,
`batched_fsai` the 256-system batch and its PCG solves, and `cpu_baseline`
the C restatement of the reference's threaded CPU assembler (oracle/, built
from this repository) on the full headline workload, with every host core
and with one thread; `side_workloads` the Taylor-Hood divergence block, the
Taylor-Hood Stokes operator and the cut-cell band.
"""

from __future__ import annotations

import argparse
import ctypes
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "scikit-fem-native_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

import numpy as np  # noqa: E402

METRIC = "assembled elements/s + % HBM roofline; learned-FSAI vs Jacobi PCG time to 1e-8"
HBM_PEAK_GBS = 8000.0
BF16_MFMA_PEAK_TFLOPS = 2500.0   # MI355X dense BF16 matrix peak
BYTES_PER_ELEMENT_C2 = 825.7  # algorithmic bytes per tet of one FP64 assembly
# the same without the int32 element->CSR map (10 x 10 entries per tet): the "floor"
FLOOR_BYTES_PER_ELEMENT_C2 = BYTES_PER_ELEMENT_C2 - 4. * 10 * 10
# rocprofv3 FETCH_SIZE/WRITE_SIZE summaries of the other workloads at their default sizes
TRAFFIC_FILES = {"th": "th_divergence_block_64_hbm_traffic.json",
                 "c1": "c1_p1_laplace_32_hbm_traffic.json",
                 "c3": "c3_vector_p2_elasticity_48_hbm_traffic.json",
                 "c4": "c4_n1_curlcurl_96_hbm_traffic.json",
                 "stokes": "stokes_taylor_hood_48_hbm_traffic.json",
                 "cut": "cut_band_p2_64_hbm_traffic.json"}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def spd_tensors(E, seed):
    """Per-element SPD A = Q diag(10^u) Q^T, Q Haar (unit quaternion), u ~ U(-2, 2)."""
    from skfemntv.batch import random_spd_tensors
    return random_spd_tensors(E, np.random.default_rng(seed))


def jittered_p2_cube(n, amount, seed):
    import skfemntv as sk
    ax = np.linspace(0., 1., n + 1)
    mesh = sk.MeshTet.init_tensor(ax, ax, ax)
    p = mesh.p.copy()
    interior = np.all((p > 1e-12) & (p < 1 - 1e-12), axis=0)
    rng = np.random.default_rng(seed)
    p[:, interior] += rng.uniform(-amount / n, amount / n, size=(3, int(interior.sum())))
    return sk.MeshTet2.from_mesh(sk.MeshTet(p, mesh.t))


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return ""


def cpu_baseline(mesh, A_host, repeats=7, warmup=2):
    """The reference's threaded CPU algorithm, restated in C (oracle/fem_oracle.c,
    gcc -O3), on the full headline workload.

    Same data flow as the reference's CrossBilinearAssembler gradient x
    gradient tensor path (cross_bilinear_assembler.cpp:19-107,207-247): the
    geometry is tabulated once (basis_geometry.cpp:49-209), the constructor
    builds the CSR pattern, the int64 element->CSR map and the greedy
    first-fit element colouring over row DOFs (cross_bilinear_assembler.cpp:
    307-341), and assemble() zeroes the values and runs the fused q x a x b
    loop with scatter, colour by colour, each colour split statically over
    the threads (parallel.hpp:21-45).  Timed warm, constructor excluded
    (reported as cold setup), median of `repeats` after `warmup` calls
    (benchmarks/hcurl_tet_n1_assembly.py:18-26), with every host core and
    with 1 thread (the reference default, runtime.cpp:10)."""
    import oracle
    from oracle import host, pipeline
    p2, t2 = mesh.p, mesh.t
    E = t2.shape[1]
    X, W = host.tet_rule_4()
    t0 = time.perf_counter()
    (shape, grads, dx, x, det, tol, cond), _ = pipeline.basis_geometry(p2, t2, X, W, True)
    geometry_s = time.perf_counter() - t0
    del shape, x, det, tol, cond
    A_eq = np.ascontiguousarray(np.broadcast_to(A_host[:, None], (E, 4, 3, 3)))
    t0 = time.perf_counter()
    N, dofs = pipeline.dofs_of(t2, p2.shape[1], 10, 1)
    rows = dofs.reshape(E, -1)
    ip, ix, sc = oracle.pattern(rows, nrows=N)
    colour, ncolours = oracle.colouring(rows, N)
    constructor_s = time.perf_counter() - t0
    nnz = len(ix)
    del ip, ix
    cores = len(os.sched_getaffinity(0))
    med = {}
    for threads in (cores, 1):
        for _ in range(warmup):
            oracle.assemble_tensor(grads, dx, A_eq, sc, nnz, threads=threads, colour=colour)
        ts = []
        for _ in range(repeats):
            t0 = time.perf_counter()
            oracle.assemble_tensor(grads, dx, A_eq, sc, nnz, threads=threads, colour=colour)
            ts.append(time.perf_counter() - t0)
        med[threads] = statistics.median(ts)
        log(f"cpu baseline: {threads} threads {med[threads]:.3f} s per assembly")
    return {
        "value": E / med[cores],
        "unit": "elements/s",
        "cores": cores,
        "kind": "port",
        "label": "reference algorithm, CPU restatement",
        "sample": (f"the full headline workload: P2 anisotropic 64^3 jittered Kuhn cube ({E} tets, nnz {nnz}), "
                   f"4-pt rule, FP64; C restatement of the reference's threaded CrossBilinearAssembler "
                   f"(greedy first-fit colouring, {ncolours} colours, static per-colour split), gcc -O3; "
                   f"warm median of {repeats} after {warmup} warm-up calls; CPU: {cpu_model()}"),
        "seconds_per_assembly": med[cores],
        "one_thread": {"value": E / med[1], "seconds_per_assembly": med[1]},
        "cold_setup_s": {"geometry": geometry_s, "pattern_map_colouring": constructor_s},
    }


def batched_fsai(systems, lo, hi, steps, warmup, timer):
    """256 P1 anisotropic systems on uneven boxes: batched assembly (one
    HIP-graph replay per step), then FP64 PCG to ||b - Ax||/||b|| <= 1e-8
    with Jacobi and with the pretrained learned FSAI."""
    import torch
    from skfemntv.batch import AnisotropicBatch, batch_shapes
    from skfemntv.fsai import FSAINetwork, FSAISystems, LearnedFSAI
    from skfemntv.solver import JacobiPreconditioner, pcg, relative_residuals
    t0 = time.perf_counter()
    shapes = batch_shapes(systems, lo, hi, 5)
    batch = AnisotropicBatch(shapes, seed=5000)
    torch.cuda.synchronize()
    setup = time.perf_counter() - t0
    log(f"batched: {systems} systems, {batch.nelements} tets, {batch.A.shape[0]} unknowns, "
        f"nnz {batch.A.nnz}, setup {setup:.1f}s ({batch.timings})")
    for _ in range(min(max(warmup, 1), 20)):   # 5 ms each; the GPU is warm by now
        batch.assemble(graph=True)
    asm_s = timer.run(lambda: batch.assemble(graph=True), steps)
    # algorithmic bytes of one batched assembly: vertices, int32 P1
    # connectivity, the 72 B tensor, the int32 4 x 4 element->CSR map and the
    # FP64 values of the full (uncondensed) matrices, plus the load vector
    plan = batch.form.plan
    E, V = batch.nelements, batch.mesh.p.shape[1]
    alg = 24. * V + (16. + 72. + 64.) * E + 8. * plan.nnz + 8. * plan.nrows
    A, b = batch.A, batch.b
    net = FSAINetwork.pretrained()
    out = {}
    for rep in range(2):   # the second repetition is reported (first one warms up)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        rj = pcg(A, b, batch.sys_ptr, JacobiPreconditioner(A), tol=1e-8)
        jac_s = time.perf_counter() - t0
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        fs = FSAISystems(A, batch.sys_ptr, batch.xyz)
        M = LearnedFSAI(net)(fs)
        torch.cuda.synchronize()
        setup_s = time.perf_counter() - t0
        rl = pcg(A, b, batch.sys_ptr, M, tol=1e-8)
        learned_s = time.perf_counter() - t0
        out = {"jacobi": rj, "jac_s": jac_s, "learned": rl, "learned_s": learned_s, "setup_s": setup_s}
    rj, rl = out["jacobi"], out["learned"]
    # measured HBM bytes of one batched assembly (committed rocprofv3 summary)
    batch_traffic = None
    tpath = os.path.join(ROOT, "profiles", "c5_batched_p1_256_hbm_traffic.json")
    if systems == 256 and (lo, hi) == (16, 48) and os.path.exists(tpath):
        with open(tpath) as f:
            batch_traffic = json.load(f)["per_assembly_bytes"]
    # the MLP kernels on the batch's lower nonzeros: inference (part of the
    # setup above) and one training backward pass (activation kernel, split-K
    # weight gradients, reduction).  MFMA flops per nonzero: forward
    # 2 (16*128 + 128*128); backward the forward again, W2^T dz2, and the
    # two weight-gradient products.
    _, go = fs.kaporin(net.forward(fs.features, fs.nnz))
    _, ws = net.backward(fs.features, go, fs.nnz, None)
    fwd_s = timer.run(lambda: net.forward(fs.features, fs.nnz), 10)
    bwd_s = timer.run(lambda: net.backward(fs.features, go, fs.nnz, ws), 5)
    f_fwd = 2. * (16 * 128 + 128 * 128)
    f_bwd = f_fwd + 2. * 128 * 128 + 2. * (128 * 128 + 16 * 128)
    mlp = {"rows": fs.nnz, "forward_ms": fwd_s * 1e3, "backward_ms": bwd_s * 1e3,
           "forward_tflops": f_fwd * fs.nnz / fwd_s / 1e12, "backward_tflops": f_bwd * fs.nnz / bwd_s / 1e12,
           "peak_tflops": BF16_MFMA_PEAK_TFLOPS, "dtype": "bf16 MFMA, fp32 accumulate",
           "flops_definition": "per nonzero: forward 2(16*128 + 128*128); backward the forward, "
                                   "2*128*128 for W2^T dz2 and 2(128*128 + 16*128) for the weight gradients"}
    mlp["forward_frac"] = mlp["forward_tflops"] / BF16_MFMA_PEAK_TFLOPS
    mlp["backward_frac"] = mlp["backward_tflops"] / BF16_MFMA_PEAK_TFLOPS
    log(f"mlp: forward {mlp['forward_ms']:.3f} ms ({mlp['forward_tflops']:.0f} TFLOP/s), "
        f"backward {mlp['backward_ms']:.3f} ms ({mlp['backward_tflops']:.0f} TFLOP/s)")
    res_j = relative_residuals(A, b, rj.x, batch.sys_ptr)
    res_l = relative_residuals(A, b, rl.x, batch.sys_ptr)
    return {
        "workload": f"{systems} P1 anisotropic systems, cells/axis iid U{{{lo}..{hi}}} (seed 5), jitter 0.2h, "
                    f"per-element SPD A (seed 5000+s), f=1, boundary condensed",
        "elements": E, "unknowns": A.shape[0], "nnz": A.nnz, "nnz_uncondensed": plan.nnz,
        "assembly_ms_per_batch": asm_s * 1e3, "assembly_elements_per_s": E / asm_s,
        "assembly": "payload + row-gather with the load vector, matrix and load written condensed through "
                    "position maps, replayed as one HIP graph",
        "assembly_roofline": {"bound": "hbm", "bytes": alg, "achieved": alg / asm_s / 1e9, "peak": HBM_PEAK_GBS,
                              "unit": "GB/s", "frac": alg / asm_s / 1e9 / HBM_PEAK_GBS,
                              "bytes_definition": "24 V + (16 conn + 72 tensor + 64 int32 map) E + 8 nnz + 8 N",
                              "traffic": batch_traffic},
        "setup_s": setup,
        "jacobi_pcg": {"wall_s": out["jac_s"], "batch_iterations": rj.batch_iterations,
                       "iterations_mean": float(rj.iterations.mean()), "iterations_max": int(rj.iterations.max()),
                       "true_residual_max": float(res_j.max())},
        "learned_fsai_pcg": {"wall_s": out["learned_s"], "setup_s": out["setup_s"],
                             "solve_s": out["learned_s"] - out["setup_s"],
                             "batch_iterations": rl.batch_iterations,
                             "iterations_mean": float(rl.iterations.mean()),
                             "iterations_max": int(rl.iterations.max()),
                             "true_residual_max": float(res_l.max()), "lower_nonzeros": fs.nnz},
        "speedup_learned_vs_jacobi": out["jac_s"] / out["learned_s"],
        "mlp": mlp,
    }


class Timer:
    """Mean seconds per call over `steps` back-to-back calls, timed with HIP
    events on the stream the kernels are launched on."""

    def __init__(self):
        from skfemntv import _native
        self.native = _native
        self.L = _native.lib()
        self.stream = _native.stream()
        self.ev0, self.ev1 = ctypes.c_void_p(), ctypes.c_void_p()
        _native.check(self.L.skfa_event_create(ctypes.byref(self.ev0)))
        _native.check(self.L.skfa_event_create(ctypes.byref(self.ev1)))

    def run(self, fn, steps):
        import torch
        torch.cuda.synchronize()
        self.native.check(self.L.skfa_event_record(self.ev0, self.stream))
        for _ in range(steps):
            fn()
        self.native.check(self.L.skfa_event_record(self.ev1, self.stream))
        ms = ctypes.c_float()
        self.native.check(self.L.skfa_event_elapsed_ms(self.ev0, self.ev1, ctypes.byref(ms)))
        torch.cuda.synchronize()
        return ms.value / 1e3 / steps


def ramp(fn, seconds):
    """Untimed back-to-back calls until `seconds` of wall time have passed:
    the GPU reaches steady clocks after ~0.1 s of continuous assembly (C2:
    3 warm-up steps measure 0.38 ms per step, 30 measure 0.33, 200 measure
    0.32).  Returns the number of calls."""
    import torch
    t0, calls = time.perf_counter(), 0
    while True:
        for _ in range(10):
            fn()
        calls += 10
        torch.cuda.synchronize()
        if time.perf_counter() - t0 >= seconds:
            return calls


def build_workload(name):
    """The other assembly configurations at full size.  Returns a dict with
    the step function (precision -> None), the FP64 values tensor (or None),
    elements, nnz, algorithmic bytes per element and a description.

    c1 P1 Laplace + load 32^3, c3 vector-P2 elasticity 48^3, c4 N1 curl-curl
    + kappa mass 96^3, th the Taylor-Hood divergence block (vector P2 x P1)
    on the 64^3 jittered cube, stokes the Taylor-Hood Stokes operator on a
    composite basis at 48^3, cut a spherical band of cut cells (12 points
    each).  Algorithmic bytes count unique vertex coordinates, int32
    connectivity, coefficients as given, the int32 element->CSR map at node
    level (vector fields interleave their components inside a node block,
    so a 10 x 10 node pair is one map entry), and the FP64 values written
    once."""
    import torch
    import skfemntv as sk
    from skfemntv._assembly import NativeBilinearForm, upload
    from skfemntv.hcurl import AffineTetN1Basis, TetN1Assembler
    t0 = time.perf_counter()
    values = None
    fp32 = False
    if name == "c1":
        ax = np.linspace(0., 1., 33)
        mesh = sk.MeshTet.init_tensor(ax, ax, ax)
        basis = sk.Basis(mesh, sk.ElementVector(sk.ElementTetP1(), dim=1))
        E = mesh.nelements
        form = NativeBilinearForm(basis)
        one = upload(np.ones(E))
        f = upload(np.array([1.]))
        rhs = torch.empty(basis.N, dtype=torch.float64, device=one.device)

        def step(precision="fp64"):
            # matrix and load vector in one call (skfa_assemble_bilinear_load)
            form.assemble_device(gradient=one, precision=precision, load=(f, rhs))
        bpe, desc, nnz = 106.7, f"P1 Laplace ddot(grad u, grad v) + load f=1, 32^3 Kuhn cube, {E} tets, FP64", form.plan.nnz
        values, fp32 = form._values, True
    elif name == "c3":
        n = 48
        ax = np.linspace(0., 1., n + 1)
        m = sk.MeshTet.init_tensor(ax, ax, ax)
        p = m.p.copy()
        interior = np.all((p > 1e-12) & (p < 1 - 1e-12), axis=0)
        rng = np.random.default_rng(0)
        p[:, interior] += rng.uniform(-0.15 / n, 0.15 / n, size=(3, int(interior.sum())))
        mesh = sk.MeshTet2.from_mesh(sk.MeshTet(p, m.t))
        X, W = sk.basis.tet_rule_4()
        basis = sk.Basis(mesh, sk.ElementVector(sk.ElementTetP2(), dim=3), quadrature=(X, W))
        E = mesh.nelements
        rng = np.random.default_rng(1)
        Ey = np.exp(rng.standard_normal(E))
        nu = rng.uniform(0.2, 0.45, E)
        lam = Ey * nu / ((1 + nu) * (1 - 2 * nu))
        mu = Ey / (2 * (1 + nu))
        form = NativeBilinearForm(basis)
        s_dev, l_dev = upload(2 * mu), upload(lam)

        def step(precision="fp64"):
            form.assemble_device(symmetric_gradient=s_dev, divergence=l_dev, precision=precision)
        bpe, nnz = 3255.0, form.plan.nnz
        desc = f"vector-P2 elasticity 2 mu eps:eps + lam div div, 48^3 Kuhn cube jittered 0.15h, {E} tets, FP64"
        values, fp32 = form._values, True
    elif name == "c4":
        ax = np.linspace(0., 1., 97)
        mesh = sk.MeshTet.init_tensor(ax, ax, ax)
        basis = AffineTetN1Basis(mesh)
        E = mesh.nelements
        asm_ = TetN1Assembler(basis)
        kappa = upload(np.random.default_rng(2).uniform(0.1, 10., E))
        one = upload(np.ones(E))

        def step(precision="fp64"):
            asm_.assemble_device(kappa, one, False)
        bpe, nnz = 356.0, asm_.plan.nnz
        desc = f"N1 curl-curl + kappa mass, 96^3 Kuhn cube, {E} tets, FP64"
    elif name == "th":
        # Taylor-Hood divergence block B = int q div u: trial vector P2,
        # test P1, on the 64^3 jittered P2 cube of the headline workload
        mesh = jittered_p2_cube(64, 0.2, 0)
        t2 = mesh.t
        X, W = sk.basis.tet_rule_4()
        vp2 = sk.Basis(mesh, sk.ElementVector(sk.ElementTetP2(), dim=3), quadrature=(X, W))
        p1 = sk.Basis(mesh, sk.ElementVector(sk.ElementTetP1(), dim=1), quadrature=(X, W))
        E = mesh.nelements
        form = sk.NativeCrossBilinearForm(p1, vp2)
        plan = form.cross_plan()
        one = upload(np.ones(1))

        def step(precision="fp64"):
            form.assemble_cross_device("value", "gradient", "column_divergence", one, (0, 0))
        nnz = plan.nnz
        V = int(np.unique(t2[:4]).size)
        # vertices (24 B each) + int32 connectivity of both fields (4 + 10
        # nodes) + int32 node-level element->CSR map (4 x 10 entries) + FP64 values
        bpe = (24. * V + 4. * 14 * E + 4. * 4 * 10 * E + 8. * nnz) / E
        desc = f"Taylor-Hood divergence block div(u) q, vector P2 x P1, 64^3 Kuhn cube jittered 0.2h, {E} tets, FP64"
    elif name == "stokes":
        # Taylor-Hood Stokes operator on a composite basis (vector P2 x P1),
        # 48^3 Kuhn cube jittered 0.15h: nu grad u : grad v on the velocity
        # block, -div(v) p and -div(u) q on the coupling blocks
        n = 48
        ax = np.linspace(0., 1., n + 1)
        m = sk.MeshTet.init_tensor(ax, ax, ax)
        p = m.p.copy()
        interior = np.all((p > 1e-12) & (p < 1 - 1e-12), axis=0)
        rng = np.random.default_rng(0)
        p[:, interior] += rng.uniform(-0.15 / n, 0.15 / n, size=(3, int(interior.sum())))
        mesh = sk.MeshTet2.from_mesh(sk.MeshTet(p, m.t))
        X, W = sk.basis.tet_rule_4()
        basis = sk.Basis(mesh, sk.ElementVector(sk.ElementTetP2(), dim=3) * sk.ElementTetP1(), quadrature=(X, W))
        E = mesh.nelements
        comp = sk.NativeCompositeBilinearForm(basis)
        velocity = comp._square(0)
        bt, b = comp._cross(0, 1), comp._cross(1, 0)
        nu = upload(np.random.default_rng(1).uniform(0.5, 2., E))
        one = upload(np.ones(1))

        def step(precision="fp64"):
            velocity.assemble_device(gradient=nu)
            bt.assemble_cross_device("gradient", "value", "row_divergence", one, (0, 0))
            b.assemble_cross_device("value", "gradient", "column_divergence", one, (0, 0))
        nnz_a, nnz_b = velocity.plan.nnz, b.cross_plan().nnz
        nnz = nnz_a + 2 * nnz_b
        V = int(np.unique(mesh.t[:4]).size)
        # per block: vertices + int32 connectivity + coefficient + int32
        # node-level element->CSR map + FP64 values (velocity 10 x 10 node
        # pairs, coupling 10 x 4)
        bpe = (3 * 24. * V + (40. + 8. + 4. * 100) * E + 8. * nnz_a
               + 2 * ((56. + 4. * 40) * E + 8. * nnz_b)) / E
        desc = (f"Taylor-Hood Stokes (vector P2 x P1 composite): nu grad:grad + both divergence blocks, 48^3 "
                f"Kuhn cube jittered 0.15h, {E} tets, FP64")
    elif name == "cut":
        # cut-cell (ragged quadrature) value + gradient form of a P2 field:
        # the cells of the 64^3 P2 cube whose centroid lies within one cell
        # width of the sphere |x - c| = 0.3, 12 random points each
        from skfemntv import _native
        from skfemntv.basis import tabulate
        from skfemntv.cut import CUT_GRADIENT, CUT_VALUE
        mesh = jittered_p2_cube(64, 0.2, 0)
        c = mesh.p[:, mesh.t[:4]].mean(axis=1)
        band = np.abs(np.linalg.norm(c - 0.5, axis=0) - 0.3) < 1. / 64
        rng = np.random.default_rng(3)
        Pc = 12
        counts = np.where(band, Pc, 0)
        off = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
        P = int(off[-1])
        X = rng.dirichlet(np.ones(4), P)[:, 1:]
        shape, ref_grad = tabulate(X.T, True)
        corners = mesh.p[:, mesh.t[:4][:, band]]
        Jinv = np.linalg.inv(np.moveaxis(corners[:, 1:4] - corners[:, [0]], -1, 0))
        grads = np.einsum("pnk,pki->pni", ref_grad, np.repeat(Jinv, Pc, axis=0))
        w = rng.uniform(0.1, 1., P) / (64. ** 3 * 6 * Pc)
        dofs = mesh.t[:10].T[:, :, None].astype(np.int64)
        asm_ = sk.CutBilinearFormAssembler(dofs, off, shape, grads, w)
        pl = asm_._plan
        dv, dg = upload(rng.uniform(0.5, 2., P)), upload(rng.uniform(0.5, 2., P))
        vals = torch.zeros(max(pl.nnz, 1), dtype=torch.float64, device=_native.device())
        dd = asm_._dev
        Lb = _native.lib()

        def step(precision="fp64"):
            _native.check(Lb.skfa_assemble_cut_bilinear(
                pl.handle, _native.ptr(pl.offsets), _native.ptr(dd["weights"]), _native.ptr(dd["shape"]),
                _native.ptr(dd["gradients"]), 3, _native.ptr(dd["shape"]), _native.ptr(dd["gradients"]), 3,
                CUT_VALUE | CUT_GRADIENT, _native.ptr(dv), _native.ptr(dg), 0, 0, None, 0, _native.ptr(vals),
                _native.stream()))
        E = int(band.sum())
        nnz = pl.nnz
        # per point: shape (10) + gradients (30) + weight + two coefficients,
        # FP64; per cell: offsets, int32 connectivity; FP64 values written once
        bpe = ((10 + 30 + 3) * 8. * P + (8. + 40.) * E + 8. * nnz) / E
        desc = (f"cut cells: P2 value + gradient form on the {E} cells of a spherical band in the 64^3 cube, "
                f"{Pc} points each ({P} points), FP64")
    else:
        raise SystemExit(f"unknown workload {name}")
    torch.cuda.synchronize()
    # first call (the autotuner times the launch shapes of a new mesh shape here)
    step("fp64")
    torch.cuda.synchronize()
    plan = {"c1": lambda: form.plan, "c3": lambda: form.plan, "c4": lambda: asm_.plan}.get(name, lambda: None)()
    return {"name": name, "step": step, "values": values, "fp32": fp32, "E": E, "nnz": nnz, "bpe": bpe,
            "desc": desc, "setup_s": time.perf_counter() - t0, "plan": plan}


def autotune_summary(plan):
    """The launch shape the autotuner chose for this plan's mesh signature
    (timed on the first assembly of a plan of this shape)."""
    from skfemntv import autotune
    rec = autotune.LAST.get(autotune.signature(plan)) or autotune._CACHE.get(autotune.signature(plan))
    if rec is None:
        return None
    return {"lds_budget": rec.budget, "row_lanes": rec.row_lanes, "settings": rec.settings,
            "best_ms": rec.seconds_per_call * 1e3, "candidates": rec.candidates, "tuning_s": rec.tuning_seconds,
            "current": {"lds_budget": plan.lds_budget, "row_lanes": plan.row_lanes,
                        **{k: plan.tuning[k] for k in rec.settings}}}


def measure(w, steps, warmup, ramp_s, timer):
    """Time one workload: ramp, `warmup` steps, `steps` timed FP64 steps;
    for workloads with an FP32-products form, FP32 and FP64 blocks are
    timed alternately (FP64, FP32, FP64, FP32) after the same warm-up."""
    import torch
    step = w["step"]
    fp32 = w["fp32"]
    ramp(lambda: (step("fp64"), step("fp32") if fp32 else None), ramp_s)
    for _ in range(warmup):
        step("fp64")
        if fp32:
            step("fp32")
    sec = timer.run(lambda: step("fp64"), steps)
    out = {"ms_per_step": sec * 1e3, "elements_per_s": w["E"] / sec, "nnz_per_s": w["nnz"] / sec,
           "bytes_per_element": w["bpe"], "achieved_gbs": w["bpe"] * w["E"] / sec / 1e9,
           "roofline_frac": w["bpe"] * w["E"] / sec / 1e9 / HBM_PEAK_GBS, "setup_s": w["setup_s"]}
    if fp32:
        s32 = timer.run(lambda: step("fp32"), steps)
        s64b = timer.run(lambda: step("fp64"), steps)
        s32b = timer.run(lambda: step("fp32"), steps)
        step("fp64")
        v64 = w["values"].clone()
        step("fp32")
        v32 = w["values"].clone()
        err = float(torch.linalg.norm(v32 - v64) / torch.linalg.norm(v64))
        step("fp64")
        out["fp32_products"] = {"ms_per_step": 0.5 * (s32 + s32b) * 1e3,
                                "speedup_vs_fp64": (sec + s64b) / (s32 + s32b),
                                "rel_frobenius_vs_fp64": err,
                                "timing": "alternating FP64/FP32 blocks of `steps` each after one warm-up"}
    return sec, out


def measured_traffic(name):
    """Measured HBM bytes of one assembly (FETCH_SIZE + WRITE_SIZE, the write
    counter calibrated on a fill of known size), from the committed
    rocprofv3 summary of that workload, or None."""
    path = os.path.join(ROOT, "profiles", TRAFFIC_FILES.get(name, ""))
    if name not in TRAFFIC_FILES or not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)["per_assembly_bytes"]


def other_workload(name, steps, warmup, ramp_s, timer, vector_mfma=False):
    """One JSON line for a single non-headline workload (--workload);
    ``vector_mfma`` puts the C3 plan on the vector-P2 MFMA gather's layout."""
    w = build_workload(name)
    log(f"{name}: setup {w['setup_s']:.1f}s")
    if vector_mfma:
        if name != "c3":
            raise SystemExit("--vector-mfma applies to --workload c3")
        w["plan"].rebuild(40 * 1024, row_lanes=4)
        w["plan"]._tuned = True
    sec, summary = measure(w, steps, warmup, ramp_s, timer)
    E, nnz, bpe = w["E"], w["nnz"], w["bpe"]
    out = {"metric": METRIC, "value": E / sec, "unit": "elements/s", "n_gpus": 1, "steps": steps, "warmup": warmup,
           "ms_per_step": sec * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f64", "data": "synthetic",
           "config": {"workload": w["desc"], "elements": E, "nnz": nnz, "parallelism": "single GPU"},
           "roofline": {"bound": "hbm", "achieved": bpe * E / sec / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": bpe * E / sec / 1e9 / HBM_PEAK_GBS, "traffic": None},
           "nnz_per_s": nnz / sec, "setup_s": w["setup_s"]}
    out["roofline"]["traffic"] = measured_traffic(name)
    if "fp32_products" in summary:
        out["fp32_products"] = summary["fp32_products"]
    return out


def headline(args, timer):
    """C2: FP64 P2 anisotropic diffusion on the 64^3 jittered cube."""
    import torch
    import skfemntv as sk
    from skfemntv._assembly import NativeCrossBilinearForm, upload
    from skfemntv.helpers import dot, grad, mul
    t0 = time.perf_counter()
    mesh = jittered_p2_cube(args.n, 0.2, 0)
    E = mesh.nelements
    log(f"mesh {E} tets, {mesh.p.shape[1]} nodes ({time.perf_counter() - t0:.1f}s)")
    X, W = sk.basis.tet_rule_4()
    A_host = spd_tensors(E, 1)

    @sk.BilinearForm
    def aniso(u, v, w):
        return dot(mul(w.diffusion, grad(u)), grad(v))

    # the recommended (3,3,E,Q) layout, a broadcast view of the per-element tensors
    diffusion = np.broadcast_to(np.moveaxis(A_host, 0, -1)[..., None], (3, 3, E, 4))
    # cold first call: geometry check, device plan (pattern + gather map),
    # the autotuner's timing of launch shapes (first plan of this mesh
    # shape only), coefficient upload, kernels and the CSR copied to the host
    t1 = time.perf_counter()
    basis_cold = sk.Basis(mesh, sk.ElementVector(sk.ElementTetP2(), dim=1), quadrature=(X, W))
    K = sk.asm(aniso, basis_cold, diffusion=diffusion)
    cold_asm_s = time.perf_counter() - t1
    del K, basis_cold
    log(f"cold first asm (new basis) {cold_asm_s:.2f}s")
    t1 = time.perf_counter()
    basis = sk.Basis(mesh, sk.ElementVector(sk.ElementTetP2(), dim=1), quadrature=(X, W))
    basis_s = time.perf_counter() - t1
    form = NativeCrossBilinearForm(basis, basis)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    plan = form.plan          # builds the pattern, incidences and gather records on the GPU
    torch.cuda.synchronize()
    plan_s = time.perf_counter() - t1
    log(f"basis + geometry check {basis_s:.2f}s; plan (pattern + gather map on the GPU) {plan_s:.3f}s, "
        f"nnz={plan.nnz}")
    A = upload(A_host)
    if args.copy_mib:
        src = torch.ones(args.copy_mib * (1 << 20) // 8, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    def step(precision="fp64"):
        form.assemble_tensor_device(A, precision=precision, symmetric=True)

    ramp_calls = ramp(lambda: (step("fp64"), step("fp32")), args.ramp)
    for _ in range(args.warmup):
        step("fp64")
        step("fp32")
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    kernel_s = timer.run(lambda: step("fp64"), args.steps)   # the headline: exactly `steps` steps
    wall = time.perf_counter() - t_start
    # FP32-products form: blocks alternate with FP64 after the same warm-up
    s32a = timer.run(lambda: step("fp32"), args.steps)
    s64b = timer.run(lambda: step("fp64"), args.steps)
    s32b = timer.run(lambda: step("fp32"), args.steps)
    v64 = form.assemble_tensor_device(A, precision="fp64", symmetric=True).clone()
    v32 = form.assemble_tensor_device(A, precision="fp32", symmetric=True).clone()
    err32 = float(torch.linalg.norm(v32 - v64) / torch.linalg.norm(v64))
    del v32
    torch.cuda.synchronize()

    algorithmic = BYTES_PER_ELEMENT_C2 * E
    achieved_gbs = algorithmic / kernel_s / 1e9
    # measured HBM bytes of one assembly (FETCH_SIZE + WRITE_SIZE of the
    # payload pass and the gather, calibrated on a copy of known size), from
    # the committed rocprofv3 summary of this workload
    traffic = None
    traffic_file = os.path.join(ROOT, "profiles", "c2_p2_aniso_64_hbm_traffic.json")
    if args.n == 64 and os.path.exists(traffic_file):
        with open(traffic_file) as f:
            traffic = json.load(f)["per_assembly_bytes"]
    result = {
        "metric": METRIC,
        "value": E / kernel_s,
        "unit": "elements/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": kernel_s * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"P2 tet anisotropic diffusion, {args.n}^3 Kuhn cube jittered 0.2h, "
                               f"{E} tets, 4-pt quadrature, FP64",
                   "elements": E, "nnz": plan.nnz, "dofs": basis.N, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved_gbs / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "k_payload_h1 + k_gather_h1 (one assembly)",
                     "floor_bytes": FLOOR_BYTES_PER_ELEMENT_C2 * E,
                     "floor_frac": FLOOR_BYTES_PER_ELEMENT_C2 * E / kernel_s / 1e9 / HBM_PEAK_GBS},
        "fp32_products": {"ms_per_step": 0.5 * (s32a + s32b) * 1e3,
                          "speedup_vs_fp64": (kernel_s + s64b) / (s32a + s32b),
                          "rel_frobenius_vs_fp64": err32,
                          "timing": "alternating FP64/FP32 blocks of `steps` each after one warm-up"},
        "nnz_per_s": plan.nnz / kernel_s,
        "wall_ms_per_step": wall / args.steps * 1e3,
        "clock_ramp": {"seconds": args.ramp, "calls": ramp_calls,
                       "note": "untimed FP64+FP32 assemblies before the warm-up, until the GPU is at steady clocks"},
        "plan_setup_s": plan_s,
        "cold_first_asm_s": cold_asm_s,
        "autotune": autotune_summary(plan),
    }
    # warm end-to-end asm on the same workload: form tracing, the coefficient
    # upload, payload + gather, and the scipy CSR copied back to the host
    asm_times = []
    for _ in range(3):
        t1 = time.perf_counter()
        K = sk.asm(aniso, basis, diffusion=diffusion)
        asm_times.append(time.perf_counter() - t1)
    result["asm_end_to_end"] = {"ms": statistics.median(asm_times) * 1e3,
                                "elements_per_s": E / statistics.median(asm_times),
                                "includes": "form tracing, coefficient upload, kernels, CSR values to host"}
    del K
    return result, mesh, A_host, (form, A)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--ramp", type=float, default=0.3, help="seconds of untimed assemblies before the warm-up")
    ap.add_argument("--n", type=int, default=64, help="cells per cube edge")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the C1/C3/C4 and side-workload keys of the default run")
    ap.add_argument("--systems", type=int, default=256, help="batched workload: number of systems (0 = skip)")
    ap.add_argument("--cells", type=int, nargs=2, default=(16, 48), help="batched workload: cells per axis range")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="run one device-to-device copy of this size first (counter calibration)")
    ap.add_argument("--workload", default="c2", choices=("c2", "c1", "c3", "c4", "th", "stokes", "cut"),
                    help="c2 (default): the headline line with the other configurations as extra keys; "
                         "the others print their own line")
    ap.add_argument("--vector-mfma", action="store_true",
                    help="with --workload c3: run the 16-row layout of the vector-P2 MFMA gather (40 KiB spans) "
                         "instead of the autotuned one")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build's assembly path runs on a single MI355X; "
              "multi-GPU sharding is a design note only", file=sys.stderr)
        sys.exit(2)

    import torch
    from skfemntv import _native

    torch.cuda.set_device(0)
    torch.zeros(1, device="cuda")
    _native.lib()
    timer = Timer()
    if args.workload != "c2":
        print(json.dumps(other_workload(args.workload, args.steps, args.warmup, args.ramp, timer,
                                        vector_mfma=args.vector_mfma)), flush=True)
        return
    result, mesh, A_host, keep = headline(args, timer)
    if not args.no_configs:
        configs = {}
        for name in ("c1", "c3", "c4"):
            log(f"config {name} ...")
            w = build_workload(name)
            _, summary = measure(w, args.steps, args.warmup, 0.2, timer)
            summary["workload"] = w["desc"]
            summary["elements"], summary["nnz"] = w["E"], w["nnz"]
            summary["traffic"] = measured_traffic(name)
            if w["plan"] is not None:
                summary["autotune"] = autotune_summary(w["plan"])
            configs[name] = summary
            del w
            torch.cuda.empty_cache()
        result["configs"] = configs
        # side workloads of the next-step rows (cross blocks, composite
        # Stokes operator, cut cells), FP64, same timing
        side = {}
        for name in ("th", "stokes", "cut"):
            log(f"side workload {name} ...")
            w = build_workload(name)
            _, summary = measure(w, args.steps, args.warmup, 0.2, timer)
            summary["workload"] = w["desc"]
            summary["elements"], summary["nnz"] = w["E"], w["nnz"]
            summary["traffic"] = measured_traffic(name)
            side[name] = summary
            del w
            torch.cuda.empty_cache()
        result["side_workloads"] = side
    del keep
    torch.cuda.empty_cache()
    if args.systems > 0:
        log("batched assembly + PCG ...")
        result["batched_fsai"] = batched_fsai(args.systems, args.cells[0], args.cells[1], args.steps, args.warmup,
                                              timer)
    if not args.no_cpu_baseline:
        log("cpu baseline (C restatement of the reference's threaded assembler, full workload) ...")
        result["cpu_baseline"] = cpu_baseline(mesh, A_host)
        result["cpu_baseline"]["gpu_speedup"] = result["value"] / result["cpu_baseline"]["value"]
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
