"""Benchmark: FP64 P2-tetrahedral anisotropic-diffusion assembly on one MI355X.

Workload (one "step" = one full assembly of the CSR values):
  64^3 Kuhn-split unit cube (1,572,864 tets, 2,146,689 P2 DOFs), interior
  vertices jittered uniformly by +-0.2h (seed 0), per-element SPD tensor
  A = Q diag(lambda) Q^T with Q a Haar rotation and lambda log-uniform in
  [1e-2, 1e2] (seed 1), 4-point quadrature, form dot(mul(A, grad u), grad v).
Inputs (vertex coordinates, connectivity, A, gather plan) are resident in
HBM before timing; the timed region is K back-to-back assemblies.

Prints one JSON line.  `value` = assembled elements/s of the FP64 path.
`roofline` uses the algorithmic bytes of one assembly (825.7 B per element:
vertices, int32 connectivity, 72 B coefficient, int32 element->CSR map,
FP64 CSR values) over the kernel's mean duration measured with HIP events on
the launch stream.  `cpu_baseline` times the reference's own compiled C++
CrossBilinearAssembler (oracle/_ref) on a bounded 16^3 sample of the same
workload with all host cores.
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
BYTES_PER_ELEMENT_C2 = 825.7  # algorithmic bytes per tet of one FP64 assembly
# the same without the int32 element->CSR map (10 x 10 entries per tet): the "floor"
FLOOR_BYTES_PER_ELEMENT_C2 = BYTES_PER_ELEMENT_C2 - 4. * 10 * 10
# rocprofv3 FETCH_SIZE/WRITE_SIZE summaries of the other workloads at their default sizes
TRAFFIC_FILES = {"th": "th_divergence_block_64_hbm_traffic.json"}


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


def cpu_baseline(n_sample=16, repeats=7, warmup=2):
    """Reference C++ (compiled from its sources into oracle/_ref) on a
    bounded sample; falls back to the C restatement when _ref is absent."""
    import oracle
    from oracle import host, pipeline
    ax = np.linspace(0., 1., n_sample + 1)
    p, t = host.kuhn_mesh(ax, ax, ax)
    p = host.jitter_interior(p, 1. / n_sample, 0.2, 0)
    p2, t2 = host.p2_mesh(p, t)
    X, W = host.tet_rule_4()
    E = t.shape[1]
    A = spd_tensors(E, 1)
    A_eq = np.ascontiguousarray(np.broadcast_to(A[:, None], (E, 4, 3, 3)))
    cores = len(os.sched_getaffinity(0))
    ref = oracle.ref_module()
    (shape, grads, dx, x, det, tol, cond), _ = pipeline.basis_geometry(p2, t2, X, W, True)
    N, dofs = pipeline.dofs_of(t2, p2.shape[1], 10, 1)
    out = {}
    if ref is not None:
        coef = np.ascontiguousarray(A_eq[:, :, None, :, None, :])
        asm_ = ref.CrossBilinearAssembler(dofs, dofs, shape, shape, dx, grads, grads)
        for threads in (cores, 1):
            ref.set_num_threads(threads)
            for _ in range(warmup):
                asm_.assemble(coef, "gradient", "gradient", 0)
            ts = []
            for _ in range(repeats):
                t0 = time.perf_counter()
                asm_.assemble(coef, "gradient", "gradient", 0)
                ts.append(time.perf_counter() - t0)
            out[threads] = statistics.median(ts)
        ref.set_num_threads(1)
        kind = "reference"
    else:
        ip, ix, sc = oracle.pattern(dofs.reshape(E, -1), nrows=N)
        colour, _ = oracle.colouring(dofs.reshape(E, -1), N)
        for threads in (cores, 1):
            for _ in range(warmup):
                oracle.assemble_tensor(grads, dx, A_eq, sc, len(ix), threads=threads, colour=colour)
            ts = []
            for _ in range(repeats):
                t0 = time.perf_counter()
                oracle.assemble_tensor(grads, dx, A_eq, sc, len(ix), threads=threads, colour=colour)
                ts.append(time.perf_counter() - t0)
            out[threads] = statistics.median(ts)
        kind = "port"
    cpu_model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    cpu_model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {
        "value": E / out[cores],
        "unit": "elements/s",
        "cores": cores,
        "kind": kind,
        "sample": (f"P2 anisotropic {n_sample}^3 Kuhn cube ({E} tets), 4-pt rule, warm median of {repeats} "
                   f"assemble() calls; CrossBilinearAssembler; 1-thread {E / out[1]:.4g} elements/s; {cpu_model}"),
    }


def batched_fsai(systems, lo, hi, steps, warmup, L, ev0, ev1, stream):
    """256 P1 anisotropic systems on uneven boxes: batched assembly (one
    HIP-graph replay per step), then FP64 PCG to ||b - Ax||/||b|| <= 1e-8
    with Jacobi and with the pretrained learned FSAI."""
    import torch
    from skfemntv import _native
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
    _native.check(L.skfa_event_record(ev0, stream))
    for _ in range(steps):
        batch.assemble(graph=True)
    _native.check(L.skfa_event_record(ev1, stream))
    ms = ctypes.c_float()
    _native.check(L.skfa_event_elapsed_ms(ev0, ev1, ctypes.byref(ms)))
    asm_s = ms.value / 1e3 / steps
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
    res_j = relative_residuals(A, b, rj.x, batch.sys_ptr)
    res_l = relative_residuals(A, b, rl.x, batch.sys_ptr)
    return {
        "workload": f"{systems} P1 anisotropic systems, cells/axis iid U{{{lo}..{hi}}} (seed 5), jitter 0.2h, "
                    f"per-element SPD A (seed 5000+s), f=1, boundary condensed",
        "elements": batch.nelements, "unknowns": A.shape[0], "nnz": A.nnz,
        "assembly_ms_per_batch": asm_s * 1e3, "assembly_elements_per_s": batch.nelements / asm_s,
        "assembly": "payload + row-gather + condensation gather + load vector, replayed as one HIP graph",
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
    }


def timed(fn, steps, warmup, L, ev0, ev1, stream):
    """Mean seconds per call of fn over `steps` calls, HIP events on the launch stream."""
    import torch
    from skfemntv import _native
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    _native.check(L.skfa_event_record(ev0, stream))
    for _ in range(steps):
        fn()
    _native.check(L.skfa_event_record(ev1, stream))
    ms = ctypes.c_float()
    _native.check(L.skfa_event_elapsed_ms(ev0, ev1, ctypes.byref(ms)))
    return ms.value / 1e3 / steps


def other_workload(name, steps, warmup, L, ev0, ev1, stream):
    """The other assembly configurations at full size (one JSON line each):
    c1 P1 Laplace + load 32^3, c3 vector-P2 elasticity 48^3, c4 N1 curl-curl +
    kappa mass 96^3, th the Taylor-Hood divergence block (vector P2 x P1) on
    the 64^3 jittered cube, stokes the Taylor-Hood Stokes operator on a
    composite basis at 48^3.  Algorithmic bytes per element: 106.7, 3255, 356,
    for th 24 V/E + 56 + 480 + 8 nnz/E (vertices, both connectivities, int32
    element->CSR map, values), for stokes the sum of its three blocks, for
    cut (a spherical band of cut cells, 12 points each) the per-point tables,
    cell offsets and connectivity, and the values."""
    import torch
    import skfemntv as sk
    from skfemntv._assembly import NativeBilinearForm, NativeLinearForm, upload
    from skfemntv.hcurl import AffineTetN1Basis, TetN1Assembler
    t0 = time.perf_counter()
    if name == "c1":
        ax = np.linspace(0., 1., 33)
        mesh = sk.MeshTet.init_tensor(ax, ax, ax)
        basis = sk.Basis(mesh, sk.ElementVector(sk.ElementTetP1(), dim=1))
        E = mesh.nelements
        form, load = NativeBilinearForm(basis), NativeLinearForm(basis)
        one = upload(np.ones(E))
        f = upload(np.array([1.]))

        def step(precision="fp64"):
            form.assemble_device(gradient=one, precision=precision)
            load.assemble_device(f, False, None, False)
        bpe, desc, nnz = 106.7, f"P1 Laplace ddot(grad u, grad v) + load f=1, 32^3 Kuhn cube, {E} tets, FP64", form.plan.nnz
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
        # nodes) + int32 element->CSR map (4 x 30 entries) + FP64 values
        bpe = (24. * V + 4. * 14 * E + 4. * 4 * 30 * E + 8. * nnz) / E
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
        # element->CSR map + FP64 values (velocity 30 x 30, coupling 30 x 4)
        bpe = (3 * 24. * V + (40. + 8. + 4. * 900) * E + 8. * nnz_a
               + 2 * ((56. + 4. * 120) * E + 8. * nnz_b)) / E
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
    setup = time.perf_counter() - t0
    log(f"{name}: setup {setup:.1f}s")
    sec = timed(step, steps, warmup, L, ev0, ev1, stream)
    out = {"metric": METRIC, "value": E / sec, "unit": "elements/s", "n_gpus": 1, "steps": steps, "warmup": warmup,
           "ms_per_step": sec * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f64", "data": "synthetic",
           "config": {"workload": desc, "elements": E, "nnz": nnz, "parallelism": "single GPU"},
           "roofline": {"bound": "hbm", "achieved": bpe * E / sec / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": bpe * E / sec / 1e9 / HBM_PEAK_GBS, "traffic": None},
           "nnz_per_s": nnz / sec, "setup_s": setup}
    # measured HBM bytes of one launch (FETCH_SIZE + WRITE_SIZE, calibrated
    # on a copy of known size), from the committed rocprofv3 summary
    traffic_file = os.path.join(ROOT, "profiles", f"{TRAFFIC_FILES.get(name, '')}")
    if name in TRAFFIC_FILES and os.path.exists(traffic_file):
        with open(traffic_file) as f:
            out["roofline"]["traffic"] = json.load(f)["per_assembly_bytes"]
    if name in ("c1", "c3"):
        sec32 = timed(lambda: step("fp32"), steps, warmup, L, ev0, ev1, stream)
        step("fp64")
        v64 = form._values.clone()
        step("fp32")
        v32 = form._values.clone()
        err32 = float(torch.linalg.norm(v32 - v64) / torch.linalg.norm(v64))
        out["fp32_products"] = {"ms_per_step": sec32 * 1e3, "speedup_vs_fp64": sec / sec32,
                                "rel_frobenius_vs_fp64": err32}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    # the GPU reaches steady clocks after ~0.1 s of back-to-back assemblies:
    # measured on C2, 3 warmup steps give 0.38 ms/step, 30 give 0.33, 200 give 0.32
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--n", type=int, default=64, help="cells per cube edge")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--systems", type=int, default=256, help="batched workload: number of systems (0 = skip)")
    ap.add_argument("--cells", type=int, nargs=2, default=(16, 48), help="batched workload: cells per axis range")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="run one device-to-device copy of this size first (counter calibration)")
    ap.add_argument("--workload", default="c2", choices=("c2", "c1", "c3", "c4", "th", "stokes", "cut"),
                    help="c2 (default): the headline line; c1/c3/c4: the other assembly configurations")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build's assembly path runs on a single MI355X; "
              "multi-GPU sharding is a design note only", file=sys.stderr)
        sys.exit(2)

    import torch
    import skfemntv as sk
    from skfemntv import _native
    from skfemntv._assembly import NativeCrossBilinearForm, upload

    torch.cuda.set_device(0)
    if args.workload != "c2":
        L = _native.lib()
        stream = _native.stream()
        ev0, ev1 = ctypes.c_void_p(), ctypes.c_void_p()
        _native.check(L.skfa_event_create(ctypes.byref(ev0)))
        _native.check(L.skfa_event_create(ctypes.byref(ev1)))
        print(json.dumps(other_workload(args.workload, args.steps, args.warmup, L, ev0, ev1, stream)), flush=True)
        return
    t0 = time.perf_counter()
    mesh = jittered_p2_cube(args.n, 0.2, 0)
    E = mesh.nelements
    log(f"mesh {E} tets, {mesh.p.shape[1]} nodes ({time.perf_counter() - t0:.1f}s)")
    X = np.array([[.5854101966249685, .1381966011250105, .1381966011250105, .1381966011250105],
                  [.1381966011250105, .5854101966249685, .1381966011250105, .1381966011250105],
                  [.1381966011250105, .1381966011250105, .5854101966249685, .1381966011250105]])
    W = np.full(4, 1. / 24.)
    t1 = time.perf_counter()
    basis = sk.Basis(mesh, sk.ElementVector(sk.ElementTetP2(), dim=1), quadrature=(X, W))
    log(f"basis + geometry check {time.perf_counter() - t1:.1f}s")
    t1 = time.perf_counter()
    form = NativeCrossBilinearForm(basis, basis)
    torch.cuda.synchronize()
    setup_plan = time.perf_counter() - t1
    log(f"plan (pattern + gather map on GPU) {setup_plan:.2f}s nnz={form.plan.nnz}")
    A_host = spd_tensors(E, 1)
    A = upload(A_host)
    if args.copy_mib:
        src = torch.ones(args.copy_mib * (1 << 20) // 8, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    L = _native.lib()
    stream = _native.stream()
    ev0, ev1 = ctypes.c_void_p(), ctypes.c_void_p()
    _native.check(L.skfa_event_create(ctypes.byref(ev0)))
    _native.check(L.skfa_event_create(ctypes.byref(ev1)))

    def step(precision="fp64"):
        form.assemble_tensor_device(A, precision=precision, symmetric=True)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    _native.check(L.skfa_event_record(ev0, stream))
    for _ in range(args.steps):
        step()
    _native.check(L.skfa_event_record(ev1, stream))
    ms = ctypes.c_float()
    _native.check(L.skfa_event_elapsed_ms(ev0, ev1, ctypes.byref(ms)))
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_start
    kernel_s = ms.value / 1e3 / args.steps
    # FP32-products form on the same inputs
    for _ in range(args.warmup):
        step("fp32")
    _native.check(L.skfa_event_record(ev0, stream))
    for _ in range(args.steps):
        step("fp32")
    _native.check(L.skfa_event_record(ev1, stream))
    _native.check(L.skfa_event_elapsed_ms(ev0, ev1, ctypes.byref(ms)))
    kernel32_s = ms.value / 1e3 / args.steps
    v64 = form.assemble_tensor_device(A, precision="fp64", symmetric=True).clone()
    v32 = form.assemble_tensor_device(A, precision="fp32", symmetric=True).clone()
    err32 = float(torch.linalg.norm(v32 - v64) / torch.linalg.norm(v64))
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
                   "elements": E, "nnz": form.plan.nnz, "dofs": basis.N, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved_gbs / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "k_payload_h1 + k_gather_h1 (one assembly)",
                     "floor_bytes": FLOOR_BYTES_PER_ELEMENT_C2 * E,
                     "floor_frac": FLOOR_BYTES_PER_ELEMENT_C2 * E / kernel_s / 1e9 / HBM_PEAK_GBS},
        "fp32_products": {"ms_per_step": kernel32_s * 1e3, "speedup_vs_fp64": kernel_s / kernel32_s,
                          "rel_frobenius_vs_fp64": err32},
        "nnz_per_s": form.plan.nnz / kernel_s,
        "wall_ms_per_step": wall / args.steps * 1e3,
        "plan_setup_s": setup_plan,
    }
    # End-to-end asm on the same workload: form tracing, the coefficient in
    # the recommended (3,3,E,Q) layout (a broadcast view, uploaded per call),
    # payload + gather, and the scipy CSR copied back to the host.  Median of 3.
    from skfemntv.helpers import dot, grad, mul

    @sk.BilinearForm
    def aniso(u, v, w):
        return dot(mul(w.diffusion, grad(u)), grad(v))

    diffusion = np.broadcast_to(np.moveaxis(A_host, 0, -1)[..., None], (3, 3, E, 4))
    asm_times = []
    for _ in range(3):
        t1 = time.perf_counter()
        K = sk.asm(aniso, basis, diffusion=diffusion)
        asm_times.append(time.perf_counter() - t1)
    result["asm_end_to_end"] = {"ms": statistics.median(asm_times) * 1e3,
                                "elements_per_s": E / statistics.median(asm_times),
                                "includes": "form tracing, coefficient upload, kernels, CSR values to host"}
    del K
    if args.systems > 0:
        log("batched assembly + PCG ...")
        result["batched_fsai"] = batched_fsai(args.systems, args.cells[0], args.cells[1], args.steps, args.warmup,
                                              L, ev0, ev1, stream)
    if not args.no_cpu_baseline:
        log("cpu baseline ...")
        result["cpu_baseline"] = cpu_baseline()
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
