"""Launch-shape autotuner of the row-gather kernels.

The gather's speed depends on choices that no formula settles well across
meshes and GPUs: the LDS budget of a wave's CSR span (it fixes how many
consecutive node rows a 64-lane wave owns and how many lanes share a row),
how the lanes of one row share its incidences (in turn, or concurrently
with the LDS atomic add resolving shared addresses), the formulation of P2
element rows and vector-P2 node blocks (quadrature tables, or the
barycentric structure of the reference gradients), the element-table gather
of 4-node forms, the 16-row layout of the vector-P2 MFMA gather and the
chunk of consecutive waves dealt to one XCD.

On the first assembly of a plan with at least ``MIN_ELEMENTS`` elements
(outside HIP-graph capture), the tuner times every candidate on the live
inputs and keeps the fastest; the choice is cached per mesh signature
(elements, local nodes, node rows, components), so later plans of the same
shape start from it without timing.  Every candidate computes the same
matrix: FP64 results differ by rounding only (the oracle tests hold each
formulation to 1e-12), and the span budget and XCD chunk do not change the
bits at all.

Timing: a 0.15 s clock ramp, then per candidate two warm-up calls and three
blocks of four calls timed with HIP events on the launch stream; the
candidate's score is its median block.  The reference has no counterpart:
its only shape choice is the static per-colour thread partition
(cpp/include/native_fem/parallel.hpp:21-41).
"""

from __future__ import annotations

import ctypes
import itertools
import time
from dataclasses import dataclass, field

from . import _native

MIN_ELEMENTS = 100_000
_enabled = True
_CACHE: dict = {}
LAST: dict = {}   # signature -> TuningRecord of the last tuning run (reporting)


@dataclass
class TuningRecord:
    budget: int
    settings: dict
    seconds_per_call: float
    candidates: int
    tuning_seconds: float
    table: list = field(default_factory=list)
    row_lanes: int = 0


def set_enabled(flag: bool) -> None:
    """Switch the autotuner on or off (process-wide)."""
    global _enabled
    _enabled = bool(flag)


def enabled() -> bool:
    return _enabled


def signature(plan):
    return (plan.E, plan.n, plan.nnodes, plan.components)


def cached_layout(E, n, nnodes, components, default):
    """(LDS budget, lanes per row) a new plan of this shape should start from."""
    rec = _CACHE.get((E, n, nnodes, components))
    return (rec.budget, rec.row_lanes) if rec is not None else (default, 0)


# span budget of the 16-row layout of the vector-P2 MFMA gather (its best
# of 24-48 KiB on the 48^3 elasticity workload; the layout measured 3x
# slower than the default one there, LDS-bound at about 3 waves per CU)
MFMA_BUDGETS = (40 * 1024,)


def candidates(plan, vector_mfma=False):
    """(layouts, setting dicts, chunks) to try for a plan's kernels; a
    layout is (LDS budget, lanes per row), a setting dict applies to the
    default layouts (the MFMA gather has no such settings)."""
    n, C = plan.n, plan.components
    chunks = (8, 16, 32)
    if C == 3:
        budgets = (12 * 1024, 16 * 1024, 20 * 1024, 24 * 1024)
        forms = (0, 1) if n == 10 else (0,)
        variants = [{"row_split": s, "row_form": f} for s, f in itertools.product((0, 3), forms)]
    elif n == 10:
        budgets = (8 * 1024, 10 * 1024, 12 * 1024, 14 * 1024, 16 * 1024)
        variants = [{"row_split": s, "row_form": f} for s, f in itertools.product((0, 3), (0, 1))]
    elif n == 4:
        budgets = (4 * 1024, 8 * 1024, 12 * 1024, 16 * 1024)
        variants = [{"element_table": t} for t in (-1, 0, 1)]
    else:   # edge plans (N1) and other element types
        budgets = (8 * 1024, 12 * 1024, 16 * 1024)
        variants = [{"row_split": s} for s in (0, 3)]
    layouts = [(b, 0) for b in budgets]
    if vector_mfma and C == 3 and n == 10:
        layouts += [(b, 4) for b in MFMA_BUDGETS]
    return layouts, variants, chunks


class _Timer:
    def __init__(self):
        L = _native.lib()
        self.L = L
        self.ev = (ctypes.c_void_p(), ctypes.c_void_p())
        for e in self.ev:
            _native.check(L.skfa_event_create(ctypes.byref(e)))

    def per_call(self, run, calls):
        L = self.L
        s = _native.stream()
        _native.check(L.skfa_event_record(self.ev[0], s))
        for _ in range(calls):
            run()
        _native.check(L.skfa_event_record(self.ev[1], s))
        ms = ctypes.c_float()
        _native.check(L.skfa_event_elapsed_ms(self.ev[0], self.ev[1], ctypes.byref(ms)))
        return ms.value * 1e-3 / calls

    def close(self):
        for e in self.ev:
            self.L.skfa_event_destroy(e)


def _score(timer, run):
    run()
    run()
    blocks = sorted(timer.per_call(run, 4) for _ in range(3))
    return blocks[1]


def maybe_tune(plan, run, vector_mfma=False):
    """Tune `plan` on its first qualifying assembly.  `run()` performs one
    assembly with the plan's current layout and settings; ``vector_mfma``
    says that the form can run on the 16-row vector-P2 MFMA layout."""
    if getattr(plan, "_tuned", False) or not _enabled or plan.E < MIN_ELEMENTS:
        return None
    import torch
    if torch.cuda.is_current_stream_capturing():
        return None
    plan._tuned = True
    sig = signature(plan)
    rec = _CACHE.get(sig)
    if rec is not None:
        apply(plan, rec)
        return rec
    rec = tune(plan, run, vector_mfma)
    _CACHE[sig] = rec
    LAST[sig] = rec
    return rec


def apply(plan, rec):
    if plan.lds_budget != rec.budget or plan.row_lanes != rec.row_lanes:
        plan.rebuild(rec.budget, row_lanes=rec.row_lanes)
    plan.set_tuning(**rec.settings)


def tune(plan, run, vector_mfma=False):
    """Time every candidate and leave `plan` at the fastest."""
    import torch
    t0 = time.perf_counter()
    layouts, variants, chunks = candidates(plan, vector_mfma)
    timer = _Timer()
    try:
        ramp_end = time.perf_counter() + 0.15
        while time.perf_counter() < ramp_end:
            for _ in range(4):
                run()
            torch.cuda.synchronize()
        table = []
        # layout x formulation at the default chunk, then the chunk
        for budget, lanes in layouts:
            plan.rebuild(budget, row_lanes=lanes)
            for v in (variants if lanes == 0 else ({},)):
                plan.set_tuning(**v)
                table.append(((budget, lanes), dict(v), _score(timer, run)))
        (budget, lanes), best_v, best_t = min(table, key=lambda r: r[2])
        plan.rebuild(budget, row_lanes=lanes)
        plan.set_tuning(**best_v)
        best_chunk = plan.tuning["xcd_chunk"]
        for c in chunks:
            if c == best_chunk:
                continue
            plan.set_tuning(xcd_chunk=c)
            t = _score(timer, run)
            table.append(((budget, lanes), {**best_v, "xcd_chunk": c}, t))
            if t < best_t:
                best_t, best_chunk = t, c
        settings = {**best_v, "xcd_chunk": best_chunk}
        plan.set_tuning(**settings)
    finally:
        timer.close()
    return TuningRecord(budget=budget, settings=settings, seconds_per_call=best_t, candidates=len(table),
                        tuning_seconds=time.perf_counter() - t0,
                        table=[(b, s, t) for b, s, t in table], row_lanes=lanes)
