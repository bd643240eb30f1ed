"""Runtime controls (python/skfemntv/runtime.py:32-64).

The reference's thread count selects how many CPU threads its assemblers
use.  In the MI355X build every assembly runs on the GPU, so the setting is
kept only for API compatibility: it is validated and stored, and has no
effect on results or on the device kernels.
"""

from __future__ import annotations

import os
from contextlib import contextmanager

BACKEND = "skfemntv-mi355x"
CAPABILITIES = frozenset({"native_backend", "gpu_bilinear_assembly", "gpu_linear_assembly", "gpu_hcurl_assembly",
                          "gpu_pattern_build", "deterministic_reduction"})

_threads = 1


def has_capability(name: str) -> bool:
    return name in CAPABILITIES


def available_num_threads() -> int:
    affinity = getattr(os, "sched_getaffinity", None)
    if affinity is not None:
        try:
            return max(1, len(affinity(0)))
        except OSError:
            pass
    return max(1, os.cpu_count() or 1)


def get_num_threads() -> int:
    return _threads


def set_num_threads(count: int) -> int:
    global _threads
    if isinstance(count, bool) or not isinstance(count, int) or count < 1:
        raise ValueError("thread count must be a positive integer")
    _threads = min(count, available_num_threads())
    return _threads


@contextmanager
def thread_limit(count: int):
    global _threads
    previous = _threads
    effective = set_num_threads(count)
    try:
        yield effective
    finally:
        _threads = previous
