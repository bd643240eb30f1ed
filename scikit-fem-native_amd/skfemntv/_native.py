"""ctypes binding of libskfem_amd.so (the MI355X C-ABI, include/skfem_amd.h).

The library is mandatory: every assembly goes through it and there is no
CPU fallback.  When the library or a HIP device is missing, the first call
that needs it raises ``NativeUnavailable``.
"""

from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("SKFEMNTV_LIBRARY") or os.path.join(_HERE, "_lib", "libskfem_amd.so")

# Exported symbols (kept in sync with include/skfem_amd.h by a test).
SYMBOLS = (
    "skfa_last_error", "skfa_abi_version", "skfa_device_name",
    "skfa_plan_create", "skfa_plan_create_budget", "skfa_plan_create_ex", "skfa_plan_create_layout", "skfa_plan_set_tuning", "skfa_plan_destroy", "skfa_plan_info", "skfa_plan_export_csr", "skfa_plan_export_map",
    "skfa_cross_plan_create", "skfa_cross_plan_destroy", "skfa_cross_plan_info", "skfa_cross_plan_export_csr",
    "skfa_assemble_cross", "skfa_assemble_facet_linear", "skfa_assemble_facet_bilinear",
    "skfa_assemble_cut_bilinear", "skfa_assemble_cut_linear",
    "skfa_assemble_bilinear", "skfa_assemble_bilinear_noalloc", "skfa_assemble_bilinear_load", "skfa_assemble_linear", "skfa_tabulate_geometry",
    "skfa_edge_topology", "skfa_assemble_tet_n1",
    "skfa_event_create", "skfa_event_destroy", "skfa_event_record", "skfa_event_elapsed_ms",
    "skfa_stream_synchronize",
    "skfa_spmv", "skfa_csr_condense", "skfa_csr_lower", "skfa_csr_transpose", "skfa_csr_transpose_lower", "skfa_gather_values",
    "skfa_gather_values_f32", "skfa_diag_inverse", "skfa_sys_dot", "skfa_pcg_solve",
    "skfa_fsai_features", "skfa_fsai_factor", "skfa_fsai_kaporin", "skfa_mlp_pack", "skfa_mlp_forward",
    "skfa_mlp_workspace_bytes", "skfa_mlp_backward", "skfa_mlp_adam",
)

FSAI_FEATURES = 16
FSAI_PARAMS = 18817
FSAI_FRAG = 34816


FLAG_SYMMETRIC_TENSOR = 1
FLAG_VECTOR3_DIAGONAL = 2
FLAG_BARYCENTRIC_TABLE = 4
TUNE_XCD_CHUNK = 1
TUNE_LDS_ATOMIC = 2
TUNE_PLACEMENT = 3
TUNE_STREAMING = 4
TUNE_ELEMENT_TABLE = 5
TUNE_ROW_SPLIT = 6
TUNE_MFMA = 7
TUNE_ROW_FORM = 8


class NativeUnavailable(RuntimeError):
    """The MI355X library or a HIP device is not available."""


class NativeError(RuntimeError):
    """A C-ABI call returned an error status."""


_lib = None

P = ctypes.c_void_p
I32 = ctypes.c_int32
I64 = ctypes.c_int64


def _declare(lib):
    sig = {
        "skfa_last_error": (ctypes.c_char_p, []),
        "skfa_abi_version": (ctypes.c_int, []),
        "skfa_device_name": (ctypes.c_int, [ctypes.c_char_p, I32]),
        "skfa_plan_create": (ctypes.c_int, [P, I64, I32, I64, I32, P, ctypes.POINTER(P)]),
        "skfa_plan_destroy": (ctypes.c_int, [P]),
        "skfa_plan_create_budget": (ctypes.c_int, [P, I64, I32, I64, I32, I32, P, ctypes.POINTER(P)]),
        "skfa_plan_create_ex": (ctypes.c_int, [P, I64, I32, I64, I32, I32, I64, P, ctypes.POINTER(P)]),
        "skfa_plan_create_layout": (ctypes.c_int, [P, I64, I32, I64, I32, I32, I64, I32, P, ctypes.POINTER(P)]),
        "skfa_plan_set_tuning": (ctypes.c_int, [P, I32, I32]),
        "skfa_plan_info": (ctypes.c_int, [P, ctypes.POINTER(I64)]),
        "skfa_plan_export_csr": (ctypes.c_int, [P, P, P, P]),
        "skfa_plan_export_map": (ctypes.c_int, [P, P, P, P]),
        "skfa_cross_plan_create": (ctypes.c_int, [P, I32, I64, I32, P, I32, I64, I32, I64, P, ctypes.POINTER(P)]),
        "skfa_cross_plan_destroy": (ctypes.c_int, [P]),
        "skfa_cross_plan_info": (ctypes.c_int, [P, ctypes.POINTER(I64)]),
        "skfa_cross_plan_export_csr": (ctypes.c_int, [P, P, P, P]),
        "skfa_assemble_cross": (ctypes.c_int, [P, P, P, P, I32, P, I32, P, I32, I32, I32, P, I64, I64, P, P]),
        "skfa_assemble_facet_linear": (ctypes.c_int, [P, P, P, I32, P, P, I32, P]),
        "skfa_assemble_facet_bilinear": (ctypes.c_int, [P, P, P, P, I32, P, P, P]),
        "skfa_assemble_cut_bilinear": (ctypes.c_int, [P, P, P, P, P, I32, P, P, I32, I32, P, P, I32, I32, P,
                                                      I32, P, P]),
        "skfa_assemble_cut_linear": (ctypes.c_int, [P, P, P, P, P, I32, P, I32, P, I32, P, P]),
        "skfa_assemble_bilinear": (ctypes.c_int, [P, P, I64, P, P, I32, P, I32, P, P, I32, I32,
                                                  P, P, P, P, P, I32, I32, I32, P, P]),
        "skfa_assemble_bilinear_noalloc": (ctypes.c_int, [P, P, I64, P, P, I32, P, I32, P, P, I32, I32,
                                                          P, P, P, P, P, I32, I32, I32, P, P]),
        "skfa_assemble_bilinear_load": (ctypes.c_int, [P, P, I64, P, P, I32, P, I32, P, P, I32, I32,
                                                       P, P, P, P, P, I32, I32, I32, P, P, P, P, P, P]),
        "skfa_assemble_linear": (ctypes.c_int, [P, P, P, P, I32, P, I32, P, I32, P, I32, P, P]),
        "skfa_tabulate_geometry": (ctypes.c_int, [P, P, P, I32, I64, P, I32, I32, P, P, P, P, P, P, P, P]),
        "skfa_edge_topology": (ctypes.c_int, [P, I64, I64, P, P, P, P, ctypes.POINTER(I64), ctypes.POINTER(I64), P]),
        "skfa_assemble_tet_n1": (ctypes.c_int, [P, P, P, P, I32, P, P, I32, I32, I32, P, P]),
        "skfa_event_create": (ctypes.c_int, [ctypes.POINTER(P)]),
        "skfa_event_destroy": (ctypes.c_int, [P]),
        "skfa_event_record": (ctypes.c_int, [P, P]),
        "skfa_event_elapsed_ms": (ctypes.c_int, [P, P, ctypes.POINTER(ctypes.c_float)]),
        "skfa_stream_synchronize": (ctypes.c_int, [P]),
        "skfa_spmv": (ctypes.c_int, [P, P, P, P, P, I64, I32, P]),
        "skfa_csr_condense": (ctypes.c_int, [P, P, P, I64, P, P, P, P, P, ctypes.POINTER(I64),
                                             ctypes.POINTER(I64), P]),
        "skfa_csr_lower": (ctypes.c_int, [P, P, I64, P, P, P, ctypes.POINTER(I64), P]),
        "skfa_csr_transpose": (ctypes.c_int, [P, P, I64, I64, P, P, P, P]),
        "skfa_csr_transpose_lower": (ctypes.c_int, [P, P, I64, P, P, P, P, P, P]),
        "skfa_gather_values": (ctypes.c_int, [P, P, I64, P, P]),
        "skfa_gather_values_f32": (ctypes.c_int, [P, P, I64, P, P]),
        "skfa_diag_inverse": (ctypes.c_int, [P, P, P, I64, P, P]),
        "skfa_sys_dot": (ctypes.c_int, [P, P, P, I32, P, P, P]),
        "skfa_pcg_solve": (ctypes.c_int, [P, P, P, I64, P, I32, P, P, I32, P, P, P, P, P, P, P,
                                          ctypes.c_double, I32, I32, I32, P, P, P, P, ctypes.POINTER(I32), P]),
        "skfa_fsai_features": (ctypes.c_int, [P, P, P, I64, P, P, P, P, P, P, P]),
        "skfa_fsai_factor": (ctypes.c_int, [P, P, P, P, I64, P, P, P, P]),
        "skfa_fsai_kaporin": (ctypes.c_int, [P, P, P, I64, P, I32, P, P, P, P, ctypes.c_double, P, P, P, P, P]),
        "skfa_mlp_pack": (ctypes.c_int, [P, P, P]),
        "skfa_mlp_forward": (ctypes.c_int, [P, P, P, I64, P, P]),
        "skfa_mlp_workspace_bytes": (ctypes.c_size_t, [I64]),
        "skfa_mlp_backward": (ctypes.c_int, [P, P, P, P, I64, P, ctypes.c_size_t, P, P]),
        "skfa_mlp_adam": (ctypes.c_int, [P, P, P, P, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                         ctypes.c_float, I32, ctypes.c_float, P, P]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args


def load_library():
    """Load libskfem_amd.so (no device needed)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise NativeUnavailable(
                f"{LIB_PATH} is not built; run `python -c 'import __graft_entry__ as g; g.build()'` "
                "or `make -C scikit-fem-native_amd/csrc`")
        lib = ctypes.CDLL(LIB_PATH)
        _declare(lib)
        _lib = lib
    return _lib


_device_checked = False


def lib():
    """The library, after checking that a HIP device is usable."""
    global _device_checked
    L = load_library()
    if not _device_checked:
        import torch
        if not torch.cuda.is_available():
            raise NativeUnavailable("skfemntv (MI355X build) requires a HIP device; none is visible")
        _device_checked = True
    return L


def check(rc):
    if rc != 0:
        msg = load_library().skfa_last_error().decode(errors="replace")
        if rc == -1:
            raise ValueError(msg)
        raise NativeError(f"libskfem_amd status {rc}: {msg}")
    return rc


def ptr(t):
    """Raw device pointer of a torch tensor (or None -> NULL)."""
    if t is None:
        return None
    return ctypes.c_void_p(t.data_ptr())


def stream():
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def device():
    import torch
    return torch.device("cuda", torch.cuda.current_device())
