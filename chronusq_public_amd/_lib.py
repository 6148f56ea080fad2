"""ctypes binding of libcqgpu.so (the C ABI in include/cqgpu.h).

The library is loaded from the package directory; importing this module never falls back
to any other implementation -- a missing library is an ImportError-style failure.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcqgpu.so")

CQ_OK = 0
CQ_ERR_INVALID = -1
CQ_ERR_NOMEM = -2
CQ_ERR_DEVICE = -3
CQ_COULOMB = 0
CQ_EXCHANGE = 1
CQ_PREC_FP64 = 0
CQ_PREC_MIXED = 1

# every symbol include/cqgpu.h declares
EXPORTED = [
    "cq_ctx_create", "cq_ctx_create_batch", "cq_ctx_destroy", "cq_last_error", "cq_set_schwarz_threshold",
    "cq_get_schwarz_threshold", "cq_set_precision_mode", "cq_get_dims", "cq_get_schwarz",
    "cq_contract_direct", "cq_contract_direct_device", "cq_contract_direct_device_noalloc", "cq_get_quartet_counts", "cq_get_precision_counts",
    "cq_num_class_pairs", "cq_class_pair_info", "cq_debug_quartets", "cq_set_profiling",
    "cq_get_kernel_times", "cq_get_screen_times", "cq_set_screen_fusion", "cq_get_work_stats", "cq_one_electron", "cq_one_electron_batch", "cq_multipole", "cq_set_launch_streams", "cq_set_lds_digestion",
    "cq_solid_harmonics", "cq_set_class_digestion", "cq_get_class_digestion",
    "cq_set_class_stream", "cq_set_class_chunk_group", "cq_set_task_budget", "cq_get_task_plan", "cq_set_class_tile", "cq_get_class_tile",
    # learned initial guess (include/cqgpu_guess.h)
    "cq_mlp_last_error", "cq_mlp_gemm_bf16", "cq_mlp_relu_backward", "cq_mlp_column_sum", "cq_mlp_to_bf16",
    "cq_mlp_mse", "cq_mlp_adam",
]

_lib = None


class CQError(RuntimeError):
    """Raised for a non-zero return code (the reference raises std::runtime_error / CErr)."""


def load():
    global _lib
    if _lib is not None:
        return _lib
    # One HIP runtime per process: torch ships its own libamdhip64 (same SONAME,
    # libamdhip64.so.7).  Loading torch first makes the dynamic linker bind this library to
    # that runtime, so device pointers and streams from torch are valid here.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    if not os.path.exists(LIB_PATH):
        raise OSError(f"{LIB_PATH} is not built; run __graft_entry__.build() or "
                      "python -m chronusq_public_amd.build")
    L = ctypes.CDLL(LIB_PATH)
    P = ctypes.POINTER
    dp, ip, vp = P(ctypes.c_double), P(ctypes.c_int), ctypes.c_void_p
    L.cq_ctx_create.argtypes = [P(vp), ctypes.c_int, ip, ip, ip, dp, dp, dp]
    L.cq_ctx_create_batch.argtypes = [P(vp), ctypes.c_int, ip, ip, ip, dp, dp, dp, ip]
    L.cq_ctx_destroy.argtypes = [vp]
    L.cq_last_error.argtypes = [vp]
    L.cq_last_error.restype = ctypes.c_char_p
    L.cq_set_schwarz_threshold.argtypes = [vp, ctypes.c_double]
    L.cq_get_schwarz_threshold.argtypes = [vp, dp]
    L.cq_set_precision_mode.argtypes = [vp, ctypes.c_int, ctypes.c_double]
    L.cq_get_dims.argtypes = [vp, ip, ip, ip, ip, ip]
    L.cq_get_schwarz.argtypes = [vp, dp]
    L.cq_contract_direct.argtypes = [vp, ctypes.c_int, P(dp), P(dp), ip, dp, ctypes.c_int, ctypes.c_int]
    L.cq_contract_direct_device.argtypes = [vp, ctypes.c_int, P(vp), P(vp), ip, dp, ctypes.c_int,
                                            ctypes.c_int, vp]
    L.cq_contract_direct_device_noalloc.argtypes = [vp, ctypes.c_int, P(vp), P(vp), ip, dp, ctypes.c_int,
                                                    ctypes.c_int, vp]
    L.cq_get_quartet_counts.argtypes = [vp, P(ctypes.c_longlong), ctypes.c_int, ip]
    L.cq_get_precision_counts.argtypes = [vp, P(ctypes.c_longlong), P(ctypes.c_longlong)]
    L.cq_num_class_pairs.argtypes = []
    L.cq_class_pair_info.argtypes = [ctypes.c_int, ip, ip, ip, ip]
    L.cq_debug_quartets.argtypes = [vp, ctypes.c_int, ip, dp]
    L.cq_set_profiling.argtypes = [vp, ctypes.c_int]
    L.cq_set_screen_fusion.argtypes = [vp, ctypes.c_int]
    L.cq_set_launch_streams.argtypes = [vp, ctypes.c_int]
    L.cq_set_lds_digestion.argtypes = [vp, ctypes.c_int]
    L.cq_one_electron.argtypes = [vp, ctypes.c_int, dp, dp, vp, vp, vp, ctypes.c_int, vp]
    L.cq_one_electron_batch.argtypes = [vp, ctypes.c_int, dp, dp, ip, vp, vp, vp, ctypes.c_int, vp]
    L.cq_multipole.argtypes = [vp, ctypes.c_int, dp, vp, ctypes.c_int, vp]
    L.cq_get_kernel_times.argtypes = [vp, dp, P(ctypes.c_longlong), ctypes.c_int]
    L.cq_get_screen_times.argtypes = [vp, dp, P(ctypes.c_longlong), ctypes.c_int]
    L.cq_get_work_stats.argtypes = [vp, dp, ctypes.c_int]
    L.cq_solid_harmonics.argtypes = [ctypes.c_int, dp]
    L.cq_set_class_digestion.argtypes = [vp, ctypes.c_int, ctypes.c_int]
    L.cq_get_class_digestion.argtypes = [vp, ctypes.c_int, ip]
    L.cq_set_class_stream.argtypes = [vp, ctypes.c_int, ctypes.c_int]
    L.cq_set_class_chunk_group.argtypes = [vp, ctypes.c_int, ctypes.c_int]
    L.cq_set_task_budget.argtypes = [vp, ctypes.c_longlong]
    L.cq_get_task_plan.argtypes = [vp, P(ctypes.c_longlong), ip]
    L.cq_set_class_tile.argtypes = [vp, ctypes.c_int, ctypes.c_longlong, ctypes.c_int, ctypes.c_int]
    L.cq_get_class_tile.argtypes = [vp, ctypes.c_int, ctypes.c_int, ip, ip, ip, P(ctypes.c_longlong)]
    ci, cl, cf = ctypes.c_int, ctypes.c_long, ctypes.c_float
    L.cq_mlp_last_error.argtypes = []
    L.cq_mlp_last_error.restype = ctypes.c_char_p
    L.cq_mlp_gemm_bf16.argtypes = [vp, ci, ci, vp, ci, ci, vp, ci, vp, ci, ci, ci, vp, ci, ci, vp]
    L.cq_mlp_relu_backward.argtypes = [vp, vp, vp, vp, cl, vp]
    L.cq_mlp_column_sum.argtypes = [vp, ci, ci, vp, vp]
    L.cq_mlp_to_bf16.argtypes = [vp, vp, cl, vp]
    L.cq_mlp_mse.argtypes = [vp, vp, vp, vp, ci, ci, cf, vp, vp, vp]
    L.cq_mlp_adam.argtypes = [vp, vp, vp, vp, vp, cl, cf, cf, cf, cf, ci, cf, vp]
    for name in EXPORTED:
        fn = getattr(L, name)
        if name not in ("cq_last_error", "cq_mlp_last_error"):
            fn.restype = ctypes.c_int
    _lib = L
    return L


def check(rc: int, ctx=None):
    if rc != CQ_OK:
        msg = load().cq_last_error(ctx)
        raise CQError(f"libcqgpu error {rc}: {msg.decode() if msg else ''}")
