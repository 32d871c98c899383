"""Loader of the HIP library (libsm2d_hip.so) with a loud failure when it is unusable."""
from __future__ import annotations

import ctypes as C
import os

from . import abi
from . import io as sio
from .engine import Library

_LIB = None


# Prototypes beyond the entry points the oracle shares (engine.abi.COMMON_PROTOS).
HIP_PROTOS = {
    "set_profiling": [C.c_void_p, C.c_int],
    "get_timing": [C.c_void_p, C.POINTER(abi.Timing)],
    "ctx_create_batch": [C.c_int, C.POINTER(abi.Grid), C.c_int, C.POINTER(C.c_void_p)],
    "run_batch": [C.c_void_p, C.POINTER(abi.Spot), C.c_int, C.POINTER(abi.Transport),
                  abi.dp, abi.fp, C.POINTER(abi.Audit)],
    "batch_spot_edep": [C.c_void_p, C.c_int, abi.dp],
    "set_transport_waves": [C.c_void_p, C.c_int],
    "set_gather_rows": [C.c_void_p, C.c_int],
    "tuner_report": [C.c_void_p, C.c_char_p, C.c_int],
    "set_graphs": [C.c_void_p, C.c_int],
    "mlp_create": [C.c_int, C.c_uint32, C.POINTER(C.c_void_p)],
    "mlp_destroy": [C.c_void_p],
    "mlp_get_params": [C.c_void_p, abi.fp, abi.fp],
    "mlp_set_params": [C.c_void_p, abi.fp, abi.fp],
    "mlp_forward": [C.c_void_p, abi.fp, C.c_int, abi.fp],
    "mlp_infer": [C.c_void_p, abi.fp, C.c_int, abi.fp],
    "mlp_grad": [C.c_void_p, abi.fp, abi.fp, C.c_int, C.c_float, C.c_float, abi.fp, abi.fp],
    "mlp_train": [C.c_void_p, C.c_void_p, C.POINTER(abi.MlpTrainCfg), abi.fp, C.c_int],
    "mlp_evaluate": [C.c_void_p, C.c_void_p, C.c_int, C.c_uint32, abi.fp],
    "surrogate_attach": [C.c_void_p, C.c_void_p, C.c_int],
    "surrogate_enable": [C.c_void_p, C.c_int],
    "surrogate_export": [C.c_void_p, abi.fp, C.c_int, abi.ip, abi.ip, abi.i32p],
    "surrogate_set_table": [C.c_void_p, abi.fp, C.c_int, C.c_int, abi.i32p],
    "surrogate_check": [C.c_void_p, C.c_float, abi.fp],
    "surrogate_stats": [C.c_void_p, abi.fp],
    "stage_clocks": [C.c_int, abi.dp, C.c_int],
    "device_name": [C.c_int, C.c_char_p, C.c_int],
}


def load_library() -> Library:
    """Loads the in-tree HIP library.  There is no CPU fallback: a missing library raises, and
    so does every transport entry point without a GPU (the host-only configuration, file
    and planning functions of sm_2d_amd.io work anywhere)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    # SM2D_LIB selects another build of the same sources (the stage-clock diagnostic build)
    path = os.environ.get("SM2D_LIB", abi.LIB_PATH)
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing; build it with `python -m sm_2d_amd.build`")
    lib = Library(path, "sm2d_", {**HIP_PROTOS, **sio.PROTOS})
    lib.lib.sm2d_abi_version.restype = C.c_int
    lib.lib.sm2d_mlp_param_count.restype = C.c_int
    if lib.lib.sm2d_abi_version() != 1:
        raise RuntimeError("libsm2d_hip.so ABI version mismatch")
    _LIB = lib
    return lib
