"""ctypes binding of hussar_amd/libhussar_hip.so (the C ABI of include/hussar_hip.h).

There is no fallback: if the library is missing or fails to load, every entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os

from .types import (c_device_info, c_frame_config, c_knobs, c_request, c_request_info, c_run_stats, c_scene,
                    c_server_config, c_server_stats)

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libhussar_hip.so")

HSR_OK = 0
STATUS_NAMES = {0: "HSR_OK", 1: "HSR_ERR_INVALID", 2: "HSR_ERR_HIP", 3: "HSR_ERR_OOM", 4: "HSR_ERR_INTERRUPTED",
                5: "HSR_ERR_UNSUPPORTED", 6: "HSR_ERR_BUSY"}
OPT_COUNTERS, OPT_RECORD_HITS, OPT_PATHS_PER_LAUNCH, OPT_PIPELINE, OPT_PHASE_MODE = 1, 2, 3, 4, 5
PHASE_FP64, PHASE_FP32, PHASE_DOUBLE_FLOAT = 0, 1, 2
OPT_TRAV_WAVES = 6
OPT_BATCH_PATHS = 7
OPT_CUBE_EVENTS = 8
OPT_CHAIN = 9
OPT_GRAPH = 10
OPT_TAIL = 11
OPT_MARKERS = 12
OPT_PRIMARY = 13
OPT_PRIMARY_MIN = 14
OPT_PRIMARY_RES = 15
OPT_PRIMARY_CAP = 16
OPT_SHADOW_SPLAT = 17

# Every exported symbol and its signature: (restype, argtypes)
_P = C.POINTER
_vp = C.c_void_p
SIGNATURES = {
    "hsr_knobs_default": (None, [_P(c_knobs)]),
    "hsr_hip_create": (C.c_int, [_P(C.c_float), C.c_int64, _P(C.c_int32), C.c_int64, C.c_int32, _P(_vp)]),
    "hsr_hip_destroy": (None, [_vp]),
    "hsr_hip_set_velocities": (C.c_int, [_vp, _P(C.c_float), C.c_int64]),
    "hsr_hip_configure_frame": (C.c_int, [_vp, _P(c_frame_config)]),
    "hsr_hip_clear_frame": (C.c_int, [_vp]),
    "hsr_hip_run": (C.c_int, [_vp, _P(c_scene), _P(c_knobs), C.c_int64, C.c_int64, C.c_float, C.c_int32, _vp,
                              _P(C.c_uint8)]),
    "hsr_hip_fetch_frame": (C.c_int, [_vp, _P(C.c_float), _P(C.c_double), _P(C.c_double)]),
    "hsr_hip_fetch_frame_fft": (C.c_int, [_vp, _P(C.c_float), _P(C.c_double)]),
    "hsr_hip_frame_peak": (C.c_int, [_vp, C.c_int32, _P(C.c_int32), _P(C.c_double), _P(C.c_double)]),
    "hsr_hip_frame_peak_noalloc": (C.c_int, [_vp, C.c_int32, _P(C.c_int32), _P(C.c_double), _P(C.c_double)]),
    "hsr_path_tracer_run": (C.c_int, [_vp, _P(c_scene), _P(c_knobs), C.c_int64, _vp, _P(C.c_uint8)]),
    "hsr_path_tracer_run_schedule": (C.c_int, [_vp, _P(c_scene), _P(c_knobs), C.c_int32, C.c_int64, C.c_int64,
                                               _vp, _P(C.c_uint8)]),
    "hsr_guide_create": (C.c_int, [_vp, C.c_float, C.c_float, C.c_int32, _P(_vp)]),
    "hsr_guide_destroy": (None, [_vp]),
    "hsr_guide_reset": (C.c_int, [_vp]),
    "hsr_guide_step": (C.c_int, [_vp]),
    "hsr_guide_sampling_size": (C.c_int, [_vp, _P(C.c_int32)]),
    "hsr_guide_get_sampling": (C.c_int, [_vp, _P(C.c_int32), _P(C.c_float), C.c_int32]),
    "hsr_guide_set_sampling": (C.c_int, [_vp, _P(C.c_int32), _P(C.c_float), C.c_int32]),
    "hsr_guide_training_size": (C.c_int, [_vp, _P(C.c_int32)]),
    "hsr_guide_get_training": (C.c_int, [_vp, _P(C.c_int32), _P(C.c_double), C.c_int32]),
    "hsr_hip_last_stats": (C.c_int, [_vp, _P(c_run_stats)]),
    "hsr_hip_set_option": (C.c_int, [_vp, C.c_int32, C.c_int64]),
    "hsr_hip_get_hit_record": (C.c_int, [_vp, _P(C.c_int32), _P(C.c_int8), C.c_int64, C.c_int32, C.c_int32]),
    "hsr_hip_device_info": (C.c_int, [C.c_int32, _P(c_device_info)]),
    "hsr_last_error": (C.c_char_p, []),
    "hsr_abi_version": (C.c_int32, []),
    "hsr_server_create": (C.c_int, [_P(c_server_config), _P(_vp)]),
    "hsr_server_destroy": (None, [_vp]),
    "hsr_server_open_scene": (C.c_int, [_vp, _P(C.c_float), C.c_int64, _P(C.c_int32), C.c_int64, _P(C.c_int32)]),
    "hsr_server_set_velocities": (C.c_int, [_vp, C.c_int32, _P(C.c_float), C.c_int64]),
    "hsr_server_close_scene": (C.c_int, [_vp, C.c_int32]),
    "hsr_server_submit": (C.c_int, [_vp, _P(c_request), _P(C.c_int64)]),
    "hsr_server_request_bytes": (C.c_int, [_vp, _P(c_request), _P(C.c_int64)]),
    "hsr_server_poll": (C.c_int, [_vp, C.c_int64, _P(C.c_int32)]),
    "hsr_server_wait": (C.c_int, [_vp, C.c_int64, _P(C.c_float), C.c_int64, _P(c_request_info)]),
    "hsr_server_get_stats": (C.c_int, [_vp, _P(c_server_stats)]),
    "hsr_server_reset_stats": (C.c_int, [_vp]),
}

_lib = None


class HsrError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {message}")
        self.status = status


ABI_VERSION = 5  # HSR_ABI_VERSION of include/hussar_hip.h


def lib():
    """Load the backend library (raises if it is missing: there is no CPU fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with hussar_amd.build.build() "
                               "(hipcc --offload-arch=gfx950); the backend has no CPU fallback")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        if L.hsr_abi_version() != ABI_VERSION:
            raise RuntimeError(f"{LIB_PATH} has ABI version {L.hsr_abi_version()}, the bindings expect "
                               f"{ABI_VERSION}: rebuild it with hussar_amd.build.build()")
        _lib = L
    return _lib


def check(status: int) -> None:
    if status != HSR_OK:
        msg = lib().hsr_last_error()
        raise HsrError(status, msg.decode() if msg else "")
