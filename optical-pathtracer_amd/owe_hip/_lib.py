"""ctypes binding of libowe_hip.so, the backend's C ABI (include/owe_hip.h).

The library is built in-tree (``make -C optical-pathtracer_amd``); importing this module without it
raises at once — there is no fallback path.
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libowe_hip.so"

OK = 0
ERR_INVALID = 1
ERR_UNAVAILABLE = 2
ERR_DEVICE = 3
ERR_UNSUPPORTED = 4

QUERY_CLOSEST = 0
QUERY_ANY = 1


class OweHipError(RuntimeError):
    """A failed C-ABI call: ``code`` is the status, the message the library's last error."""

    def __init__(self, code: int, message: str):
        super().__init__(f"owe_hip error {code}: {message}")
        self.code = code


class DeviceInfo(ctypes.Structure):
    _fields_ = [
        ("index", ctypes.c_int),
        ("name", ctypes.c_char * 128),
        ("arch", ctypes.c_char * 32),
        ("type", ctypes.c_char * 16),
        ("hbm_bytes", ctypes.c_uint64),
        ("compute_units", ctypes.c_int),
        ("xcds", ctypes.c_int),
    ]


class MeshStats(ctypes.Structure):
    _fields_ = [
        ("triangles", ctypes.c_uint64),
        ("nodes", ctypes.c_uint64),
        ("depth", ctypes.c_uint32),
        ("stack_bound", ctypes.c_uint32),
        ("build_seconds", ctypes.c_double),
        ("device_bytes", ctypes.c_uint64),
    ]


_lib = None


def _declare(lib, name, restype, argtypes):
    fn = getattr(lib, name)
    fn.restype = restype
    fn.argtypes = argtypes


def load() -> ctypes.CDLL:
    """The loaded library (built in-tree); raises FileNotFoundError when it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise FileNotFoundError(
            f"{LIB_PATH} is missing: build it with `make -C optical-pathtracer_amd` "
            "(or __graft_entry__.build()); the hip backend has no fallback")
    # PyTorch ships its own HIP runtime. Loading it first makes the process use one runtime for both
    # (the library resolves libamdhip64 to the copy already loaded); loading this library first
    # would leave torch unable to see the device.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = ctypes.CDLL(os.fspath(LIB_PATH))
    c_int, c_u32, c_u64, c_vp = ctypes.c_int, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_void_p
    _declare(lib, "owe_hip_last_error", ctypes.c_char_p, [])
    _declare(lib, "owe_hip_backend_name", ctypes.c_char_p, [])
    _declare(lib, "owe_hip_backend_summary", ctypes.c_char_p, [])
    _declare(lib, "owe_hip_backend_integrators", ctypes.c_char_p, [])
    _declare(lib, "owe_hip_available", c_int, [ctypes.c_char_p, ctypes.c_size_t])
    _declare(lib, "owe_hip_device_count", c_int, [ctypes.POINTER(c_int)])
    _declare(lib, "owe_hip_device_info_get", c_int, [c_int, ctypes.POINTER(DeviceInfo)])
    _declare(lib, "owe_hip_mesh_create", c_int,
             [c_int, c_vp, c_u32, c_vp, c_u32, ctypes.POINTER(c_vp)])
    _declare(lib, "owe_hip_mesh_destroy", c_int, [c_vp])
    _declare(lib, "owe_hip_mesh_stats_get", c_int, [c_vp, ctypes.POINTER(MeshStats)])
    _declare(lib, "owe_hip_trace_batch", c_int, [c_vp, c_int, c_vp, c_u64, c_vp, c_vp, c_vp])
    _declare(lib, "owe_hip_mesh_set_launch", c_int, [c_vp, c_int, c_int, c_int, c_int])
    _declare(lib, "owe_hip_mesh_get_launch", c_int, [c_vp, c_int, ctypes.POINTER(c_int), ctypes.POINTER(c_int),
                                                     ctypes.POINTER(c_int)])
    _declare(lib, "owe_hip_renderer_create", c_int, [c_int, c_vp, c_vp, ctypes.POINTER(c_vp)])
    _declare(lib, "owe_hip_renderer_destroy", c_int, [c_vp])
    _declare(lib, "owe_hip_renderer_pass", c_int, [c_vp, c_int])
    _declare(lib, "owe_hip_renderer_resolve", c_int, [c_vp, c_vp])
    _declare(lib, "owe_hip_renderer_stats", c_int, [c_vp, c_vp])
    _declare(lib, "owe_hip_renderer_first_inconsistency", c_int, [c_vp, c_vp, c_vp])
    _declare(lib, "owe_hip_renderer_sppm_state", c_int, [c_vp, c_vp, c_vp, c_vp])
    _declare(lib, "owe_hip_renderer_progress", c_int, [c_vp, ctypes.POINTER(c_int), ctypes.POINTER(ctypes.c_longlong)])
    _declare(lib, "owe_hip_renderer_reset", c_int, [c_vp])
    _declare(lib, "owe_hip_renderer_camera_paths", c_int, [c_vp, c_int, c_int, c_vp, c_vp, c_u64, c_vp])
    _declare(lib, "owe_hip_renderer_particles", c_int, [c_vp, c_int, c_vp, c_u64, c_vp])
    _declare(lib, "owe_hip_renderer_set_mode", c_int, [c_vp, c_int, c_int, c_int, c_int, c_int, c_int])
    _declare(lib, "owe_hip_renderer_get_mode", c_int, [c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp])
    _declare(lib, "owe_hip_renderer_set_vcm_in_flight", c_int, [c_vp, c_int])
    _declare(lib, "owe_hip_renderer_get_vcm_in_flight", c_int, [c_vp, c_vp, c_vp])
    _declare(lib, "owe_hip_set_tuning_cache", c_int, [ctypes.c_char_p])
    _declare(lib, "owe_hip_renderer_tuning", c_int, [c_vp, ctypes.POINTER(c_int), ctypes.POINTER(c_u64), ctypes.POINTER(c_int)])
    _declare(lib, "owe_hip_renderer_autotune", c_int, [c_vp, c_int, c_vp])
    _declare(lib, "owe_hip_lens_batch_create", c_int, [c_int, c_vp, c_u32, c_vp, c_u32, c_vp, c_u32, c_vp,
                                                        ctypes.POINTER(c_vp)])
    _declare(lib, "owe_hip_lens_batch_destroy", c_int, [c_vp])
    _declare(lib, "owe_hip_lens_batch_run", c_int, [c_vp])
    _declare(lib, "owe_hip_lens_batch_results", c_int, [c_vp, c_vp, c_vp, c_vp])
    _declare(lib, "owe_hip_lens_batch_set_shape", c_int, [c_vp, c_int, c_int])
    _declare(lib, "owe_hip_lens_batch_autotune", c_int, [c_vp, c_vp])
    _declare(lib, "owe_hip_lens_batch_set_params", c_int, [c_vp, c_vp, ctypes.c_uint32])
    _declare(lib, "owe_hip_lens_batch_gradients", c_int, [c_vp, c_vp])
    _declare(lib, "owe_hip_lens_batch_trace", c_int, [c_vp, c_u32, c_int, c_vp, c_u64, c_vp])
    _declare(lib, "owe_hip_lens_batch_last_ms", c_int, [c_vp, c_vp])
    _declare(lib, "owe_hip_caustic_create", c_int, [c_int, c_vp, ctypes.POINTER(c_vp)])
    _declare(lib, "owe_hip_caustic_destroy", c_int, [c_vp])
    _declare(lib, "owe_hip_caustic_forward", c_int, [c_vp, c_vp, c_vp, c_u64, c_vp])
    _declare(lib, "owe_hip_caustic_backward", c_int, [c_vp, c_vp, c_vp, c_vp, c_u64, c_vp])
    _declare(lib, "owe_hip_renderer_set_timing", c_int, [c_vp, c_int])
    _declare(lib, "owe_hip_renderer_kernel_times", c_int, [c_vp, c_vp, c_vp])
    _declare(lib, "owe_hip_batch_create", c_int, [c_int, c_vp, ctypes.c_int32, c_vp, ctypes.POINTER(c_vp)])
    _declare(lib, "owe_hip_batch_destroy", c_int, [c_vp])
    _declare(lib, "owe_hip_batch_member", c_int, [c_vp, ctypes.c_int32, ctypes.POINTER(c_vp)])
    _declare(lib, "owe_hip_batch_pass", c_int, [c_vp, c_int])
    _declare(lib, "owe_hip_batch_stats", c_int, [c_vp, c_vp])
    _declare(lib, "owe_hip_batch_set_timing", c_int, [c_vp, c_int])
    _declare(lib, "owe_hip_batch_kernel_time", c_int, [c_vp, c_vp, c_vp])
    _lib = lib
    return lib


def check(status: int) -> None:
    """Raises OweHipError for a non-zero status, with the library's description."""
    if status != OK:
        msg = load().owe_hip_last_error()
        raise OweHipError(status, msg.decode() if msg else "")


def available() -> tuple[bool, str]:
    buf = ctypes.create_string_buffer(256)
    ok = load().owe_hip_available(buf, len(buf))
    return bool(ok), buf.value.decode()


def device_info(index: int = 0) -> DeviceInfo:
    info = DeviceInfo()
    check(load().owe_hip_device_info_get(index, ctypes.byref(info)))
    return info
