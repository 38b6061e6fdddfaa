"""The renderer through the C ABI (owe_hip_renderer_*).

The Python face of the reference's Renderer (src/owe/render/renderer.hpp:19-79) for the hip
backend, with the path or the light integrator: ``run_pass(spp)`` adds a pass (spp samples, or
spp particles, per pixel), ``resolve()`` returns the XYZ image [H, W, 3] in double, ``stats()`` the
TransportStats. ``camera_paths`` and ``particles`` return single paths' or particles' XYZ for
per-path parity. ``Batch`` renders many scenes (variants of one family) as one persistent launch
per pass (owe_hip_batch_*).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from .scene import CRenderSettings, CScene, CTransportStats, Scene, settings as make_settings


def set_tuning_cache(path) -> None:
    """The autotuner's per-device tuning cache file (None: none); process-wide."""
    _lib.check(_lib.load().owe_hip_set_tuning_cache(None if path is None else str(path).encode()))


class Renderer:
    def __init__(self, scene: Scene, settings: CRenderSettings | None = None, device: int = 0, _handle=None):
        self.scene = scene
        self.settings = settings if settings is not None else make_settings()
        self.width, self.height = scene.observer["width"], scene.observer["height"]
        self._owned = _handle is None
        if _handle is not None:  # a batch member: the batch owns it
            self._h = _handle
            return
        c = scene.c()
        h = ctypes.c_void_p()
        _lib.check(_lib.load().owe_hip_renderer_create(device, ctypes.byref(c), ctypes.byref(self.settings),
                                                       ctypes.byref(h)))
        self._h = h

    def close(self) -> None:
        if self._h and self._owned:
            _lib.check(_lib.load().owe_hip_renderer_destroy(self._h))
        self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run_pass(self, spp: int) -> None:
        _lib.check(_lib.load().owe_hip_renderer_pass(self._h, int(spp)))

    def resolve(self) -> np.ndarray:
        out = np.zeros((self.height, self.width, 3), dtype=np.float64)
        _lib.check(_lib.load().owe_hip_renderer_resolve(self._h, out.ctypes.data))
        return out

    def sppm_state(self) -> dict:
        """The SPPM state: per pixel τ (XYZ), N and r, photons per iteration and iterations so far."""
        st = np.zeros((self.height, self.width, 5), dtype=np.float64)
        photons, its = ctypes.c_uint64(), ctypes.c_longlong()
        _lib.check(_lib.load().owe_hip_renderer_sppm_state(self._h, st.ctypes.data, ctypes.byref(photons),
                                                            ctypes.byref(its)))
        return dict(tau=st[..., :3], N=st[..., 3], r=st[..., 4], photons=photons.value, iterations=its.value)

    def stats(self) -> dict:
        s = CTransportStats()
        _lib.check(_lib.load().owe_hip_renderer_stats(self._h, ctypes.byref(s)))
        return {"paths": s.paths, "segments": s.segments, "inconsistencies": s.inconsistencies, "leaks": s.leaks}

    def first_inconsistency(self) -> dict | None:
        """TransportStats::first: the first region inconsistency (boundary, ray_region, boundary_region,
        previous, p, origin, direction, primitive; world coordinates), or None."""
        class _Rec(ctypes.Structure):
            _fields_ = [("boundary", ctypes.c_uint32), ("ray_region", ctypes.c_uint32),
                        ("boundary_region", ctypes.c_uint32), ("previous", ctypes.c_uint32),
                        ("p", ctypes.c_double * 3), ("origin", ctypes.c_double * 3), ("direction", ctypes.c_double * 3),
                        ("previous_primitive", ctypes.c_uint32), ("primitive", ctypes.c_uint32)]
        rec, found = _Rec(), ctypes.c_int()
        _lib.check(_lib.load().owe_hip_renderer_first_inconsistency(self._h, ctypes.byref(rec), ctypes.byref(found)))
        if not found.value:
            return None
        return dict(boundary=rec.boundary, ray_region=rec.ray_region, boundary_region=rec.boundary_region,
                    previous=rec.previous, p=tuple(rec.p), origin=tuple(rec.origin), direction=tuple(rec.direction),
                    primitive=rec.primitive)

    def progress(self) -> tuple[int, int]:
        p, n = ctypes.c_int(), ctypes.c_longlong()
        _lib.check(_lib.load().owe_hip_renderer_progress(self._h, ctypes.byref(p), ctypes.byref(n)))
        return p.value, n.value

    def reset(self) -> None:
        _lib.check(_lib.load().owe_hip_renderer_reset(self._h))

    KERNELS = ("generate", "intersect", "shade", "shadow", "splat", "persistent")
    WAVEFRONT, PERSISTENT = 0, 1

    def set_mode(self, mode: int, blocks_per_cu: int = 0, chunk: int = 0, use_graph: bool = False,
                 waves_per_simd: int = 0, refill: int = 0) -> None:
        """The wavefront (0) or persistent batched (1) form of the camera loop; 0 keeps a shape value."""
        _lib.check(_lib.load().owe_hip_renderer_set_mode(self._h, int(mode), int(blocks_per_cu), int(chunk),
                                                         int(use_graph), int(waves_per_simd), int(refill)))

    def mode(self) -> dict:
        v = [ctypes.c_int() for _ in range(6)]
        _lib.check(_lib.load().owe_hip_renderer_get_mode(self._h, *[ctypes.byref(x) for x in v]))
        return dict(mode=v[0].value, blocks_per_cu=v[1].value, chunk=v[2].value, use_graph=bool(v[3].value),
                    waves_per_simd=v[4].value, refill=v[5].value)

    def set_vcm_in_flight(self, iterations: int) -> None:
        """vcm/bdpt: iterations run together (1 up to what the caches hold; 0: all of those). The
        image is the same bit for bit whatever the number."""
        _lib.check(_lib.load().owe_hip_renderer_set_vcm_in_flight(self._h, int(iterations)))

    def vcm_in_flight(self) -> tuple[int, int]:
        """(iterations run together, the most the renderer's caches hold)."""
        a, b = ctypes.c_int(), ctypes.c_int()
        _lib.check(_lib.load().owe_hip_renderer_get_vcm_in_flight(self._h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def autotune(self, spp: int) -> float:
        """Picks the fastest launch shape of the current form over passes of spp; returns its ms."""
        ms = ctypes.c_double()
        _lib.check(_lib.load().owe_hip_renderer_autotune(self._h, int(spp), ctypes.byref(ms)))
        return ms.value

    def tuning(self) -> dict:
        """The wavefront form's tuned bounce rounds per sync and paths in flight; whether the last
        autotune was a tuning-cache hit."""
        k, q, c = ctypes.c_int(), ctypes.c_uint64(), ctypes.c_int()
        _lib.check(_lib.load().owe_hip_renderer_tuning(self._h, ctypes.byref(k), ctypes.byref(q), ctypes.byref(c)))
        return {"rounds_per_sync": k.value, "queue_capacity": q.value, "from_cache": bool(c.value)}

    def set_timing(self, enable: bool = True) -> None:
        _lib.check(_lib.load().owe_hip_renderer_set_timing(self._h, int(enable)))

    def kernel_times(self) -> dict:
        """{kernel: (milliseconds, launches)} since timing was enabled (HIP events on the render stream)."""
        ms = np.zeros(6)
        n = np.zeros(6, dtype=np.uint64)
        _lib.check(_lib.load().owe_hip_renderer_kernel_times(self._h, ms.ctypes.data, n.ctypes.data))
        return {k: (float(ms[i]), int(n[i])) for i, k in enumerate(self.KERNELS)}

    def camera_paths(self, spp: int, pixel, sample, pass_index: int = 0) -> np.ndarray:
        pixel = np.ascontiguousarray(pixel, dtype=np.uint32)
        sample = np.ascontiguousarray(sample, dtype=np.uint32)
        out = np.zeros((len(pixel), 3), dtype=np.float64)
        _lib.check(_lib.load().owe_hip_renderer_camera_paths(self._h, int(spp), int(pass_index), pixel.ctypes.data,
                                                             sample.ctypes.data, len(pixel), out.ctypes.data))
        return out

    def particles(self, index, pass_index: int = 0) -> np.ndarray:
        """XYZ each particle index[i] of a light-tracing pass deposits (summed over its splats)."""
        index = np.ascontiguousarray(index, dtype=np.uint64)
        out = np.zeros((len(index), 3), dtype=np.float64)
        _lib.check(_lib.load().owe_hip_renderer_particles(self._h, int(pass_index), index.ctypes.data, len(index),
                                                          out.ctypes.data))
        return out


class Batch:
    """Many scenes rendered with camera paths as one persistent launch per pass (owe_hip_batch_*):
    ``members[i]`` is scene i's Renderer (owned by the batch: resolve, stats, progress; member 0's
    launch shape, set or autotuned on it, is the batch's)."""

    def __init__(self, scenes, settings: CRenderSettings | None = None, device: int = 0):
        self.scenes = list(scenes)
        self.settings = settings if settings is not None else make_settings()
        cs = [sc.c() for sc in self.scenes]
        arr = (CScene * len(cs))(*cs) if cs else None
        h = ctypes.c_void_p()
        _lib.check(_lib.load().owe_hip_batch_create(device, arr, len(cs), ctypes.byref(self.settings), ctypes.byref(h)))
        self._h = h
        self.members = []
        for i, sc in enumerate(self.scenes):
            m = ctypes.c_void_p()
            _lib.check(_lib.load().owe_hip_batch_member(h, i, ctypes.byref(m)))
            self.members.append(Renderer(sc, self.settings, _handle=m))

    def close(self) -> None:
        if self._h:
            for m in self.members:
                m._h = ctypes.c_void_p()
            _lib.check(_lib.load().owe_hip_batch_destroy(self._h))
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run_pass(self, spp: int) -> None:
        _lib.check(_lib.load().owe_hip_batch_pass(self._h, int(spp)))

    def stats(self) -> dict:
        s = CTransportStats()
        _lib.check(_lib.load().owe_hip_batch_stats(self._h, ctypes.byref(s)))
        return {"paths": s.paths, "segments": s.segments, "inconsistencies": s.inconsistencies, "leaks": s.leaks}

    def set_timing(self, enable: bool = True) -> None:
        _lib.check(_lib.load().owe_hip_batch_set_timing(self._h, int(enable)))

    def kernel_time(self) -> tuple[float, int]:
        """(milliseconds, launches) of the batch's persistent launches since timing was enabled."""
        ms, n = ctypes.c_double(), ctypes.c_uint64()
        _lib.check(_lib.load().owe_hip_batch_kernel_time(self._h, ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value
