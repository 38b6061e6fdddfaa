"""Batch ray queries on a triangle mesh through the C ABI (owe_hip_mesh_*, owe_hip_trace_batch).

The Python face of ``MeshShape::intersect`` (reference src/owe/scene/mesh.cpp:73-97) over many rays
at once, and of the float kernels' any-hit ``occludedMesh`` (shaders/geometry.slang:577-615). Rays
are ``[n, 8]`` float32 device tensors ``(o.x, o.y, o.z, tmin, d.x, d.y, d.z, tmax)``; torch only
holds device memory and the stream.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib


class Mesh:
    """A triangle mesh resident on one device with its 4-wide BVH."""

    def __init__(self, positions: np.ndarray, triangles: np.ndarray, device: int = 0):
        pos = np.ascontiguousarray(positions, dtype=np.float32).reshape(-1, 3)
        tri = np.ascontiguousarray(triangles, dtype=np.uint32).reshape(-1, 3)
        if tri.size and int(tri.max()) >= len(pos):
            raise ValueError("triangle index out of range")
        lib = _lib.load()
        handle = ctypes.c_void_p()
        _lib.check(lib.owe_hip_mesh_create(device, pos.ctypes.data, len(pos), tri.ctypes.data, len(tri),
                                           ctypes.byref(handle)))
        self._h = handle
        self.device = device
        self.n_triangles = len(tri)

    def stats(self) -> _lib.MeshStats:
        s = _lib.MeshStats()
        _lib.check(_lib.load().owe_hip_mesh_stats_get(self._h, ctypes.byref(s)))
        return s

    PER_LANE, REFILL = 0, 1

    def set_launch(self, form: int, blocks_per_cu: int = 0, refill: int = 0, mode: int = -1) -> None:
        """One lane per ray (0) or persistent lanes with refill (1) for query mode `mode` (closest 0,
        any 1, both −1); 0 keeps a shape value."""
        _lib.check(_lib.load().owe_hip_mesh_set_launch(self._h, int(mode), int(form), int(blocks_per_cu),
                                                       int(refill)))

    def launch(self, mode: int = _lib.QUERY_CLOSEST) -> dict:
        v = [ctypes.c_int() for _ in range(3)]
        _lib.check(_lib.load().owe_hip_mesh_get_launch(self._h, int(mode), *[ctypes.byref(x) for x in v]))
        return dict(form=v[0].value, blocks_per_cu=v[1].value, refill=v[2].value)

    def close(self) -> None:
        if self._h:
            _lib.check(_lib.load().owe_hip_mesh_destroy(self._h))
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def trace(self, rays: torch.Tensor, mode: int = _lib.QUERY_CLOSEST, prim: torch.Tensor | None = None,
              t: torch.Tensor | None = None, stream: torch.cuda.Stream | None = None):
        """Queries every ray; returns ``(prim int32 [n], t float32 [n])`` (prim -1 on a miss).

        Outputs may be passed in (preallocated, same device) to keep a timed loop allocation-free.
        """
        if not (rays.is_cuda and rays.dtype == torch.float32 and rays.dim() == 2 and rays.shape[1] == 8
                and rays.is_contiguous()):
            raise ValueError("rays must be a contiguous [n, 8] float32 tensor on the device")
        if rays.device.index != self.device:
            raise ValueError("rays live on another device than the mesh")
        n = rays.shape[0]
        if prim is None:
            prim = torch.empty(n, dtype=torch.int32, device=rays.device)
        if t is None:
            t = torch.empty(n, dtype=torch.float32, device=rays.device)
        for name, buf, dt in (("prim", prim, torch.int32), ("t", t, torch.float32)):
            if not (buf.is_cuda and buf.dtype == dt and buf.is_contiguous() and buf.numel() >= n
                    and buf.device == rays.device):
                raise ValueError(f"{name} must be a contiguous {dt} device tensor with at least {n} entries")
        s = stream if stream is not None else torch.cuda.current_stream(rays.device)
        _lib.check(_lib.load().owe_hip_trace_batch(self._h, int(mode), rays.data_ptr(), n, prim.data_ptr(),
                                                   t.data_ptr(), ctypes.c_void_p(s.cuda_stream)))
        return prim, t
