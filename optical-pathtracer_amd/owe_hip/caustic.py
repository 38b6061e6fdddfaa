"""Differentiable caustic design (include/owe_hip_caustic.h) for the Python tooling.

``Caustic(settings)`` holds the GPU state; ``irradiance(heights, pass_index)`` is a
``torch.autograd.Function``: its forward traces the pass's particles through the height field into
the wall grid, and its backward replays the same particles (path-replay backpropagation) to turn
∂L/∂irradiance into ∂L/∂heights — nothing per path is stored between the two.

``caustic_settings()`` builds the inverse-caustic workload: a 100 × 100 × 10 mm PMMA slab
(Sellmeier, src/owe/core/medium.cpp:135) whose exit face is a 256 × 256 height field, a collimated
5778 K sun, 1,048,576 particles per pass, a 512 × 512 irradiance grid on a wall 1 m beyond the slab.
``initial_heights`` is seeded smooth noise of 0.1 mm amplitude and ``target`` a seeded sum of twelve
Gaussians carrying the same flux as the slab transmits.
"""
from __future__ import annotations

import ctypes
import math

import numpy as np

from . import _lib
from .scene import CMedium, INDEX_SELLMEIER

PMMA_B, PMMA_C = (1.1819,), (0.011313,)


class CCausticSettings(ctypes.Structure):
    _fields_ = [("nodes", ctypes.c_uint32), ("pixels", ctypes.c_uint32), ("particles", ctypes.c_uint64),
                ("seed", ctypes.c_uint64), ("slab_half", ctypes.c_double), ("thickness", ctypes.c_double),
                ("wall_z", ctypes.c_double), ("wall_half", ctypes.c_double), ("sun_kelvin", ctypes.c_double),
                ("medium", CMedium)]


def caustic_settings(nodes=256, pixels=512, particles=1 << 20, seed=1, slab_half=0.05, thickness=0.01,
                     wall_distance=1.0, wall_half=0.1, kelvin=5778.0) -> CCausticSettings:
    s = CCausticSettings()
    s.nodes, s.pixels, s.particles, s.seed = nodes, pixels, particles, seed
    s.slab_half, s.thickness, s.wall_z, s.wall_half, s.sun_kelvin = (slab_half, thickness, thickness + wall_distance,
                                                                     wall_half, kelvin)
    s.medium.index_kind, s.medium.terms = INDEX_SELLMEIER, 1
    s.medium.coeff_a[0], s.medium.coeff_b[0] = PMMA_B[0], PMMA_C[0]
    return s


def initial_heights(nodes=256, amplitude=1e-4, seed=7, feature=24) -> np.ndarray:
    """Seeded smooth noise: bicubic-interpolated random values every `feature` nodes (and half that,
    at half weight), scaled to ±amplitude metres."""
    rng = np.random.default_rng(seed)
    total = np.zeros((nodes, nodes))
    for f, wgt in ((feature, 1.0), (feature // 2, 0.5)):
        n = (nodes - 1) // f + 4
        g = rng.uniform(-1, 1, (n, n))
        t = np.arange(nodes) / f
        i = np.floor(t).astype(int)
        u = t - i

        def cubic(p0, p1, p2, p3, x):
            return p1 + 0.5 * x * (p2 - p0 + x * (2 * p0 - 5 * p1 + 4 * p2 - p3 + x * (3 * (p1 - p2) + p3 - p0)))

        rows = cubic(g[i], g[i + 1], g[i + 2], g[i + 3], u[:, None])
        total += wgt * cubic(rows[:, i], rows[:, i + 1], rows[:, i + 2], rows[:, i + 3], u[None, :])
    total *= amplitude / np.abs(total).max()
    return total.astype(np.float32)


def target(s: CCausticSettings, flux_fraction: float = 0.92, n=12, seed=11) -> np.ndarray:
    """A seeded sum of n Gaussians on the wall grid, scaled to carry flux_fraction of the flux the
    slab's aperture receives (its transmittance for a flat slab)."""
    rng = np.random.default_rng(seed)
    p = s.pixels
    c = (np.arange(p) + 0.5) / p * 2 * s.wall_half - s.wall_half
    X, Y = np.meshgrid(c, c)
    img = np.zeros((p, p))
    for _ in range(n):
        cx, cy = rng.uniform(-0.8, 0.8, 2) * s.slab_half
        sig = rng.uniform(0.03, 0.12) * s.slab_half
        img += rng.uniform(0.5, 1.5) * np.exp(-((X - cx) ** 2 + (Y - cy) ** 2) / (2 * sig * sig))
    pix = 2 * s.wall_half / p
    img *= flux_fraction * (2 * s.slab_half) ** 2 / (img.sum() * pix * pix)
    return img.astype(np.float32)


class Caustic:
    """The GPU state of one caustic workload (owe_hip_caustic_*)."""

    def __init__(self, settings: CCausticSettings, device: int = 0):
        self.settings = settings
        h = ctypes.c_void_p()
        _lib.check(_lib.load().owe_hip_caustic_create(device, ctypes.byref(settings), ctypes.byref(h)))
        self._h = h

    def close(self):
        if self._h:
            _lib.check(_lib.load().owe_hip_caustic_destroy(self._h))
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def forward_into(self, heights, out, pass_index: int = 0):
        """heights, out: contiguous float32 CUDA tensors (nodes², pixels²)."""
        import torch

        stream = torch.cuda.current_stream().cuda_stream
        _lib.check(_lib.load().owe_hip_caustic_forward(self._h, heights.data_ptr(), out.data_ptr(), int(pass_index),
                                                       stream))

    def backward_into(self, heights, grad_out, grad_heights, pass_index: int = 0):
        import torch

        stream = torch.cuda.current_stream().cuda_stream
        _lib.check(_lib.load().owe_hip_caustic_backward(self._h, heights.data_ptr(), grad_out.data_ptr(),
                                                        grad_heights.data_ptr(), int(pass_index), stream))

    def irradiance(self, heights, pass_index: int = 0):
        """The wall irradiance as a differentiable function of the heights (a torch tensor)."""
        return _IrradianceFn.apply(heights, self, int(pass_index))


def _make_fn():
    import torch

    class IrradianceFn(torch.autograd.Function):
        @staticmethod
        def forward(ctx, heights, caustic, pass_index):
            s = caustic.settings
            h = heights.detach().contiguous().float()
            out = torch.empty((s.pixels, s.pixels), dtype=torch.float32, device=heights.device)
            caustic.forward_into(h, out, pass_index)
            ctx.caustic, ctx.pass_index = caustic, pass_index
            ctx.save_for_backward(h)
            return out

        @staticmethod
        def backward(ctx, grad_out):
            (h,) = ctx.saved_tensors
            g = torch.empty_like(h)
            ctx.caustic.backward_into(h, grad_out.contiguous().float(), g, ctx.pass_index)
            return g, None, None

    return IrradianceFn


class _Lazy:
    fn = None

    @classmethod
    def apply(cls, *args):
        if cls.fn is None:
            cls.fn = _make_fn()
        return cls.fn.apply(*args)


_IrradianceFn = _Lazy
