"""The particle pass's lane-refill form against its one-thread-per-particle form.

Both run the same device code per particle on the same (seed, index, pass) streams, and the
particle film is exact integer bins, so every pass's image must be equal bit for bit, with equal
statistics — across shapes (refill threshold, blocks per compute unit, chunk), scenes with a
sensor, an observer, media, contacts and hero or single wavelengths, and several passes. The
autotuner's choice for this form must also leave no trace in the film.
"""
from __future__ import annotations

import numpy as np
import pytest

from owe_hip import render
from owe_hip import scene as S
from test_gpu_particles import CASES, light

pytestmark = pytest.mark.gpu

MORE = {
    "prism-standing": (lambda: S.dispersion_prism_scene(48, 48, lift=0.0),
                       dict(seed=21, max_depth=16, rr_depth=4, spectral=S.SPECTRAL_UNIFORM1, lambda_min=380.0,
                            lambda_max=780.0)),
    "fog-bubble": (lambda: S.scattering_furnace_scene(3.0, 0.3, 0.8, 0.5, size=8, bubble=0.3),
                   dict(seed=22, max_depth=1000, rr_depth=6)),
    "rough-glass": (lambda: S.rough_glass_scene(0.2, size=16), dict(seed=23, max_depth=32, rr_depth=6)),
}


def _render(sc, st, ppp, passes, refill=None):
    r = render.Renderer(sc, st)
    if refill is not None:
        r.set_mode(render.Renderer.PERSISTENT, **refill)
    for _ in range(passes):
        r.run_pass(ppp)
    return r.resolve(), r.stats()


@pytest.mark.parametrize("case", sorted(CASES) + sorted(MORE))
@pytest.mark.parametrize("shape", [dict(blocks_per_cu=8, chunk=4, refill=32), dict(blocks_per_cu=2, chunk=1, refill=1),
                                   dict(blocks_per_cu=16, chunk=16, refill=48)])
def test_refill_form_equals_one_thread_per_particle(cuda, case, shape):
    make, kw = {**CASES, **MORE}[case]
    sc, st = make(), light(**kw)
    ppp = 64 if sc.observer["width"] * sc.observer["height"] <= 4096 else 16
    a, sa = _render(sc, st, ppp, 2)
    b, sb = _render(sc, st, ppp, 2, shape)
    print(f"{case} {shape}: mean Y {a[..., 1].mean():.6g} / {b[..., 1].mean():.6g}; {sa} / {sb}", flush=True)
    assert a[..., 1].mean() > 0
    assert np.array_equal(a, b)
    assert sa == sb


def test_refill_form_autotune_leaves_no_trace(cuda):
    sc = S.dispersion_prism_scene(64, 64)
    st = light(seed=31, max_depth=16, rr_depth=4, spectral=S.SPECTRAL_UNIFORM1, lambda_min=380.0, lambda_max=780.0)
    r = render.Renderer(sc, st)
    r.set_mode(render.Renderer.PERSISTENT)
    ms = r.autotune(16)
    shape = r.mode()
    print(f"tuned particle refill shape {shape}: {ms:.3f} ms per pass", flush=True)
    assert ms > 0 and r.stats()["paths"] == 0 and r.progress() == (0, 0)
    r.run_pass(16)
    ref, _ = _render(sc, st, 16, 1)
    assert np.array_equal(r.resolve(), ref)
