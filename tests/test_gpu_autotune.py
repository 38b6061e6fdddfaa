"""The autotuner: two-stage timing (every shape once, the fastest again), the wavefront's
paths-in-flight knob, and the per-device tuning cache (a hit applies the recorded shape without a
sweep; another scene or pass size misses). Tuned shapes change only the launch, never the film."""
from __future__ import annotations

import numpy as np
import pytest

from owe_hip import render
from owe_hip import scene as S

pytestmark = pytest.mark.gpu

PRISM = dict(max_depth=16, rr_depth=4, spectral=S.SPECTRAL_UNIFORM1, lambda_min=380.0, lambda_max=780.0)


def test_tuning_cache_round_trip(cuda, tmp_path):
    path = tmp_path / "tuning.txt"
    render.set_tuning_cache(path)
    try:
        sc, st = S.dispersion_prism_scene(64, 64), S.settings(seed=3, **PRISM)
        a = render.Renderer(sc, st)
        a.set_mode(render.Renderer.PERSISTENT, use_graph=True)
        ms = a.autotune(16)
        shape = a.mode()
        assert not a.tuning()["from_cache"] and ms > 0
        assert len(path.read_text().splitlines()) == 1
        b = render.Renderer(sc, st)
        b.set_mode(render.Renderer.PERSISTENT, use_graph=True)
        ms2 = b.autotune(16)
        print(f"sweep {ms:.4f} ms, cache hit {ms2:.4f} ms, shape {shape}", flush=True)
        assert b.tuning()["from_cache"] and b.mode() == shape and ms2 == pytest.approx(ms, rel=1e-5)
        c = render.Renderer(S.dispersion_prism_scene(48, 48), st)
        c.set_mode(render.Renderer.PERSISTENT, use_graph=True)
        c.autotune(16)
        assert not c.tuning()["from_cache"]
        assert len(path.read_text().splitlines()) == 2
        # The tuned and the untuned renderers draw the same film.
        for r in (a, b):
            r.run_pass(16)
        plain = render.Renderer(sc, st)
        plain.set_mode(render.Renderer.PERSISTENT)
        plain.run_pass(16)
        assert np.array_equal(a.resolve(), plain.resolve()) and np.array_equal(b.resolve(), plain.resolve())
    finally:
        render.set_tuning_cache(None)


def test_wavefront_tunes_paths_in_flight(cuda):
    sc, st = S.dispersion_prism_scene(64, 64), S.settings(seed=4, **PRISM)
    w = render.Renderer(sc, st)
    full = w.tuning()["queue_capacity"]
    w.autotune(16)
    t = w.tuning()
    print(f"wavefront tuning {t} of {full} slots", flush=True)
    assert not t["from_cache"] and 1024 <= t["queue_capacity"] <= full and t["rounds_per_sync"] in (1, 2, 4, 8)
    w.run_pass(16)
    ref = render.Renderer(sc, st)
    ref.run_pass(16)
    assert np.array_equal(w.resolve(), ref.resolve())
