"""The persistent batched form of the camera loop against the plain wavefront form.

Both forms run the same device functions on the same (seed, pixel, sample, pass) streams, and the
film is integer, so the images must be equal bit for bit (the batched form's parity bar: equal to
the plain kernel per item for the same seeds), with equal statistics. Also: HIP-graph replay of
the persistent pass, replayed several times with eager work between replays, against eager passes;
and the autotuner, which must leave no trace in the film.
"""
from __future__ import annotations

import numpy as np
import pytest

from owe_hip import render
from owe_hip import scene as S

pytestmark = pytest.mark.gpu

PRISM = dict(max_depth=16, rr_depth=4, spectral=S.SPECTRAL_UNIFORM1, lambda_min=380.0, lambda_max=780.0)


def _sky_prism():
    sc = S.dispersion_prism_scene(64, 64)
    sc.set_uniform_sky(S.Spectrum.constant(0.3))
    return sc


CASES = {
    "prism-uniform1": (lambda: S.dispersion_prism_scene(64, 64), S.settings(seed=3, **PRISM), 16),
    "prism-sky-hero4": (_sky_prism, S.settings(seed=4, max_depth=16, rr_depth=4), 8),
    "furnace": (lambda: S.white_furnace_scene(0.8), S.settings(seed=5), 16),
    "ball-caustic": (lambda: S.ball_caustic_scene(32), S.settings(seed=6), 16),
    "submerged-sensor": (lambda: S.submerged_sensor_scene(8), S.settings(seed=7), 64),
}


def render_with(sc, st, spp, passes, mode, **kw):
    r = render.Renderer(sc, st)
    if mode:
        r.set_mode(render.Renderer.PERSISTENT, **kw)
    for _ in range(passes):
        r.run_pass(spp)
    img, stats = r.resolve(), r.stats()
    r.close()
    return img, stats


@pytest.mark.parametrize("waves,refill", [(2, 1), (3, 1), (2, 32)])
@pytest.mark.parametrize("case", sorted(CASES))
def test_persistent_form_equals_wavefront_bit_for_bit(cuda, case, waves, refill):
    # Both register budgets of the persistent kernel (3 waves per SIMD spills registers) and a
    # refill threshold (finished lanes wait for half the wave) compute every path exactly as the
    # wavefront kernels do.
    make, st, spp = CASES[case]
    sc = make()
    a, sa = render_with(sc, st, spp, 2, False)
    b, sb = render_with(sc, st, spp, 2, True, blocks_per_cu=2, chunk=4, waves_per_simd=waves, refill=refill)
    diff = np.abs(a - b)
    print(f"{case} ({waves} waves/SIMD, refill {refill}): max |diff| {diff.max():.3e}, pixels differing "
          f"{(diff > 0).any(axis=2).sum()}, stats {sa} / {sb}", flush=True)
    assert np.array_equal(a, b)
    assert sa == sb


@pytest.mark.parametrize("shape", [(1, 1), (3, 16), (8, 2)])
def test_persistent_launch_shape_does_not_change_the_image(cuda, shape):
    sc = S.dispersion_prism_scene(64, 64)
    st = S.settings(seed=9, **PRISM)
    a, _ = render_with(sc, st, 8, 1, True, blocks_per_cu=2, chunk=4)
    b, _ = render_with(sc, st, 8, 1, True, blocks_per_cu=shape[0], chunk=shape[1])
    assert np.array_equal(a, b)


def test_graph_replays_equal_eager_passes(cuda):
    sc = S.dispersion_prism_scene(64, 64)
    st = S.settings(seed=10, **PRISM)
    eager = render.Renderer(sc, st)
    eager.set_mode(render.Renderer.PERSISTENT, blocks_per_cu=2, chunk=4)
    graph = render.Renderer(sc, st)
    graph.set_mode(render.Renderer.PERSISTENT, blocks_per_cu=2, chunk=4, use_graph=True)
    for k in range(4):
        eager.run_pass(8)
        graph.run_pass(8)
        # eager work of the program between replays
        other = render.Renderer(S.white_furnace_scene(0.5, size=8), S.settings(seed=k))
        other.run_pass(4)
        other.close()
        ea, ga = eager.resolve(), graph.resolve()
        print(f"replay {k}: equal {np.array_equal(ea, ga)}", flush=True)
        assert np.array_equal(ea, ga)
        assert eager.stats() == graph.stats()


def test_autotune_picks_a_shape_and_leaves_no_trace(cuda):
    sc = S.dispersion_prism_scene(64, 64)
    st = S.settings(seed=11, **PRISM)
    ref, sref = render_with(sc, st, 16, 1, True, blocks_per_cu=2, chunk=4)
    r = render.Renderer(sc, st)
    r.set_mode(render.Renderer.PERSISTENT)
    ms = r.autotune(16)
    m = r.mode()
    print(f"autotune: {m}, best pass {ms:.3f} ms", flush=True)
    assert ms > 0 and m["mode"] == render.Renderer.PERSISTENT
    assert not r.resolve().any() and r.stats()["paths"] == 0
    r.run_pass(16)
    assert np.array_equal(r.resolve(), ref)
    w = render.Renderer(sc, st)
    assert w.autotune(16) > 0 and w.mode()["mode"] == render.Renderer.WAVEFRONT
    w.run_pass(16)
    assert np.array_equal(w.resolve(), ref)
