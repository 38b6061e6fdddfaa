"""Batches of scenes: many scenes' camera-path passes as one persistent launch.

Each member of a batch is its scene's renderer; the launch drains the members' passes one after the
other from their own path-id counters into their own films, so every member's image must equal,
bit for bit, what a persistent pass of its own renderer gives from the same (seed, pixel, sample,
pass) streams, with the same statistics — for members of different sizes, with and without
touching bodies, at both register budgets. The workload this serves: perturbed Gregorian telescopes
(the differentiable tolerancing family) rendered in their room.
"""
from __future__ import annotations

import numpy as np
import pytest

from owe_hip import _lib
from owe_hip import bodies as B
from owe_hip import lens as L
from owe_hip import render
from owe_hip import scene as S

pytestmark = pytest.mark.gpu

PRISM = dict(max_depth=16, rr_depth=4, spectral=S.SPECTRAL_UNIFORM1, lambda_min=380.0, lambda_max=780.0)


def _alone(sc, st, spp, passes, shape):
    r = render.Renderer(sc, st)
    r.set_mode(render.Renderer.PERSISTENT, **shape)
    for _ in range(passes):
        r.run_pass(spp)
    return r.resolve(), r.stats()


def _check(scenes, st, spp, passes, shape):
    b = render.Batch(scenes, st)
    b.members[0].set_mode(render.Renderer.PERSISTENT, **shape)
    for _ in range(passes):
        b.run_pass(spp)
    total = {"paths": 0, "segments": 0}
    for i, sc in enumerate(scenes):
        img, stats = b.members[i].resolve(), b.members[i].stats()
        ref, ref_stats = _alone(sc, st, spp, passes, shape)
        print(f"member {i}: {sc.observer['width']}x{sc.observer['height']}, mean Y {img[..., 1].mean():.6g} "
              f"(alone {ref[..., 1].mean():.6g}), {stats}", flush=True)
        assert img[..., 1].mean() > 0
        assert np.array_equal(img, ref)
        assert stats == ref_stats
        assert b.members[i].progress() == (passes, passes * spp)
        total["paths"] += stats["paths"]
        total["segments"] += stats["segments"]
    s = b.stats()
    assert s["paths"] == total["paths"] and s["segments"] == total["segments"]
    b.close()


@pytest.mark.parametrize("waves", [2, 3])
def test_prism_family_equals_its_members_alone(cuda, waves):
    scenes = [S.dispersion_prism_scene(48, 32, lift=0.0), S.dispersion_prism_scene(32, 32),
              S.dispersion_prism_scene(40, 24, sun_elevation=0.7, lift=0.0)]
    shape = dict(blocks_per_cu=2, chunk=4, waves_per_simd=waves, refill=16)
    print(f"prism family, {waves} waves per SIMD", flush=True)
    _check(scenes, S.settings(seed=11, **PRISM), 16, 2, shape)


def test_telescope_family_equals_its_members_alone(cuda):
    variants = L.gregorian_variants(3, rings=20)
    scenes = [B.telescope_room_scene(v, size=32, statue_frequency=24) for v in variants]
    print(f"telescope family: {len(scenes)} Gregorians in the room, {len(scenes[0].boundaries)} boundaries", flush=True)
    _check(scenes, S.settings(seed=12, max_depth=32, rr_depth=4), 64, 2,
           dict(blocks_per_cu=4, chunk=1, waves_per_simd=3, refill=48))


def test_batch_refuses_other_integrators(cuda):
    with pytest.raises(_lib.OweHipError):
        render.Batch([S.dispersion_prism_scene(8, 8)], S.settings(integrator=S.INTEGRATOR_LIGHT))
    with pytest.raises(_lib.OweHipError):
        render.Batch([], S.settings())
