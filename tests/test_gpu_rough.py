"""The rough dielectric lobe (GGX, src/owe/core/shared/optics.hpp:110-240 through
src/owe/transport/scattering.cpp:65-188) on the GPU: frosted glass in sunlight, seen directly, by
next-event estimation at every rough vertex (the reflection lobe toward the sun) and through the
glass (sampled transmission), for camera paths, particles and SPPM.

* per path and per particle, from the same streams, the float kernels and the fp64 oracle agree
  except where a float decision (a roulette draw, the GGX branch choice) falls the other way;
* camera paths and particles render the same mean image in the emitting dome (radiance and
  importance read the BSDF in opposite directions; only radiance takes the (nt/ni)² factor).
"""
from __future__ import annotations

import numpy as np
import pytest

from owe_hip import render
from owe_hip import scene as S
from pyoracle import World

pytestmark = pytest.mark.gpu

ROUGH = dict(max_depth=32, rr_depth=6)


@pytest.mark.parametrize("alpha,spectral", [(0.1, S.SPECTRAL_HERO4), (0.3, S.SPECTRAL_UNIFORM1)])
def test_rough_glass_paths_match_the_oracle_path_by_path(cuda, oracle, alpha, spectral):
    sc = S.rough_glass_scene(alpha, size=32)
    st = S.settings(seed=13, spectral=spectral, **ROUGH)
    spp = 16
    pixel = np.repeat(np.arange(32 * 32, dtype=np.uint32), spp)
    sample = np.tile(np.arange(spp, dtype=np.uint32), 32 * 32)
    gpu = render.Renderer(sc, st).camera_paths(spp, pixel, sample)
    ora, segs = World(oracle, sc).camera_paths(st, spp, pixel, sample)
    gy, oy = gpu[:, 1], ora[:, 1]
    close = np.abs(gy - oy) <= 1e-3 * np.maximum(np.abs(oy), np.abs(gy)) + 1e-6 * np.abs(oy).max()
    print(f"rough glass alpha={alpha}: {close.mean():.5f} of {len(gy)} paths within 1e-3; mean Y gpu {gy.mean():.5f} "
          f"oracle {oy.mean():.5f}; mean segments {segs.mean():.2f}", flush=True)
    assert close.mean() > 0.99
    assert gy.mean() == pytest.approx(oy.mean(), rel=0.02)


@pytest.mark.parametrize("alpha", [0.1, 0.3])
def test_rough_glass_particles_match_the_oracle_particle_by_particle(cuda, oracle, alpha):
    sc = S.rough_glass_scene(alpha, size=16, dome=True)
    st = S.settings(seed=14, integrator=S.INTEGRATOR_LIGHT, **ROUGH)
    idx = np.arange(1 << 17, dtype=np.uint64)
    gpu = render.Renderer(sc, st).particles(idx)
    ora, segs = World(oracle, sc).particles(st, idx)
    gy, oy = gpu[:, 1], ora[:, 1]
    live = (gy != 0) | (oy != 0)
    close = np.abs(gy - oy) <= 1e-3 * np.maximum(np.abs(oy), np.abs(gy))
    frac = close[live].mean()
    tot = abs(gy.sum() - oy.sum()) / abs(oy.sum())
    print(f"rough glass particles alpha={alpha}: {live.sum()} contributing, {frac:.5f} agree to 1e-3, totals differ "
          f"by {tot:.2e}", flush=True)
    assert live.sum() > 1000
    assert frac >= 0.98
    assert tot < 2e-2


def test_rough_glass_camera_paths_and_particles_agree(cuda):
    sc = S.rough_glass_scene(0.2, size=16, dome=True)
    cam = render.Renderer(sc, S.settings(seed=15, **ROUGH))
    cam.run_pass(4096)
    ys = []
    for s in range(8):
        r = render.Renderer(sc, S.settings(seed=16 + s, integrator=S.INTEGRATOR_LIGHT, **ROUGH))
        r.run_pass(65536)
        ys.append(r.resolve()[..., 1].mean())
    m, e = np.mean(ys), np.std(ys, ddof=1) / np.sqrt(len(ys))
    c = cam.resolve()[..., 1].mean()
    print(f"rough glass dome: camera {c:.5f}, particles {m:.5f} ± {e:.5f}", flush=True)
    assert c == pytest.approx(m, abs=max(4 * e, 0.005 * m))


def test_rough_glass_sppm_first_iteration_matches_the_oracle(cuda, oracle):
    sc = S.rough_glass_scene(0.2, size=32)
    st = S.settings(seed=17, integrator=S.INTEGRATOR_SPPM, photons_per_pixel=32.0, **ROUGH)
    r = render.Renderer(sc, st)
    r.run_pass(1)
    g = r.sppm_state()
    o = World(oracle, sc).sppm(st, 1)
    seen = (o["r"] > 0) & (g["r"] > 0)
    rclose = (np.abs(g["r"] - o["r"]) <= 1e-4 * o["r"])[seen].mean()
    nclose = (np.abs(g["N"] - o["N"]) <= 1e-6 * np.maximum(o["N"], 1.0)).mean()
    tclose = (np.abs(g["tau"][..., 1] - o["tau"][..., 1]) <= 1e-3 * np.maximum(o["tau"][..., 1], 1e-30) + 1e-12).mean()
    print(f"rough glass sppm: {seen.mean():.3f} seen; radius agrees on {rclose:.4f}, N on {nclose:.4f}, "
          f"tau on {tclose:.4f}", flush=True)
    assert seen.mean() > 0.5
    assert rclose > 0.99 and nclose > 0.95 and tclose > 0.95
