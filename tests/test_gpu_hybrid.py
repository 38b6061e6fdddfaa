"""The hybrid integrator on the GPU (camera paths and particles with the disjoint partition of
particle.slang:358-417 and transport.slang:745-906: light tracing owns eye → D → S⁺ → light):

* per camera path and per particle, from the same streams, the float kernels and the fp64 oracle
  agree with the partition applied (the reference's caustic agreement scene, the ball caustic);
* the furnace with a glass ball stays at 1/(1 − ρ) (test_backends.cpp:540-574: every integrator);
* both launch forms give the same image bit for bit, and the image matches the oracle's hybrid
  image from the same streams;
* a surface sensor is refused, as the reference refuses it (renderer.hpp:34-35).
"""
from __future__ import annotations

import numpy as np
import pytest

from owe_hip import _lib, render
from owe_hip import loader as LD
from owe_hip import scene as S
from pyoracle import World

from test_oracle_hybrid import CAUSTIC_TEXT

pytestmark = pytest.mark.gpu


def _caustic():
    ls = LD.load_scene_text(CAUSTIC_TEXT)
    return ls.use(), ls.settings(integrator="hybrid", seed=21, max_depth=16)


def _ball():
    return S.ball_caustic_scene(size=16), S.settings(seed=22, integrator=S.INTEGRATOR_HYBRID, max_depth=32)


@pytest.mark.parametrize("case", ["caustic", "ball"])
def test_camera_paths_match_the_oracle_path_by_path(cuda, oracle, case):
    sc, st = _caustic() if case == "caustic" else _ball()
    w, h = sc.observer["width"], sc.observer["height"]
    spp = 16
    pixel = np.repeat(np.arange(w * h, dtype=np.uint32), spp)
    sample = np.tile(np.arange(spp, dtype=np.uint32), w * h)
    print(f"hybrid camera paths, {case}: {len(pixel)} paths", flush=True)
    gpu = render.Renderer(sc, st).camera_paths(spp, pixel, sample)
    ora, segs = World(oracle, sc).camera_paths(st, spp, pixel, sample)
    gy, oy = gpu[:, 1], ora[:, 1]
    close = np.abs(gy - oy) <= 1e-3 * np.maximum(np.abs(oy), np.abs(gy)) + 1e-6 * max(float(np.abs(oy).max()), 1e-30)
    print(f"  {close.mean():.5f} within 1e-3; mean Y gpu {gy.mean():.6g} oracle {oy.mean():.6g}", flush=True)
    assert close.mean() > 0.995


@pytest.mark.parametrize("case", ["caustic", "ball"])
def test_particles_match_the_oracle_particle_by_particle(cuda, oracle, case):
    sc, st = _caustic() if case == "caustic" else _ball()
    idx = np.arange(1 << 17, dtype=np.uint64)
    print(f"hybrid particles, {case}: {len(idx)}", flush=True)
    gpu = render.Renderer(sc, st).particles(idx)
    ora, segs = World(oracle, sc).particles(st, idx)
    gy, oy = gpu[:, 1], ora[:, 1]
    live = (gy != 0) | (oy != 0)
    close = np.abs(gy - oy) <= 1e-3 * np.maximum(np.abs(oy), np.abs(gy))
    print(f"  {live.sum()} contributing, {close[live].mean():.5f} agree to 1e-3; mean segments {segs.mean():.2f}",
          flush=True)
    assert live.sum() > 100
    assert close[live].mean() >= 0.98


def test_furnace_with_a_glass_ball(cuda):
    rho = 0.5
    sc = S.white_furnace_scene(rho, size=16, glass=0.4)
    print("hybrid furnace with a glass ball: 16x16, 2048 spp", flush=True)
    r = render.Renderer(sc, S.settings(seed=23, integrator=S.INTEGRATOR_HYBRID, max_depth=200))
    r.run_pass(2048)
    y = r.resolve()[..., 1].mean()
    s = r.stats()
    print(f"  {y:.5f} (expected 2); {s}", flush=True)
    assert y == pytest.approx(1 / (1 - rho), rel=0.01)
    assert s["leaks"] == 0 and s["inconsistencies"] == 0


def test_forms_agree_bit_for_bit_and_with_the_oracle(cuda, oracle):
    sc, st = _caustic()
    images = []
    for mode in (render.Renderer.WAVEFRONT, render.Renderer.PERSISTENT):
        r = render.Renderer(sc, st)
        r.set_mode(mode)
        r.run_pass(64)
        r.run_pass(64)
        images.append(r.resolve())
        r.close()
    assert np.array_equal(images[0], images[1])
    # The oracle from the same streams: its camera passes and its particle passes, summed.
    w = World(oracle, sc)
    cam, _ = w.render(st, 64, passes=2, layout=1)
    lig, _ = w.render_particles(st, 64, passes=2)
    g, o = images[0][..., 1], (cam + lig)[..., 1]
    close = np.abs(g - o) <= 1e-3 * np.maximum(o, 1e-30) + 1e-6 * o.max()
    print(f"hybrid, both forms: mean Y {g.mean():.6g}, oracle {o.mean():.6g}; {close.mean():.4f} of pixels within 1e-3",
          flush=True)
    assert g.mean() == pytest.approx(o.mean(), rel=2e-3)
    assert close.mean() > 0.95


def test_sensors_are_refused(cuda):
    with pytest.raises(_lib.OweHipError):
        render.Renderer(S.disk_irradiance_scene(screen=True), S.settings(integrator=S.INTEGRATOR_HYBRID))
