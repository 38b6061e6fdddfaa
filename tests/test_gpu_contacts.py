"""Touching bodies on the GPU: the contact rules of the world walk, against the oracle and the
reference's contact regressions (benchmarks/regressions/README.md, tests/test_far_visibility.cpp:143-200).

* a PMMA slab resting on an opaque support: no region inconsistencies or leaks, camera paths equal to
  the oracle's path by path (1e-3 on at least 99 %), the same mean as the slab lifted by 100 µm, and
  the persistent form bit-identical to the wavefront form;
* a surface sensor in the plane of an opaque cover: exactly dark under a uniform sky; raised by 100 nm
  or 1 mm it reads π within 2 %;
* light tracing onto the covered sensor from a finite disk: exactly dark; apart, lit.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

from owe_hip import loader as LD
from owe_hip import render
from pyoracle import World
from test_contacts import PARTICLES, SENSOR, SLAB

pytestmark = pytest.mark.gpu


def gpu_render(text, edits=(), spp=16, mode=None):
    ls = LD.load_scene_text(text, edits=edits)
    r = render.Renderer(ls.scene, ls.settings(integrator=ls.render.get("integrator", "path")))
    if mode is not None:
        r.set_mode(mode, use_graph=mode == render.Renderer.PERSISTENT)
    r.run_pass(spp)
    return r.resolve(), r.stats()


def test_slab_resting_on_a_support(cuda, oracle):
    img, st = gpu_render(SLAB)
    lifted, st2 = gpu_render(SLAB, ["Slab.position=(0,0,0.0501)"])
    print(f"contact {img[..., 1].mean():.6f} {st}; lifted {lifted[..., 1].mean():.6f} {st2}", flush=True)
    assert st["inconsistencies"] == 0 and st["leaks"] == 0
    assert img[..., 1].mean() == pytest.approx(lifted[..., 1].mean(), rel=0.02)
    ls = LD.load_scene_text(SLAB)
    stn = ls.settings()
    spp = 8
    pixel = np.repeat(np.arange(32 * 32, dtype=np.uint32), spp)
    sample = np.tile(np.arange(spp, dtype=np.uint32), 32 * 32)
    g = render.Renderer(ls.scene, stn).camera_paths(spp, pixel, sample)
    o, _ = World(oracle, ls.scene).camera_paths(stn, spp, pixel, sample)
    close = np.abs(g[:, 1] - o[:, 1]) <= 1e-3 * np.maximum(np.abs(o[:, 1]), np.abs(g[:, 1])) + 1e-9
    print(f"slab: {close.mean():.5f} of {len(close)} paths agree to 1e-3", flush=True)
    assert close.mean() > 0.99
    a, _ = gpu_render(SLAB, mode=render.Renderer.WAVEFRONT)
    b, _ = gpu_render(SLAB, mode=render.Renderer.PERSISTENT)
    assert np.array_equal(a, b)


def test_sensor_under_an_opaque_cover_is_dark(cuda):
    img, st = gpu_render(SENSOR, spp=64)
    print(f"covered sensor max {img.max()}, {st}", flush=True)
    assert (img == 0).all() and st["inconsistencies"] == 0
    for gap in ("0.0000001", "0.001"):
        lit, _ = gpu_render(SENSOR, [f"Screen.position=(0,0,{gap})"], spp=64)
        print(f"gap {gap}: {lit[..., 1].mean():.6f}", flush=True)
        assert lit[..., 1].mean() == pytest.approx(math.pi, rel=0.02)


def test_particles_never_reach_a_covered_sensor(cuda):
    img, _ = gpu_render(PARTICLES, spp=128)
    lit, _ = gpu_render(PARTICLES, ["Screen.position=(0,0,0.001)"], spp=128)
    print(f"particles: contact {img.max()}, separated {lit[..., 1].mean():.6g}", flush=True)
    assert (img == 0).all()
    assert lit[..., 1].mean() > 1e-3


def test_first_inconsistency_is_recorded(cuda):
    """TransportStats::first (transport_stats.hpp:15-23): a slab sunk 100 µm into its support meets
    the overlap (the support's top at z = 0 inside the slab, the slab's bottom at z = −100 µm inside
    the support); the renderer records the first such crossing with its regions and world point, and
    a scene in contact records none."""
    for edits, expect in (((), False), (["Slab.position=(0,0,0.0499)"], True)):
        ls = LD.load_scene_text(SLAB, edits=edits)
        r = render.Renderer(ls.scene, ls.settings())
        r.run_pass(16)
        first, st = r.first_inconsistency(), r.stats()
        print(f"edits {list(edits)}: {st['inconsistencies']} inconsistencies, first {first}", flush=True)
        assert (first is not None) == expect == (st["inconsistencies"] > 0)
        if first:
            assert first["ray_region"] != first["boundary_region"]
            assert first["boundary"] < len(ls.scene.boundaries)
            assert -2e-4 <= first["p"][2] <= 2e-4  # on the overlap's faces
            assert abs(np.linalg.norm(first["direction"]) - 1.0) < 1e-5
            r.reset()
            assert r.first_inconsistency() is None
        r.close()
