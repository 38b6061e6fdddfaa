"""Scenes read from `.owe` files render on the GPU as the oracle renders them.

The scene-file benchmark's cases (owe_hip/scenes: the reference's five benchmark scenes — the prism
in low sunlight between black cards, the lens caustic seen through a tilted plate and in a flat
mirror, the two slabs behind a stop, the afocal refractor in its tube read from a `.lens` file with
the eye at its exit pupil — and the package's rippled pool) are loaded through owe_hip.loader and
rendered with camera paths at reduced resolution:

* per path: the same (seed, pixel, sample, pass) camera streams through the float32 kernels and the
  fp64 oracle agree to 1e-3 on at least 99 % of paths (the rest are float decisions that fall the
  other way), with no region inconsistencies or leaks on the GPU;
* the persistent form renders each scene bit-identically to the wavefront form, and its digest (the
  ``owe bench`` FNV-1a of the raw XYZ) is therefore the same.
"""
from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest

from owe_hip import loader as LD
from owe_hip import output as O
from owe_hip import render
from pyoracle import World

pytestmark = pytest.mark.gpu

SCENES = Path(LD.__file__).parent / "scenes"
# Each scene at 1/(2k) of its own resolution.
CASES = {"dispersion_prism": 4, "caustic_seen_through_glass": 2, "caustic_seen_in_mirror": 2, "pool": 3,
         "double_refraction_slabs": 3, "telescope_exit_pupil_walk": 1}


def _scene(name):
    ls = LD.load_scene(str(SCENES / f"{name}.owe"))
    w, h = ls.detectors[ls.detector_index()]["resolution"]
    k = CASES[name]
    return ls, ls.use(resolution=(w // k // 2, h // k // 2))


@pytest.mark.parametrize("name", sorted(CASES))
def test_loaded_scene_paths_match_the_oracle(cuda, oracle, name):
    ls, sc = _scene(name)
    st = ls.settings(integrator="path", seed=3)
    w, h = sc.observer["width"], sc.observer["height"]
    spp = 8
    pixel = np.repeat(np.arange(w * h, dtype=np.uint32), spp)
    sample = np.tile(np.arange(spp, dtype=np.uint32), w * h)
    print(f"{name}: {w}x{h}, {spp} spp, {len(sc.boundaries)} boundaries", flush=True)
    r = render.Renderer(sc, st)
    gpu = r.camera_paths(spp, pixel, sample)
    ora, _ = World(oracle, sc).camera_paths(st, spp, pixel, sample)
    gy, oy = gpu[:, 1], ora[:, 1]
    close = np.abs(gy - oy) <= 1e-3 * np.maximum(np.abs(oy), np.abs(gy)) + 1e-9 * max(1.0, float(np.abs(oy).max()))
    print(f"{name}: {close.mean():.5f} of {len(gy)} paths agree to 1e-3; mean Y gpu {gy.mean():.6g} oracle "
          f"{oy.mean():.6g}; lit paths {(oy > 0).mean():.3f}", flush=True)
    assert (oy > 0).mean() > 0.001
    assert close.mean() > 0.99
    r.run_pass(spp)
    s = r.stats()
    assert s["inconsistencies"] == 0 and s["leaks"] == 0, s


@pytest.mark.parametrize("name", sorted(CASES))
def test_loaded_scene_forms_are_bit_identical(cuda, name):
    ls, sc = _scene(name)
    st = ls.settings(integrator="path", seed=5)
    digests = []
    imgs = []
    for mode in (render.Renderer.WAVEFRONT, render.Renderer.PERSISTENT):
        r = render.Renderer(sc, st)
        r.set_mode(mode, use_graph=mode == render.Renderer.PERSISTENT)
        r.run_pass(32)
        img = r.resolve()
        imgs.append(img)
        digests.append(O.image_digest(img))
    print(f"{name}: digests {digests}, mean Y {imgs[0][..., 1].mean():.6g}", flush=True)
    assert np.array_equal(imgs[0], imgs[1])
    assert digests[0] == digests[1]


def test_suite_digests_are_reproducible(cuda):
    from owe_hip import suite

    runs = [suite.run(spp=4, resolution=(64, 48), mode=m) for m in
            (render.Renderer.PERSISTENT, render.Renderer.PERSISTENT, render.Renderer.WAVEFRONT)]
    for rows in zip(*(r["rows"] for r in runs)):
        print(f"{Path(rows[0]['scene']).name}: {rows[0]['digest']} {rows[0]['mpaths_per_s']:.1f} M paths/s", flush=True)
        assert rows[0]["digest"] == rows[1]["digest"] == rows[2]["digest"]
        assert rows[0]["paths"] == 64 * 48 * 4 and rows[0]["inconsistencies"] == 0
    print(suite.format_table(runs[0]), flush=True)
