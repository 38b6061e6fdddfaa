"""Touching bodies and coincident surfaces: the contact rules, on the oracle (CPU).

The scenes restate the reference's contact regressions (benchmarks/regressions/README.md,
tests/test_far_visibility.cpp:143-200), written here in the `.owe` language:

* a PMMA slab resting on an opaque support (their faces share z = 0): zero region inconsistencies,
  and the same image as with the slab lifted by 100 µm within noise (the same streams: the paths that
  never meet the contact are identical) — a 100 nm lift is resolved as contact; a 100 µm overlap
  still reports inconsistencies (overlapping matter is not a contact);
* a surface sensor in the plane of an opaque cover under a uniform sky: exactly dark; raised by
  100 nm or 1 mm it reads the sky's irradiance π within 2 %;
* the same sensor under a finite emitting disk, light tracing: exactly dark at contact, lit apart;
* the contact flag: scenes whose boundaries do not touch (the dispersion prism) keep the plain walk.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

from owe_hip import loader as LD
from pyoracle import World

SLAB = """
units = m
world { sky = uniform(1) }
material wood { type = diffuse  reflectance = 0.3 }
body Support { type = box  size = (1, 1, 0.2)  position = (0, 0, -0.1)  material = wood }
body Slab { type = box  size = (0.6, 0.6, 0.1)  position = (0, 0, 0.05)  medium = PMMA }
observer Eye { position = (0, -0.35, 0.8)  look_at = (0, 0, 0)  up = (0, 1, 0)  fov = 35deg  resolution = (32, 32) }
render { detector = Eye  integrator = path  seed = 7  max_depth = 8 }
"""

SENSOR = """
units = m
world { sky = uniform(1) }
material black { type = absorber }
sensor Screen { size = (0.01, 0.01)  resolution = (4, 4)  position = (0, 0, 0)  axis = (0, 0, 1)  up = (0, 1, 0) }
body Cover { type = sheet  size = (100, 100)  position = (0, 0, 0)  axis = (0, 0, 1)  up = (0, 1, 0)  material = black }
render { detector = Screen  integrator = path  seed = 11  max_depth = 2 }
"""

PARTICLES = """
units = m
world { sky = none }
material black { type = absorber }
sensor Screen { size = (1, 1)  resolution = (4, 4)  position = (0, 0, 0)  axis = (0, 0, 1)  up = (0, 1, 0) }
body Cover { type = sheet  size = (100, 100)  position = (0, 0, 0)  axis = (0, 0, 1)  up = (0, 1, 0)  material = black }
body Source { type = disk  radius = 0.1  position = (0, 0, 0.2)  axis = (0, 0, -1)  material = black  emission = 1 }
render { detector = Screen  integrator = light  seed = 11  max_depth = 2 }
"""


def render(oracle, text, edits=(), spp=16):
    ls = LD.load_scene_text(text, edits=edits)
    integ = ls.render.get("integrator", "path")
    st = ls.settings(integrator=integ)
    w = World(oracle, ls.scene)
    img, stats = w.render_particles(st, spp) if integ == "light" else w.render(st, spp)
    return img, [int(x) for x in stats]


def test_slab_resting_on_a_support(oracle):
    img, st = render(oracle, SLAB)
    lifted, st2 = render(oracle, SLAB, ["Slab.position=(0,0,0.0501)"])
    near, st3 = render(oracle, SLAB, ["Slab.position=(0,0,0.0500001)"])
    over, st4 = render(oracle, SLAB, ["Slab.position=(0,0,0.0499)"])
    print(f"contact {img[..., 1].mean():.6f} {st}; 100 um {lifted[..., 1].mean():.6f} {st2}; 100 nm "
          f"{near[..., 1].mean():.6f} {st3}; overlap {over[..., 1].mean():.6f} {st4}")
    assert st[2] == 0 and st[3] == 0 and st3[2] == 0 and st2[2] == 0
    assert img[..., 1].mean() == pytest.approx(lifted[..., 1].mean(), rel=0.02)
    assert np.array_equal(img, near)
    assert st4[2] > 0


def test_sensor_under_an_opaque_cover_is_dark(oracle):
    img, st = render(oracle, SENSOR, spp=64)
    assert (img == 0).all() and st[2] == 0
    for gap in ("0.0000001", "0.001"):
        lit, _ = render(oracle, SENSOR, [f"Screen.position=(0,0,{gap})"], spp=64)
        print(f"gap {gap}: {lit[..., 1].mean():.6f}")
        assert lit[..., 1].mean() == pytest.approx(math.pi, rel=0.02)


def test_particles_never_reach_a_covered_sensor(oracle):
    img, st = render(oracle, PARTICLES, spp=128)
    lit, _ = render(oracle, PARTICLES, ["Screen.position=(0,0,0.001)"], spp=128)
    print(f"contact {img.max()}, separated {lit[..., 1].mean():.6g}")
    assert (img == 0).all()
    assert lit[..., 1].mean() > 1e-3


def test_contact_flag_follows_touching_boxes(oracle):
    from owe_hip import scene as S
    assert World(oracle, LD.load_scene_text(SLAB).scene).contacts
    assert World(oracle, LD.load_scene_text(SLAB, edits=["Slab.position=(0,0,0.06)"]).scene).contacts is False
    assert World(oracle, S.dispersion_prism_scene(8, 8)).contacts is False
