"""Visibility of sources a million metres away (the reference's far-visibility regressions,
tests/test_far_visibility.cpp:63-141), on the oracle: a blocker in front of a far emitter makes the
receiver exactly dark; moved aside it lets the light through; moved 5 cm (NEE) or 50 cm (particles)
behind the emitter it changes nothing, sample for sample.
"""
from __future__ import annotations

import numpy as np
import pytest

from owe_hip import loader as LD
from pyoracle import World

NEE = """
units = m
world { up = (0, 1, 0)  sky = none }
material white { type = diffuse  reflectance = 0.8 }
material black { type = absorber }
body Receiver { type = sheet  size = (0.4, 0.4)  position = (0, 0, 0)  axis = (0, 0, 1)  up = (0, 1, 0)  material = white }
body FarEmitter { type = sheet  size = (0.1, 0.1)  position = (0, 0, 1000000)  axis = (0, 0, -1)  up = (0, 1, 0)
                  material = black  emission = 1e12 }
body SourceBlocker { type = sheet  size = (0.4, 0.4)  position = (0, 0, 999999.95)  axis = (0, 0, -1)  up = (0, 1, 0)
                     material = black }
observer Eye { position = (0, -0.1, 0.3)  look_at = (0, 0, 0)  up = (0, 1, 0)  fov = 20deg  resolution = (16, 16) }
render { detector = Eye  integrator = path  max_depth = 2  seed = 11 }
"""

SOURCE = """
units = m
world { up = (0, 1, 0)  sky = none }
material black { type = absorber }
body FarEmitter { type = sheet  size = (0.1, 0.1)  position = (0, 0, 1000000)  axis = (0, 0, -1)  up = (0, 1, 0)
                  material = black  emission = 1 }
body SourceBlocker { type = sheet  size = (0.4, 0.4)  position = (0, 0, 999999.5)  axis = (0, 0, -1)  up = (0, 1, 0)
                     material = black }
observer Eye { position = (0, 0, 0)  look_at = (0, 0, 1000000)  up = (0, 1, 0)  fov = 0.00001deg  pupil = 1mm
               focus = 1000000  resolution = (16, 16) }
render { detector = Eye  integrator = light  max_depth = 3  seed = 11 }
"""

CASES = {"nee": (NEE, "path", 128, "SourceBlocker.position=(1,0,999999.95)", "SourceBlocker.position=(0,0,1000000.05)"),
         "particles": (SOURCE, "light", 128, "SourceBlocker.position=(1,0,999999.5)", "SourceBlocker.position=(0,0,1000000.5)")}


def run(oracle, text, integ, spp, edits=()):
    ls = LD.load_scene_text(text, edits=edits)
    st = ls.settings(integrator=integ)
    w = World(oracle, ls.scene)
    img, s = w.render_particles(st, spp) if integ == "light" else w.render(st, spp)
    assert int(s[2]) == 0 and int(s[3]) == 0
    return img


@pytest.mark.parametrize("case", sorted(CASES))
def test_far_sources_respect_blockers_near_them(oracle, case):
    text, integ, spp, aside, behind = CASES[case]
    blocked = run(oracle, text, integ, spp)
    open_ = run(oracle, text, integ, spp, [aside])
    beyond = run(oracle, text, integ, spp, [behind])
    print(f"{case}: blocked {blocked[..., 1].mean():.9g}, open {open_[..., 1].mean():.9g}, beyond {beyond[..., 1].mean():.9g}")
    assert (blocked == 0).all()
    assert open_[..., 1].mean() > 1e-6
    assert np.allclose(beyond, open_, rtol=1e-6, atol=1e-12)
