"""Sources a million metres away on the GPU (tests/test_far_visibility.cpp:63-141 restated): in
float32 camera-relative coordinates a metre at 10⁶ m is sixteen ulps. A blocker in front of the far
emitter makes the receiver exactly dark and moved aside lets the light through; a blocker behind
the emitter changes nothing within 1e-6 — 50 cm behind for particles; the NEE case's 5 cm (under one
ulp of 10⁶ in float32) is checked on the oracle only.
"""
from __future__ import annotations

import numpy as np
import pytest

from owe_hip import loader as LD
from owe_hip import render
from test_far_visibility import CASES

pytestmark = pytest.mark.gpu


def run(text, integ, spp, edits=()):
    ls = LD.load_scene_text(text, edits=edits)
    r = render.Renderer(ls.scene, ls.settings(integrator=integ))
    r.run_pass(spp)
    s = r.stats()
    assert s["inconsistencies"] == 0 and s["leaks"] == 0
    return r.resolve()


@pytest.mark.parametrize("case", sorted(CASES))
def test_far_sources_respect_blockers_near_them(cuda, case):
    text, integ, spp, aside, behind = CASES[case]
    blocked = run(text, integ, spp)
    open_ = run(text, integ, spp, [aside])
    print(f"{case}: blocked {blocked[..., 1].mean():.9g}, open {open_[..., 1].mean():.9g}", flush=True)
    assert (blocked == 0).all()
    assert open_[..., 1].mean() > 1e-6
    if case == "particles":
        beyond = run(text, integ, spp, [behind])
        print(f"{case}: beyond {beyond[..., 1].mean():.9g}", flush=True)
        assert np.allclose(beyond, open_, rtol=1e-6, atol=1e-12)
