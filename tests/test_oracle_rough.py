"""The oracle's rough dielectric (GGX, src/owe/core/shared/optics.hpp:110-240, sampled and evaluated
through src/owe/transport/scattering.cpp:65-188) in a scene, beside the lobe-level pins of
test_oracle_core (reference goldens of D, G1, the VNDF and both lobes; projected-area
normalisation; reciprocity).

Camera paths carry radiance and light particles carry importance: at a rough dielectric vertex
the two read the BSDF in opposite directions, and radiance crossing into glass gains the factor
(nt/ni)² that flux does not. Frosted glass closed in an emitting dome (no distant light, so light
tracing starts every path on a real surface) must render the same mean image either way.
"""
from __future__ import annotations

import numpy as np
import pytest

from owe_hip import scene as S
from pyoracle import World


@pytest.mark.parametrize("alpha", [0.1, 0.3])
def test_rough_glass_camera_paths_and_particles_agree(oracle, alpha):
    w = World(oracle, S.rough_glass_scene(alpha, size=16, dome=True))
    cam, cst = w.render(S.settings(seed=1, max_depth=32, rr_depth=6), 1024, layout=0)
    ys = [w.render_particles(S.settings(seed=s, max_depth=32, rr_depth=6, integrator=S.INTEGRATOR_LIGHT), 16384)[0]
          [..., 1].mean() for s in range(1, 9)]
    m, e = np.mean(ys), np.std(ys, ddof=1) / np.sqrt(len(ys))
    assert cam[..., 1].mean() == pytest.approx(m, abs=max(4 * e, 0.005 * m))
    assert e < 0.02 * m
    assert cst[2] == 0 and cst[3] == 0
