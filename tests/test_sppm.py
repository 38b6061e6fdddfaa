"""The oracle's SPPM (stochastic progressive photon mapping, the reference's sppm integrator,
src/owe/backends/gpu/shaders/sppm*.slang) against the reference's own SPPM cases.

The reference implements SPPM only in its float GPU shaders; the oracle restates it in double
precision (oracle_tracer.cpp: the camera pass's visible-point rule in Tracer::radiance, photons in
Tracer::particleWalk, a point's value in Tracer::sppmValue; oracle_abi_scene.cpp: iterations,
reach, update) and is pinned by:

* the white furnace, L = Le / (1 − ρ) within 2 % for ρ ∈ {0, 0.5} at 8×8 pixels and 4096
  iterations (tests/test_backends.cpp:540-574): the camera pass's direct part (Le + ρLe) and the
  photons' part (ρ²Le / (1 − ρ)) must both be right; every visible point there is seen directly;
* a lens caustic seen through a glass plate and in a mirror (tests/test_references.cpp:212-250):
  the ratio to the direct view is the plate's ȳ-weighted (1 − R)²/(1 − R²) and the mirror's 0.9,
  within 0.015, at 48×48 pixels and 2048 iterations. These points are seen through glass or a
  mirror, bound to their wavelength group: a group share not scaled back up would be off by the
  number of groups;
* the initial radius in closed form (GpuRenderer::resetSppm, gpu_renderer_passes.cpp:238-257);
* photon guiding (particle.slang:149-232) learns only the useful patch of the sun's emission disc,
  and its state carries across calls like the rest.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

from owe_hip import caustic_scenes as C
from owe_hip import scene as S
from pyoracle import World

SPPM = S.INTEGRATOR_SPPM


@pytest.mark.parametrize("rho", [0.0, 0.5])
def test_sppm_white_furnace(oracle, rho):
    w = World(oracle, S.white_furnace_scene(rho, size=8))
    st = w.sppm(S.settings(seed=1, integrator=SPPM), 4096)
    img = World.sppm_resolve(st)
    assert img[..., 1].mean() == pytest.approx(1 / (1 - rho), abs=0.02 / (1 - rho))
    direct = st["direct"][..., 1].mean() / st["iterations"]
    assert direct == pytest.approx(1 + rho, rel=0.01)  # emission seen, and the direct light at the point
    assert st["stats"][2] == 0 and st["stats"][3] == 0
    if rho == 0:
        assert np.all(st["N"] == 0)  # no photon survives its first surface: nothing to gather


def test_sppm_sees_a_caustic_through_a_plate_and_in_a_mirror(oracle):
    means = {}
    for name, ls in C.caustic_cases().items():
        sc = ls.use()
        st = World(oracle, sc).sppm(ls.settings(), 2048)
        means[name] = World.sppm_resolve(st)[..., 1].mean()
        assert st["stats"][2] == 0 and st["stats"][3] == 0
    plate, mirror = means["plate"] / means["direct"], means["mirror"] / means["virtual"]
    print(f"through the plate {plate:.4f} (Fresnel {C.plate_transmission():.4f}); in the mirror {mirror:.4f}")
    assert plate == pytest.approx(C.plate_transmission(), abs=0.015)
    assert mirror == pytest.approx(0.9, abs=0.015)


def test_sppm_initial_radius(oracle):
    st = S.settings(integrator=SPPM, photon_radius=3.0)
    sc = S.white_furnace_scene(0.5, size=8)
    r0, floor = World(oracle, sc).sppm_radius(st)
    assert r0 == pytest.approx(3.0 * 2 * math.tan(math.radians(35)) / 8)
    assert floor == 1e-9
    # A sensor aimed at a disc sees a pixel from the disc's centre; a bare one takes a fixed radius.
    for aimed in (True, False):
        sc = S.disk_irradiance_scene(0.3, 0.5, 0.15, aimed=aimed)
        w = World(oracle, sc)
        r0, floor = w.sppm_radius(st)
        pixel = 2 * 0.15 / 2  # the sensor's half-size over its pixels; the disc's centre is 0.5 above it
        if aimed:
            assert r0 == pytest.approx(3.0 * pixel / 0.5) and floor == 1e-9
        else:
            assert r0 == 0 and floor == pytest.approx(3.0 * pixel)


def test_sppm_sensor_measures_disk_irradiance(oracle):
    """A surface sensor aimed at a Lambertian disc (tests/analytic.hpp:28-60): every camera path meets
    the emitter, so the camera pass alone carries the irradiance, within 1 %."""
    a, h, half = 0.3, 0.5, 0.15
    w = World(oracle, S.disk_irradiance_scene(a, h, half, aimed=True))
    st = w.sppm(S.settings(seed=1, integrator=SPPM), 20000)
    assert World.sppm_resolve(st)[..., 1].mean() == pytest.approx(S.mean_disk_irradiance(2.0, a, h, half), rel=0.01)
    assert np.all(st["r"] == 0)  # no visible points: the sensor sees only the emitter


def test_sppm_photon_guide_learns_the_lens(oracle):
    """The guide counts only sun photons that reached a visible point: in the caustic scene after the
    floor's direct light (the camera pass's), those are the photons through the 40 mm lens, a small
    patch of the emission disc — and the guided estimate agrees with the unguided one."""
    ls = C.caustic_cases(resolution=32)["direct"]
    w = World(oracle, ls.use())
    st = w.sppm(ls.settings(seed=2), 64)
    g = st["guide"]
    assert g.sum() > 0
    assert (g > 0).mean() < 0.02  # a lens of radius 20 mm in a disc that covers the 300 mm floor
    assert World.sppm_resolve(st)[..., 1].mean() > 0


def test_sppm_iterations_continue(oracle):
    """Two calls of n iterations are one call of 2n: the state carries everything."""
    ls = C.caustic_cases(resolution=16)["direct"]  # sun photons: the guide's counts are state too
    for w, st in ((World(oracle, S.white_furnace_scene(0.5, size=8)), S.settings(seed=4, integrator=SPPM)),
                  (World(oracle, ls.use()), ls.settings(seed=4))):
        a = w.sppm(st, 6)
        b = w.sppm(st, 3)
        b = w.sppm(st, 3, state=b)
        assert np.array_equal(a["pix"], b["pix"]) and np.array_equal(a["direct"], b["direct"])
        assert np.array_equal(a["guide"], b["guide"])
