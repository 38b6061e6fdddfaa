"""The oracle's participating-medium transport (σs with a Henyey–Greenstein phase, σa, hero
wavelengths reweighted by their own extinction) against closed forms and an independent walk.

The reference has no test that sets a scattering coefficient (its scenes and tests leave σs = 0);
what it computes is src/owe/transport/transport.cpp:95-113 (sampleMedium: distance sampling by
the hero wavelength's extinction, secondaries reweighted by σs(λ)e^(−σt(λ)t) / σt(λ₀)e^(−σt(λ₀)t)),
transport.cpp:379-397 (scatterInMedium: next-event estimation from the medium point, then an HG
direction), core/shared/warps.hpp:145-160 (hgPhase, sampleHg), and the particle walk's medium
branch (shaders/particle.slang:356-427). Pins:

* the scattering furnace: a sphere of conservative medium (σa = 0) inside a wall that emits 1 and
  reflects ρ keeps the radiance uniform at 1 / (1 − ρ) for every σs and g, for camera paths, light
  tracing and SPPM (tests/test_backends.cpp:540-574 states the furnace for every integrator);
* with absorption there is no closed form: tests/media_walk.py walks the radiative transfer
  equation by brute force (no next-event estimation, no MIS, no wavelength sharing), and the
  oracle's radiance at the sphere's centre must match it;
* wavelength sharing: σs and σa step at 580 nm, so hero-4 secondaries on the other side of the step
  carry weights of another medium; the ȳ-weighted mixture of the two walks' radiances is what hero-4
  and uniform-1 sampling must both estimate.
"""
from __future__ import annotations

import numpy as np
import pytest

import media_walk
from owe_hip import scene as S
from pyoracle import World

DEEP = dict(max_depth=1000, rr_depth=6)


@pytest.mark.parametrize("g", [-0.5, 0.0, 0.8])
@pytest.mark.parametrize("rho", [0.0, 0.5])
def test_scattering_furnace_camera_paths(oracle, g, rho):
    w = World(oracle, S.scattering_furnace_scene(2.0, 0.0, g, rho))
    img, st = w.render(S.settings(seed=1, **DEEP), 256, layout=0)
    assert img[..., 1].mean() == pytest.approx(1 / (1 - rho), rel=0.01)
    assert st[1] > 2 * st[0]  # the paths scatter in the medium
    assert st[2] == 0 and st[3] == 0


@pytest.mark.parametrize("g", [-0.5, 0.0, 0.8])
def test_scattering_furnace_light_tracing(oracle, g):
    rho = 0.5
    w = World(oracle, S.scattering_furnace_scene(3.0, 0.0, g, rho, size=8, bubble=0.3))
    ys = [w.render_particles(S.settings(seed=s, integrator=S.INTEGRATOR_LIGHT, **DEEP), 4096)[0][..., 1].mean()
          for s in (1, 2, 3, 4)]
    assert np.mean(ys) == pytest.approx(1 / (1 - rho), rel=0.03)


@pytest.mark.parametrize("rho", [0.0, 0.5])
def test_scattering_furnace_sppm(oracle, rho):
    w = World(oracle, S.scattering_furnace_scene(3.0, 0.0, 0.8, rho, size=8, bubble=0.3))
    st = w.sppm(S.settings(seed=1, integrator=S.INTEGRATOR_SPPM, **DEEP), 512)
    assert World.sppm_resolve(st)[..., 1].mean() == pytest.approx(1 / (1 - rho), rel=0.02)
    assert st["stats"][2] == 0 and st["stats"][3] == 0


def test_brute_force_walk_closed_forms():
    """The walk itself: a conservative medium with black walls gives exactly 1, and with walls of
    albedo ρ the furnace value 1 / (1 − ρ)."""
    m, e = media_walk.sphere_radiance(3.0, 0.0, 0.5, 0.0, walks=20000)
    assert m == 1.0 and e == 0.0
    m, e = media_walk.sphere_radiance(3.0, 0.0, -0.3, 0.5, walks=100000)
    assert m == pytest.approx(2.0, abs=4 * e + 1e-3)


@pytest.mark.parametrize("sigma_s,sigma_a,g,rho", [(3.0, 1.0, 0.0, 0.0), (3.0, 1.0, 0.6, 0.0), (1.0, 2.0, 0.8, 0.3),
                                                   (3.0, 1.0, -0.4, 0.5)])
def test_absorbing_fog_matches_the_brute_force_walk(oracle, sigma_s, sigma_a, g, rho):
    want, err = media_walk.sphere_radiance(sigma_s, sigma_a, g, rho, walks=400000, seed=7)
    w = World(oracle, S.scattering_furnace_scene(sigma_s, sigma_a, g, rho, at_centre=True))
    img, st = w.render(S.settings(seed=2, **DEEP), 1024, layout=0)
    assert img[..., 1].mean() == pytest.approx(want, rel=0.015)
    assert st[2] == 0 and st[3] == 0


STEP_NM = 580.0
FOG_A = (3.0, 0.5)  # (σs, σa) per metre below the step
FOG_B = (2.0, 1.0)  # and above it


def step_spectrum(a: float, b: float) -> S.Spectrum:
    return S.Spectrum.tabulated([380.0, STEP_NM - 0.01, STEP_NM + 0.01, 780.0], [a, a, b, b])


def stepped_fog_scene(**kw):
    return S.scattering_furnace_scene(step_spectrum(FOG_A[0], FOG_B[0]), step_spectrum(FOG_A[1], FOG_B[1]),
                                      g=0.5, rho=0.3, at_centre=True, **kw)


def stepped_fog_expectation(lo: float, hi: float, norm_lo: float = 360.0, norm_hi: float = 830.0,
                            walks: int = 400000) -> float:
    """Y of the stepped fog seen from the centre for wavelengths in [lo, hi], normalised by ∫ ȳ
    over [norm_lo, norm_hi]: the two media's walk radiances weighted by the ȳ mass on each side."""
    la, _ = media_walk.sphere_radiance(*FOG_A, 0.5, 0.3, walks=walks, seed=11)
    lb, _ = media_walk.sphere_radiance(*FOG_B, 0.5, 0.3, walks=walks, seed=12)
    lam = np.linspace(norm_lo, norm_hi, 47001)
    y = S._cie_y(lam)
    below = (lam >= lo) & (lam < STEP_NM)
    above = (lam >= STEP_NM) & (lam <= hi)
    return float((la * y[below].sum() + lb * y[above].sum()) / y.sum())


def test_hero_wavelengths_share_paths_through_a_stepped_fog(oracle):
    want = stepped_fog_expectation(360.0, 830.0)
    w = World(oracle, stepped_fog_scene())
    img, st = w.render(S.settings(seed=3, spectral=S.SPECTRAL_HERO4, **DEEP), 1024, layout=0)
    assert img[..., 1].mean() == pytest.approx(want, rel=0.015)
    img1, _ = w.render(S.settings(seed=4, spectral=S.SPECTRAL_UNIFORM1, **DEEP), 1024, layout=0)
    assert img1[..., 1].mean() == pytest.approx(stepped_fog_expectation(380.0, 780.0), rel=0.015)
