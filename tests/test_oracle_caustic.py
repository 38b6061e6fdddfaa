"""The oracle's caustic slab (fp64) against closed forms.

* a flat slab transmits (1 − R)² of the luminous flux at normal incidence (Fresnel, optics.hpp:42-54,
  luminous-weighted over the 5778 K sun), all of it landing inside the slab's shadow square;
* a uniformly tilted exit face deviates the beam by θt − θi with n₁ sin θi = n₂ sin θt (Snell), so
  the image's centroid moves by the wall distance × tan(θt − θi), luminous-weighted over λ;
* the loss is continuously differentiable (quadratic B-spline splat): central differences converge
  as the step shrinks, which is what the GPU adjoint is checked against;
* the target carries the flux it is scaled to.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

from owe_hip import caustic as C
from pyoracle import CausticOracle

SMALL = dict(particles=1 << 17)


def pmma(lam):
    l2 = (lam * 1e-3) ** 2
    return math.sqrt(1 + C.PMMA_B[0] * l2 / (l2 - C.PMMA_C[0]))


def air(lam):
    s2 = (1e3 / lam) ** 2
    return 1 + 0.05792105 / (238.0185 - s2) + 0.00167917 / (57.362 - s2)


def luminous_average(f):
    lam = np.linspace(360, 830, 4701)

    def planck(l):
        c1, c2, x = 1.191042972e8, 1.438776877e4, l * 1e-3
        return 1e6 * c1 / (x ** 5 * (np.exp(c2 / (x * 5778.0)) - 1))

    def cie_y(l):
        def lobe(x, mu, s1, s2):
            t = (x - mu) / np.where(x < mu, s1, s2)
            return np.exp(-0.5 * t * t)

        return 0.821 * lobe(l, 568.8, 46.9, 40.5) + 0.286 * lobe(l, 530.9, 16.3, 31.1)

    w = planck(lam) * cie_y(lam)
    vals = np.array([f(l) for l in lam])
    return float((w * vals).sum() / w.sum())


def fresnel_normal(n1, n2):
    return ((n1 - n2) / (n1 + n2)) ** 2


def test_flat_slab_transmits_two_fresnel_factors(oracle):
    s = C.caustic_settings(**SMALL)
    co = CausticOracle(oracle, s)
    E = co.forward(np.zeros((s.nodes, s.nodes)))
    pix = 2 * s.wall_half / s.pixels
    flux = E.sum() * pix * pix
    expect = luminous_average(lambda l: (1 - fresnel_normal(pmma(l), air(l))) ** 2) * (2 * s.slab_half) ** 2
    assert flux == pytest.approx(expect, rel=5e-3)
    c = (np.arange(s.pixels) + 0.5) * pix - s.wall_half
    outside = (np.abs(c)[None, :] > s.slab_half + 2 * pix) | (np.abs(c)[:, None] > s.slab_half + 2 * pix)
    assert E[outside].sum() == 0


def test_tilted_exit_face_deviates_the_beam_by_snell(oracle):
    s = C.caustic_settings(**SMALL)
    co = CausticOracle(oracle, s)
    slope = 0.02
    x = np.linspace(-s.slab_half, s.slab_half, s.nodes)
    pix = 2 * s.wall_half / s.pixels
    c = (np.arange(s.pixels) + 0.5) * pix - s.wall_half

    def centroid(H):
        E = co.forward(H)
        return (E.sum(axis=0) * c).sum() / E.sum()

    # The same particles through a flat and a tilted face (h = slope · x, row = y): the shift of
    # the centroid is the deviation alone, free of the particles' own sampling noise.
    cx = centroid(np.tile(slope * x, (s.nodes, 1))) - centroid(np.zeros((s.nodes, s.nodes)))
    ti = math.atan(slope)

    def shift(l):
        tt = math.asin(pmma(l) / air(l) * math.sin(ti))
        return math.tan(tt - ti)

    # Leaving the denser medium the ray bends away from the normal, which leans toward −x: the beam
    # moves toward +x; the wall is (wall_z − thickness) beyond the face.
    expect = luminous_average(shift) * (s.wall_z - s.thickness)
    assert cx == pytest.approx(expect, rel=2e-3)


def test_loss_is_smooth_and_its_finite_differences_converge(oracle):
    s = C.caustic_settings(**SMALL)
    co = CausticOracle(oracle, s)
    H = C.initial_heights().astype(np.float64)
    T = C.target(s)
    nodes = np.random.default_rng(3).integers(0, s.nodes * s.nodes, 16)
    g1 = co.fd(H, T, nodes, 1e-9)
    g2 = co.fd(H, T, nodes, 3e-10)
    assert np.abs(g1 - g2).max() <= 1e-3 * np.abs(g2).max()
    assert np.abs(g2).max() > 0


def test_target_carries_its_flux():
    s = C.caustic_settings()
    T = C.target(s)
    pix = 2 * s.wall_half / s.pixels
    assert T.sum() * pix * pix == pytest.approx(0.92 * (2 * s.slab_half) ** 2, rel=1e-5)
    H = C.initial_heights()
    assert np.abs(H).max() == pytest.approx(1e-4, rel=1e-6)


@pytest.mark.parametrize("particles", [1 << 17, 100_003])
def test_stratified_particles_cover_the_entry_face_evenly(oracle, particles):
    """Particles are stratified over tiles of the entry face (tile t holds ⌊t·N/T⌋ … ⌊(t+1)·N/T⌋ − 1,
    uniform in the tile, weighted by its area over its own count). Through a flat slab each tile
    lights exactly its own square of the wall, whether or not T divides N: the flux is the flat slab's,
    and no tile's square of pixels inside the shadow is doubled or empty."""
    s = C.caustic_settings(particles=particles)
    E = CausticOracle(oracle, s).forward(np.zeros((s.nodes, s.nodes)))
    pix = 2 * s.wall_half / s.pixels
    expect = luminous_average(lambda l: (1 - fresnel_normal(pmma(l), air(l))) ** 2) * (2 * s.slab_half) ** 2
    assert E.sum() * pix * pix == pytest.approx(expect, rel=5e-3)
    tiles = 1
    while tiles < 64 and 4 * tiles * tiles * 64 <= particles:
        tiles *= 2
    edge = int(round(s.slab_half / pix))        # pixels from the centre to the shadow's edge
    step = int(round(2 * s.slab_half / tiles / pix))  # a tile's side in pixels
    assert step * tiles == 2 * edge
    c0 = s.pixels // 2
    shadow = E[c0 - edge:c0 + edge, c0 - edge:c0 + edge]
    blocks = shadow.reshape(tiles, step, tiles, step).sum(axis=(1, 3))[1:-1, 1:-1]  # away from the rim
    # The spectral weights B ȳ / pdf spread widely (a few particles carry much), so a tile's flux varies
    # by several per cent; a doubled or a missing tile is off by 100 %.
    med = np.median(blocks)
    assert blocks.min() > 0.5 * med and blocks.max() < 1.5 * med
