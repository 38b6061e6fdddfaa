"""The batched lens analysis on the GPU against the fp64 oracle.

* per ray: the same Fibonacci-lattice rays traced through the same lenses by the float32 kernel
  (roots polished in float64) and the fp64 oracle agree in acceptance, point (1e-8 m), direction
  (1e-6) and weight (1e-6), except rays within float rounding of an aperture edge;
* per lens: the spot (arrived rays, best focus, RMS radius, centroid) from the kernel's summed
  moments agrees with the oracle's over every ray of the lens at every wavelength;
* the irradiance grid holds the transmitted flux that lands on it, and is identical run to run and
  under every launch shape (integer bins);
* the reference's doublet focuses all rays and its paraxial numbers equal the oracle's.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

from owe_hip import lens as L
from pyoracle import LensOracle

pytestmark = pytest.mark.gpu

LENSES = [L.DOUBLET, L.SINGLET] + L.seeded_achromats(6, seed=11, rays_max=200_000)


@pytest.fixture(scope="module")
def batch(cuda):
    b = L.LensBatch(LENSES)
    b.run()
    yield b
    b.close()


@pytest.mark.parametrize("li", range(len(LENSES)))
def test_rays_match_oracle_ray_by_ray(batch, oracle, li):
    lo = LensOracle(oracle, LENSES)
    n = min(LENSES[li].rays, 8192)
    idx = np.arange(n, dtype=np.uint64)
    for w in (0, 4, 6):
        g = batch.trace(li, w, idx)
        c = lo.trace(li, L.FRAUNHOFER7[w], idx)
        okg, okc = g[:, 6] > 0, c[:, 6] > 0
        both = okg & okc
        flips = int((okg != okc).sum())
        dp = np.abs(g[both, :3] - c[both, :3]).max(initial=0)
        dd = np.abs(g[both, 3:6] - c[both, 3:6]).max(initial=0)
        dT = np.abs(g[both, 6] - c[both, 6]).max(initial=0)
        print(f"lens {li} λ{w}: {both.sum()} both arrive, {flips} flips, |Δp| {dp:.2e} |Δd| {dd:.2e} |ΔT| {dT:.2e}",
              flush=True)
        assert flips <= max(2, n // 2000)
        assert dp < 1e-8 and dd < 1e-6 and dT < 1e-6


@pytest.mark.parametrize("li", range(len(LENSES)))
def test_spot_matches_oracle(batch, oracle, li):
    per, mom, _ = batch.results()
    lo = LensOracle(oracle, LENSES)
    ms = np.array([lo.moments(li, lam)[0] for lam in L.FRAUNHOFER7])
    _, px = lo.moments(li, L.LAMBDA_D)
    lv = px["last_vertex_z"]
    s = L.spot(ms.sum(axis=0), lv + px["bfl"], lv)
    r = per[li]
    print(f"lens {li}: arrived {r['arrived']} / {s['arrived']}, rms {r['rms_best']:.6e} / {s['rms_best']:.6e}, "
          f"focus {r['best_focus_z']:.9f} / {s['best_focus_z']:.9f}", flush=True)
    assert r["efl"] == pytest.approx(px["efl"], rel=1e-12)
    assert r["bfl"] == pytest.approx(px["bfl"], rel=1e-12)
    assert r["entrance_pupil_radius"] == pytest.approx(px["entrance_pupil_radius"], rel=1e-12)
    assert abs(r["arrived"] - s["arrived"]) <= max(2, s["arrived"] // 2000)
    assert r["rms_best"] == pytest.approx(s["rms_best"], rel=2e-3)
    assert r["best_focus_z"] == pytest.approx(s["best_focus_z"], abs=2e-7)
    assert r["centroid_y"] == pytest.approx(s["centroid_y"], abs=1e-8)
    assert np.abs(mom[li, :, 0] - ms[:, 0]).max() <= max(2, s["arrived"] // 2000)


def test_ring_pupil_matches_oracle(cuda, oracle):
    # spotAt's hexapolar rings: six (the reference's analyzeLens call) and 300 (270,901 rays).
    lenses = [L.SINGLET.with_rings(6), L.DOUBLET.with_rings(6), LENSES[2].with_rings(300)]
    b = L.LensBatch(lenses)
    b.run()
    per, mom, _ = b.results()
    lo = LensOracle(oracle, lenses)
    for li, p in enumerate(lenses):
        idx = np.arange(min(p.rays, 20000), dtype=np.uint64)
        g = b.trace(li, 4, idx)
        c = lo.trace(li, L.FRAUNHOFER7[4], idx)
        okg, okc = g[:, 6] > 0, c[:, 6] > 0
        both = okg & okc
        flips = int((okg != okc).sum())
        dp = np.abs(g[both, :3] - c[both, :3]).max(initial=0)
        ms = np.array([lo.moments(li, lam)[0] for lam in L.FRAUNHOFER7])
        print(f"rings lens {li}: {p.rays} rays, {both.sum()} both arrive, {flips} flips, |Δp| {dp:.2e}, "
              f"arrived {per[li]['arrived']} / {int(ms[:, 0].sum())}", flush=True)
        assert flips <= 2 and dp < 1e-8
        assert abs(per[li]["arrived"] - ms[:, 0].sum()) <= max(2, ms[:, 0].sum() // 2000)
    b.close()


def test_doublet_focuses_every_ray_on_axis(cuda):
    # On axis every ray of the reference's achromat arrives (tests/test_instruments.cpp:333).
    b = L.LensBatch([L.DOUBLET], L.c_settings(field_deg=0.0))
    b.run()
    r = b.results()[0][0]
    b.close()
    assert r["arrived"] == r["launched"]
    assert 0.9 < r["transmitted"] < 1.0


def test_irradiance_holds_the_flux_that_lands_and_is_reproducible(cuda):
    half = 3e-3  # wide enough for the seven wavelengths' defocus
    b = L.LensBatch(LENSES[:3], L.c_settings(half_size=half))
    b.run()
    per, mom, irr = b.results(irradiance=True)
    b.set_shape(blocks_per_cu=1, rays_per_item=64)
    b.run()
    per2, mom2, irr2 = b.results(irradiance=True)
    assert np.array_equal(irr, irr2)
    assert [r["arrived"] for r in per] == [r["arrived"] for r in per2]
    for li, p in enumerate(LENSES[:3]):
        pix = 2 * half / irr.shape[1]
        flux = irr[li].sum() * pix * pix
        beam = math.pi * (0.999 * per[li]["entrance_pupil_radius"]) ** 2 * math.cos(math.radians(0.5))
        print(f"lens {li}: flux on the grid {flux:.6e}, transmitted {per[li]['transmitted'] * beam:.6e}", flush=True)
        # All the transmitted flux lands on the grid, up to the fixed-point rounding of each ray.
        assert flux == pytest.approx(per[li]["transmitted"] * beam, rel=1e-4)
    b.close()


def test_autotune_keeps_the_results(cuda):
    b = L.LensBatch(LENSES[:4])
    b.run()
    per, _, irr = b.results(irradiance=True)
    ms = b.autotune()
    b.run()
    per2, _, irr2 = b.results(irradiance=True)
    print(f"autotune: {ms:.3f} ms", flush=True)
    assert ms > 0
    assert np.array_equal(irr, irr2)
    for a, c in zip(per, per2):
        assert a["arrived"] == c["arrived"]
        assert c["rms_best"] == pytest.approx(a["rms_best"], rel=1e-9)
    b.close()
