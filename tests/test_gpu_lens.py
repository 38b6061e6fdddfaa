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
        pix = 2 * half / irr.shape[-1]
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


def _oracle_grid(lo, li, lens, per, settings, planes, step, half, px):
    """The through-focus grids binned in double from the oracle's rays: every arrived ray of every
    wavelength carries T·π(0.999 r_EP)² cos θ / (rays · nλ) of flux into the bin its line crosses on
    each plane (plane k at the paraxial image + (k − (P − 1)/2)·step)."""
    r = per[li]
    idx = np.arange(lens.rays, dtype=np.uint64)
    theta = math.radians(settings.field_deg)
    nw = settings.n_wavelengths
    grid = np.zeros((planes, px, px))
    z0 = r["last_vertex_z"] + r["bfl"]
    cy = r["efl"] * math.tan(theta)
    flux = math.pi * (0.999 * r["entrance_pupil_radius"]) ** 2 * math.cos(theta) / (lens.rays * nw)
    pix = 2 * half / px
    for w in range(nw):
        t = lo.trace(li, settings.wavelengths[w], idx)
        t = t[t[:, 6] > 0]
        for k in range(planes):
            z = z0 + (k - 0.5 * (planes - 1)) * step
            s = (z - t[:, 2]) / t[:, 5]
            gx = (t[:, 0] + t[:, 3] * s) / half
            gy = (t[:, 1] + t[:, 4] * s - cy) / half
            fx, fy = (gx + 1) * 0.5 * px, (1 - gy) * 0.5 * px
            ok = (fx >= 0) & (fy >= 0) & (fx < px) & (fy < px)
            np.add.at(grid[k], (fy[ok].astype(int), fx[ok].astype(int)), t[ok, 6] * flux / (pix * pix))
    return grid


def test_through_focus_stack_matches_the_oracle_bin_by_bin(cuda, oracle):
    """Five planes 40 µm apart about the paraxial image of two lenses: every bin of every plane
    against the oracle's rays binned in double (a ray within float rounding of a bin edge may land in
    its neighbour); the middle plane equals a one-plane run bit for bit; the sharpest plane lies at, or next to, the plane nearest the best focus."""
    lenses = [LENSES[1], LENSES[2]]
    half, planes, step = 0.25e-3, 5, 40e-6
    st = L.c_settings(half_size=half, planes=planes, focus_step=step, pixels=64)
    b = L.LensBatch(lenses, st)
    b.run()
    per, _, irr = b.results(irradiance=True)
    one = L.LensBatch(lenses, L.c_settings(half_size=half, pixels=64))
    one.run()
    _, _, irr1 = one.results(irradiance=True)
    one.close()
    b.close()
    assert irr.shape == (2, planes, 64, 64)
    assert np.array_equal(irr[:, planes // 2], irr1[:, 0])
    lo = LensOracle(oracle, lenses)
    for li, p in enumerate(lenses):
        want = _oracle_grid(lo, li, p, per, st, planes, step, half, 64)
        got = irr[li].astype(np.float64)
        close = np.abs(got - want) <= 1e-5 * np.maximum(np.abs(want), np.abs(got)).max()
        flux = got.sum(axis=(1, 2))
        print(f"lens {li}: {close.mean():.5f} of bins agree; plane flux {flux / flux.max()}; peak per plane "
              f"{got.max(axis=(1, 2))}", flush=True)
        assert close.mean() > 0.995
        assert np.abs(got.sum() - want.sum()) <= 1e-4 * want.sum()
