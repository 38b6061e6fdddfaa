"""The oracle's camera-path transport against the reference's closed forms, and the comparison logic.

Closed forms from the reference's own tests: white furnace 1/(1 − ρ) (tests/test_transport.cpp:83-114),
a lossless dielectric sphere invisible in a uniform sky (116-142), pinned here on the oracle; the
two stream layouts (one stream per pixel — the reference's — and one per sample — the hip backend's)
estimate the same image (block-z consistency, src/owe/backends/compare.cpp:64-119).
"""
from __future__ import annotations

import math

import numpy as np
import pytest

from owe_hip import compare
from owe_hip import scene as S
from pyoracle import World


@pytest.mark.parametrize("rho", [0.0, 0.5, 0.8])
def test_white_furnace(oracle, rho):
    w = World(oracle, S.white_furnace_scene(rho))
    img, st = w.render(S.settings(seed=1), 256, layout=0)
    assert img[..., 1].mean() == pytest.approx(1 / (1 - rho), abs=0.01 / (1 - rho))
    assert st[2] == 0 and st[3] == 0  # no inconsistencies, no leaks


def test_dielectric_sphere_is_invisible(oracle):
    w = World(oracle, S.dielectric_sphere_scene())
    img, st = w.render(S.settings(seed=1), 256, layout=0)
    assert img[..., 1].mean() == pytest.approx(1.0, abs=0.004)
    assert img[6:18, 6:18, 1].mean() == pytest.approx(1.0, abs=0.01)
    assert st[2] == 0 and st[3] == 0


def test_prism_scene_light_table_and_layout_consistency(oracle):
    w = World(oracle, S.dispersion_prism_scene(32, 32))
    pdf, b = w.light_table()
    assert len(pdf) == 1 and pdf[0] == 1.0  # the sun alone
    st = S.settings(max_depth=16, rr_depth=4, spectral=S.SPECTRAL_UNIFORM1)

    def run(layout):
        def f(seed):
            st.seed = seed
            return w.render(st, 32, layout=layout)[0]
        return f

    rep = compare.compare_renderers(run(0), run(1), runs=4, B=4)
    assert rep.consistent(), rep.text()


def test_uniform_wavelength_estimator_of_a_white_screen(oracle):
    """A uniform sky of radiance 1 seen directly: L = 1 at every wavelength. One wavelength per path
    uniform on [380, 780] nm estimates Y = ∫_{380}^{780} ȳ / ∫_{360}^{830} ȳ (the CMF tails outside
    the range are dropped); the reference's four hero wavelengths estimate Y = 1."""
    sc = S.Scene()
    air = sc.add_medium(S.Medium.constant(1.0))
    sc.add_region(air)
    sc.set_uniform_sky(S.Spectrum.constant(1.0))
    sc.set_observer(8, 8, position=(0, 0, 0), look_at=(0, 1, 0))
    w = World(oracle, sc)
    img, _ = w.render(S.settings(spectral=S.SPECTRAL_UNIFORM1), 256, layout=1)
    lam = np.linspace(380, 780, 40001)
    cie_y = np.array([oracle.cie_xyz(l)[1] for l in lam[::40]])
    expect = np.trapezoid(cie_y, lam[::40]) / oracle.cie_y_integral()
    assert img[..., 1].mean() == pytest.approx(expect, rel=0.01)
    img4, _ = w.render(S.settings(spectral=S.SPECTRAL_HERO4), 64, layout=1)
    assert img4[..., 1].mean() == pytest.approx(1.0, rel=0.01)


def test_compare_detects_a_bias_and_accepts_noise():
    rng = np.random.default_rng(1)
    base = rng.uniform(0.5, 1.5, size=(32, 32, 3))
    noisy = lambda s: base + np.random.default_rng(s).normal(0, 0.05, base.shape)  # noqa: E731
    rep = compare.compare_renderers(noisy, noisy, runs=6, B=4)
    assert rep.consistent(), rep.text()
    biased = lambda s: noisy(s) * 1.03  # noqa: E731
    rep = compare.compare_renderers(noisy, biased, runs=6, B=4)
    assert not rep.consistent()


def test_blackbody_scale_matches_reference():
    import json

    from conftest import GOLDEN

    g = json.loads((GOLDEN / "formulas.json").read_text())
    for T, y, lam, v in g["blackbody"]:
        s = S.blackbody_scale(T, y)
        assert s * S.planck(lam, T) == pytest.approx(v, rel=1e-12)


def test_prism_mesh_matches_reference_builder():
    pos, tri = S.prism_mesh(math.radians(60), 0.040, 0.020)
    assert tri.shape == (8, 3)
    c = pos.mean(axis=0)
    for t in tri:
        n = np.cross(pos[t[1]] - pos[t[0]], pos[t[2]] - pos[t[0]])
        assert np.dot(n, pos[t].mean(axis=0) - c) > 0
    # Equilateral cross-section with 40 mm sides, centroid at the origin, 20 mm long.
    xy = np.unique(np.round(pos[:, :2], 12), axis=0)
    assert len(xy) == 3
    d = [np.linalg.norm(xy[i] - xy[j]) for i, j in ((0, 1), (1, 2), (0, 2))]
    assert np.allclose(d, 0.040)
    assert np.allclose(xy.mean(axis=0), 0, atol=1e-15)
    assert np.ptp(pos[:, 2]) == pytest.approx(0.020)
