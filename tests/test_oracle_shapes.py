"""The oracle's cylinder, sag and wave shapes in the world walk, against closed forms, brute force
and the reference's own forward-light tests.

* a cylinder's hits: t = |o_xy| − R for radial rays from outside, the chord 2R cos α for a ray
  leaving the wall (the other root), nothing beyond its z trim;
* a sag with a central hole (an annular mirror): hits on z = sag(r) to 1e-12 with the sag's normal,
  none inside the hole; a ray leaving the concave face finds its other root;
* a wave surface: hits on z = h(x, y) with normal ∝ (−∇h, 1); the first crossing equals a brute-force
  scan of f(t) = z(t) − h along the ray; a departing ray does not find its own point; the sampled
  bicubic Hermite representation follows the analytic sum to 1e-3 of its amplitude (WaveSurface's
  stated accuracy, geometry.hpp:113-115);
* light tracing (particles) reproduces the power budgets of tests/test_references.cpp: a lens
  focusing sunlight (77-120), the Fresnel share through water (122-147), a concave mirror (149-175)
  within 1 %, and the ripple caustic's ray-map Jacobian (322-403): total within 1 %, L1 < 6 %, the
  fold in the same bin (or its twin a wave away), peak above 3 E₀.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

from owe_hip import bodies as B
from owe_hip import scene as S
from pyoracle import World


def one_shape_world(oracle, add):
    sc = S.Scene()
    air = sc.add_region(sc.add_medium(S.Medium.air()))
    glass = sc.add_region(sc.add_medium(S.Medium.constant(1.5)))
    diel = sc.add_optics(S.Optics(S.SURF_DIELECTRIC))
    add(sc, air, glass, diel)
    sc.set_observer(2, 2, (0, 0, 5), (0, 0, 0))
    return World(oracle, sc)


def test_cylinder_closed_forms(oracle):
    R, z0, z1 = 0.1, -0.05, 0.05
    w = one_shape_world(oracle, lambda sc, a, g, d: sc.add_cylinder(R, z0, z1, S.transform(), a, g, d))
    rng = np.random.default_rng(1)
    phi = rng.uniform(0, 2 * math.pi, 200)
    dist = rng.uniform(0.2, 3.0, 200)
    zs = rng.uniform(z0 + 1e-3, z1 - 1e-3, 200)
    o = np.stack([dist * np.cos(phi), dist * np.sin(phi), zs], axis=1)
    d = -np.stack([np.cos(phi), np.sin(phi), np.zeros(200)], axis=1)
    h = w.intersect(o, d)
    assert np.abs(h[:, 0] - (dist - R)).max() < 1e-12
    assert np.abs(np.hypot(h[:, 1], h[:, 2]) - R).max() < 1e-12
    assert np.abs(h[:, 4:6] - np.stack([np.cos(phi), np.sin(phi)], axis=1)).max() < 1e-12
    # Leaving the wall inward at angle α to the inward normal: the chord 2R cos α.
    alpha = rng.uniform(-1.2, 1.2, 200)
    p = np.stack([R * np.cos(phi), R * np.sin(phi), zs], axis=1)
    inward = -np.stack([np.cos(phi), np.sin(phi), np.zeros(200)], axis=1)
    side = np.stack([-np.sin(phi), np.cos(phi), np.zeros(200)], axis=1)
    dd = inward * np.cos(alpha)[:, None] + side * np.sin(alpha)[:, None]
    h = w.intersect(p, dd, np.zeros(200, np.uint32))
    assert np.abs(h[:, 0] - 2 * R * np.cos(alpha)).max() < 1e-12
    # Above the trim: no hit.
    h = w.intersect(o + np.array([0, 0, 0.2]), d)
    assert np.isinf(h[:, 0]).all()


def test_annular_sag_hits_and_other_root(oracle):
    Rm, rmax, rmin = -0.4, 0.05, 0.01   # concave toward +z (a mirror facing up), central hole
    c = 1 / Rm

    def add(sc, a, g, d):
        sc.add_sag(-c, 0.0, (), rmax, rmin, S.transform(), a, a, d)

    w = one_shape_world(oracle, add)
    rng = np.random.default_rng(2)
    n = 400
    x, y = rng.uniform(-0.06, 0.06, n), rng.uniform(-0.06, 0.06, n)
    o = np.stack([x, y, np.full(n, 0.3)], axis=1)
    d = np.tile([0.0, 0.0, -1.0], (n, 1))
    h = w.intersect(o, d)
    r = np.hypot(x, y)
    inside = (r <= rmax) & (r >= rmin)
    assert (np.isfinite(h[:, 0]) == inside).all()
    hit = h[inside]
    rr = np.hypot(hit[:, 1], hit[:, 2])
    sag = np.array([B.sag_of(-c, 0.0, v) for v in rr])
    assert np.abs(hit[:, 3] - sag).max() < 1e-12
    # The normal of z = sag(r): (−s' x/r, −s' y/r, 1) normalised, s' by central differences.
    e = 1e-7
    slope = np.array([(B.sag_of(-c, 0, v + e) - B.sag_of(-c, 0, v - e)) / (2 * e) for v in rr])
    nz = 1 / np.sqrt(1 + slope ** 2)
    assert np.abs(hit[:, 6] - nz).max() < 1e-6
    # A ray leaving the concave face toward the far side finds the face again there.
    p0 = np.array([[0.03, 0.0, B.sag_of(-c, 0, 0.03)]])
    dl = np.array([[-1.0, 0.0, -0.02]])
    dl /= np.linalg.norm(dl)
    h = w.intersect(p0, dl, np.zeros(1, np.uint32))
    assert np.isfinite(h[0, 0]) and h[0, 0] > 0.03
    assert abs(h[0, 3] - B.sag_of(-c, 0, math.hypot(h[0, 1], h[0, 2]))) < 1e-12


def wave_height(waves, hx, hy, margin, x, y):
    """The tapered plane-wave sum (WaveSurface::analytic, geometry.cpp:226-242), vectorised."""
    x, y = np.asarray(x, float), np.asarray(y, float)

    def taper(u, half):
        e = half - np.abs(u)
        t = np.clip(e / margin, 0.0, 1.0)
        return np.where(e >= margin, 1.0, np.where(e <= 0, 0.0, t * t * (3 - 2 * t)))

    g = sum(a * np.sin(kx * x + ky * y + ph) for a, kx, ky, ph in waves)
    inside = (np.abs(x) <= hx) & (np.abs(y) <= hy)
    return np.where(inside, taper(x, hx) * taper(y, hy) * g, 0.0)


WAVES = [(0.0005, 2 * math.pi / 0.02, 0.0, 0.0), (0.0002, 60.0, 240.0, 1.0)]


@pytest.mark.parametrize("spacing", [0.0, 0.0008])
def test_wave_hits_lie_on_the_surface_and_are_the_first_crossing(oracle, spacing):
    hx = hy = 0.06
    margin = 0.02

    def add(sc, a, g, d):
        sc.add_wave(WAVES, hx, hy, margin, S.transform(), a, g, d, spacing)

    w = one_shape_world(oracle, add)
    rng = np.random.default_rng(3)
    n = 300
    o = np.stack([rng.uniform(-0.05, 0.05, n), rng.uniform(-0.05, 0.05, n), np.full(n, 0.02)], axis=1)
    d = np.stack([rng.uniform(-0.9, 0.9, n), rng.uniform(-0.9, 0.9, n), -np.ones(n)], axis=1)
    d /= np.linalg.norm(d, axis=1)[:, None]
    h = w.intersect(o, d)
    amp = sum(abs(a) for a, *_ in WAVES)
    tol = 1e-12 if spacing == 0 else 1e-3 * amp
    hit = np.isfinite(h[:, 0])
    assert hit.sum() > n // 2
    assert np.abs(h[hit, 3] - wave_height(WAVES, hx, hy, margin, h[hit, 1], h[hit, 2])).max() < tol
    if spacing == 0:
        # Brute force: the first sign change of z(t) − h(x(t), y(t)) on a scan of step 1e-6.
        ts = np.linspace(0, 0.1, 100001)
        for i in range(0, n, 10):
            P = o[i] + d[i] * ts[:, None]
            f = P[:, 2] - wave_height(WAVES, hx, hy, margin, P[:, 0], P[:, 1])
            inside = (np.abs(P[:, 0]) <= hx) & (np.abs(P[:, 1]) <= hy)
            ch = np.flatnonzero((np.sign(f[1:]) != np.sign(f[:-1])) & inside[1:])
            if len(ch):
                assert h[i, 0] == pytest.approx(ts[ch[0]], abs=2e-6)
            else:
                assert not np.isfinite(h[i, 0])
    # Leaving a hit point upward (reflected): not the same point again.
    ok = np.flatnonzero(np.isfinite(h[:, 0]))[:50]
    up = d[ok] * np.array([1, 1, -1])
    again = w.intersect(h[ok, 1:4], up, np.zeros(len(ok), np.uint32))
    assert (~np.isfinite(again[:, 0]) | (again[:, 0] > 1e-6)).all()


# ---------------------------------------------------------------- the reference's power budgets

LIGHT = S.settings(seed=1, max_depth=32, rr_depth=6, integrator=S.INTEGRATOR_LIGHT)


def sensor_power(img, size):
    return img[..., 1].sum() * size[0] * size[1] / (img.shape[0] * img.shape[1])


@pytest.mark.parametrize("case", ["lens", "mirror", "water"])
def test_particles_meet_the_reference_power_budgets(oracle, case):
    mk, ppp, size, budget = {
        "lens": (B.sun_lens_scene, 200, (0.024, 0.024), B.lens_budget()),
        "mirror": (B.sun_mirror_scene, 1000, (0.010, 0.010), B.mirror_budget()),
        "water": (B.sun_water_scene, 8000, (0.2, 0.2), B.water_budget() * 0.04),
    }[case]
    img, stats = World(oracle, mk()).render_particles(LIGHT, ppp)
    P = sensor_power(img, size)
    print(f"{case}: power {P:.6e}, budget {budget:.6e}, ratio {P / budget:.4f}")
    assert P / budget == pytest.approx(1.0, abs=0.01)
    assert int(stats[2]) == 0
    peak = img[..., 1].max() / B.sun_irradiance()
    if case == "lens":
        assert peak > 150
    if case == "mirror":
        assert peak > 1000


def test_ripple_caustics_follow_the_ray_map_jacobian(oracle):
    A = B.ripple_analytic(n=2_000_000)
    img, _ = World(oracle, B.ripple_scene()).render_particles(LIGHT, 20000)
    R = img[..., 1].mean(axis=0)
    E0 = B.sun_irradiance()
    ratio, l1 = R.sum() / A.sum(), np.abs(R - A).sum() / A.sum()
    print(f"ripple: ratio {ratio:.4f}, L1 {l1:.3f}, peak {R.max() / E0:.1f} E0 at {R.argmax()} "
          f"(analytic {A.max() / E0:.1f} at {A.argmax()})")
    assert ratio == pytest.approx(1.0, abs=0.01)
    assert l1 < 0.06
    da = abs(int(A.argmax()) - int(R.argmax()))
    assert da <= 1 or abs(da - 80) <= 1
    assert A.max() > 3 * E0


@pytest.mark.parametrize("kind", ["lens", "waves", "waves-sampled", "meniscus"])
def test_uniform_sky_through_lossless_optics_stays_uniform(oracle, kind):
    img, stats = World(oracle, B.uniform_sky_optics_scene(kind, 16)).render(S.settings(seed=3, max_depth=64, rr_depth=8),
                                                                             256)
    y = img[..., 1]
    print(f"{kind}: mean {y.mean():.4f}, range [{y.min():.3f}, {y.max():.3f}], stats {stats}")
    assert y.mean() == pytest.approx(1.0, abs=0.01)
    assert int(stats[2]) == 0 and int(stats[3]) == 0
