"""The oracle's sequential lens trace against the reference's own closed forms and tests.

* paraxial EFL against the thick-lens formula and BFL against f(1 − (n−1)t/(nR₁)) to 1e-12 at the
  F, d and C lines (tests/test_instruments.cpp:224-234);
* a near-axis real ray (height 1 µm) through the singlet crosses the axis at the paraxial focus
  within 5e-11 m (tests/test_instruments.cpp:250-266);
* the hit point of a conic sag lies on z = sag(r) to 1e-12 and its normal agrees with the finite
  difference of the sag; the numeric asphere path with a vanishing A4 agrees with the quadric path
  to 1e-12 (tests/test_geometry.cpp sag cases);
* an achromat reduces the F–C focal shift per unit power tenfold against the singlet, and every
  launched ray arrives (tests/test_instruments.cpp:320-333); the singlet's spherical aberration is
  undercorrected;
* focusedSpot from the nine sums equals the direct computation over the traced rays;
* the ring pupil visits spotAt's hexapolar points in its order (analysis.cpp:124-131), and with
  six rings on axis at F, d and C — the reference's own analyzeLens call — the achromat and
  singlet trends above hold as they do under the Fibonacci lattice.
"""
from __future__ import annotations

import ctypes
import math

import numpy as np
import pytest

from owe_hip import lens as L
from pyoracle import LensOracle


def thick_lens_efl(n, R1, R2, t):
    P = (n - 1) * (1 / R1 - 1 / R2 + (n - 1) * t / (n * R1 * R2))
    return 1 / P


@pytest.mark.parametrize("lam", [L.LAMBDA_F, L.LAMBDA_D, L.LAMBDA_C])
def test_paraxial_efl_and_bfl_match_thick_lens_formulas(oracle, lam):
    _, px = LensOracle(oracle, [L.SINGLET], field_deg=0.0, reference=lam).moments(0, lam)
    n = L.index("N-BK7", lam) / L.ciddor_air(lam)
    f = thick_lens_efl(n, 0.05, -0.05, 0.008)
    assert px["efl"] == pytest.approx(f, rel=1e-12)
    assert px["bfl"] == pytest.approx(f * (1 - (n - 1) * 0.008 / (n * 0.05)), rel=1e-12)
    assert px["entrance_pupil_z"] == 0 and px["entrance_pupil_radius"] == pytest.approx(0.0125, rel=1e-15)


@pytest.mark.parametrize("lam", [L.LAMBDA_F, L.LAMBDA_D, L.LAMBDA_C])
def test_near_axis_ray_reaches_the_paraxial_focus(oracle, lam):
    lo = LensOracle(oracle, [L.SINGLET], field_deg=0.0, reference=lam)
    _, px = lo.moments(0, lam)
    out = lo.trace_ray(0, lam, (0.0, 1e-6, -0.05), (0.0, 0.0, 1.0))
    assert out[6] > 0.9
    p, d = out[:3], out[3:6]
    z_cross = p[2] - p[1] * d[2] / d[1]
    assert z_cross == pytest.approx(0.008 + px["bfl"], abs=5e-11)


def test_sag_hit_lies_on_the_surface_with_the_sag_normal(oracle):
    lo = LensOracle(oracle, [L.SINGLET])
    rng = np.random.default_rng(3)
    for R, k in ((0.05, 0.0), (-0.08, -0.6), (0.12, 1.5), (0.03, -1.0)):
        row = L.c_surface(L.Surface(R * 1e3, 5.0, "N-BK7", 0.2 * abs(R) * 1e3, k=k))
        for _ in range(200):
            x, y = rng.uniform(-0.15, 0.15, 2) * abs(R)
            o = np.array([x, y, -0.02])
            d = np.array([rng.normal(0, 0.05), rng.normal(0, 0.05), 1.0])
            d /= np.linalg.norm(d)
            out = lo.sag_intersect(row, o, d)
            if np.isnan(out[0]):
                continue
            p, n = out[1:4], out[4:7]
            r = math.hypot(p[0], p[1])
            assert p[2] == pytest.approx(oracle.sag_value(ctypes.byref(row), r), abs=1e-12)
            h = 1e-6
            slope = (oracle.sag_value(ctypes.byref(row), r + h) - oracle.sag_value(ctypes.byref(row), r - h)) / (2 * h)
            expect = np.array([-slope * p[0] / r, -slope * p[1] / r, 1.0])
            expect /= np.linalg.norm(expect)
            assert np.abs(n - expect).max() < 1e-7


def test_numeric_asphere_path_equals_the_quadric_path(oracle):
    lo = LensOracle(oracle, [L.SINGLET])
    rng = np.random.default_rng(5)
    quad = L.c_surface(L.Surface(60.0, 5.0, "N-BK7", 15.0, k=-0.4))
    num = L.c_surface(L.Surface(60.0, 5.0, "N-BK7", 15.0, k=-0.4, A=(1e-40,)))
    for _ in range(300):
        x, y = rng.uniform(-0.012, 0.012, 2)
        o = np.array([x, y, -0.01])
        d = np.array([rng.normal(0, 0.1), rng.normal(0, 0.1), 1.0])
        d /= np.linalg.norm(d)
        a, b = lo.sag_intersect(quad, o, d), lo.sag_intersect(num, o, d)
        assert np.isnan(a[0]) == np.isnan(b[0])
        if not np.isnan(a[0]):
            assert np.abs(a[:4] - b[:4]).max() < 1e-12


def ring_points(rings, r_ep):
    """spotAt's pupil pattern written out (analysis.cpp:124-131): centre, then 6k points on ring k."""
    pts = []
    for k in range(rings + 1):
        n = 1 if k == 0 else 6 * k
        for j in range(n):
            rr = r_ep * k / rings * 0.999
            a = 2 * math.pi * j / n
            pts.append((rr * math.cos(a), rr * math.sin(a)))
    return np.array(pts)


def test_ring_pupil_walks_spot_at_pattern_in_order(oracle):
    # A flat window on axis passes every ray straight: its exit points are the pupil points.
    window = L.Prescription([L.Surface(math.inf, 5.0, "N-BK7", 10.0, stop=True),
                             L.Surface(math.inf, 10.0, "air", 10.0)]).with_rings(6)
    assert window.rays == 127
    lo = LensOracle(oracle, [window], field_deg=0.0)
    out = lo.trace(0, L.LAMBDA_D, np.arange(window.rays, dtype=np.uint64))
    _, px = lo.moments(0, L.LAMBDA_D)
    assert px["entrance_pupil_radius"] == pytest.approx(0.010, rel=1e-12)
    assert np.abs(out[:, :2] - ring_points(6, px["entrance_pupil_radius"])).max() < 1e-15
    assert (out[:, 6] > 0).all()


@pytest.mark.parametrize("rings", [0, 6])
def test_achromat_cuts_the_chromatic_focal_shift_and_keeps_every_ray(oracle, rings):
    # rings = 6 is the reference's own configuration: analyzeLens(p, {0}, {F, d, C}, 6).
    def focus(p, lam):
        p = p.with_rings(rings) if rings else p
        lo = LensOracle(oracle, [p], field_deg=0.0, reference=lam)
        m, px = lo.moments(0, lam)
        lv = px["last_vertex_z"]
        return L.spot(m, lv + px["bfl"], lv), px, int(p.rays)

    shift = {}
    for name, p in (("singlet", L.SINGLET), ("doublet", L.DOUBLET)):
        sF, pxF, n = focus(p, L.LAMBDA_F)
        sC, _, _ = focus(p, L.LAMBDA_C)
        shift[name] = abs(sF["best_focus_z"] - sC["best_focus_z"]) / pxF["efl"]
        assert sF["arrived"] == n and sC["arrived"] == n
        if name == "singlet":
            assert sF["best_focus_z"] < sC["best_focus_z"]  # normal dispersion: blue focuses closer
    assert shift["doublet"] < 0.1 * shift["singlet"]


def test_singlet_spherical_aberration_is_undercorrected(oracle):
    lo = LensOracle(oracle, [L.SINGLET], field_deg=0.0)
    _, px = lo.moments(0, L.LAMBDA_D)
    zs = []
    for h in (0.5, 0.999):
        out = lo.trace_ray(0, L.LAMBDA_D, (0.0, h * 0.0125, -0.05), (0.0, 0.0, 1.0))
        p, d = out[:3], out[3:6]
        zs.append(p[2] - p[1] * d[2] / d[1] - (0.008 + px["bfl"]))
    assert zs[1] < 0 and abs(zs[1]) > 3 * abs(zs[0])


def test_spot_from_sums_equals_direct_computation(oracle):
    p = L.seeded_achromats(1, seed=7, rays_min=3000, rays_max=3000)[0]
    lo = LensOracle(oracle, [p])
    idx = np.arange(p.rays, dtype=np.uint64)
    rays = lo.trace(0, L.LAMBDA_D, idx)
    ok = rays[:, 6] > 0
    P, D = rays[ok, :3], rays[ok, 3:6]
    B = D[:, :2] / D[:, 2:3]
    A = P[:, :2] - B * P[:, 2:3]
    m, px = lo.moments(0, L.LAMBDA_D)
    assert m[0] == ok.sum()
    lv = px["last_vertex_z"]
    z_img = lv + px["bfl"]
    s = L.spot(m, z_img, lv)
    cov = ((A - A.mean(0)) * (B - B.mean(0))).sum()
    varb = ((B - B.mean(0)) ** 2).sum()
    zb = -cov / varb
    rms = math.sqrt((((A + B * zb) - (A + B * zb).mean(0)) ** 2).sum(1).mean())
    assert s["best_focus_z"] + lv == pytest.approx(zb, rel=1e-9)
    assert s["rms_best"] == pytest.approx(rms, rel=1e-6)


def test_seeded_achromats_follow_the_configured_ranges():
    lenses = L.seeded_achromats(256)
    assert len(lenses) == 256
    rays = np.array([p.rays for p in lenses])
    assert rays.min() >= 4096 and rays.max() <= 1 << 20
    assert 100_000 < rays.mean() < 300_000
    for p in lenses:
        assert 2.0 <= p.surfaces[0].t and 2.0 <= p.surfaces[1].t
        assert 30.0 <= abs(p.surfaces[0].R) <= 120.0
        assert p.surfaces[0].medium in L.CROWNS and p.surfaces[1].medium in L.FLINTS
    assert lenses[0].name == L.seeded_achromats(1)[0].name
