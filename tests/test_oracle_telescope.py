"""The oracle's sequential trace through mirrors (reflecting rows, central holes, obstruction by
mirror bodies) and afocal systems, against the reference's own tests and closed forms.

* the Keplerian telescope of tests/test_instruments.cpp:336-355: solveAfocal's gap makes the system
  afocal, within 10 % of f_o + f_e, and the angular magnification equals −f_o/f_e of the thick-lens
  focal lengths to 1e-9;
* the reference's Gregorian (lenses/gregorian_54x.lens) solved afocal at its design wavelength
  magnifies about 54 times, and its apparent field angle is M times the field (to its 0.2 % distortion
  at 0.05°);
* its paraboloid–ellipsoid mirror pair alone images an on-axis star without spherical aberration
  (RMS spot radius below a nanometre), and the rays it loses are exactly those that fall into the
  primary's hole;
* the perturbed trace moves with its parameter smoothly: central differences converge.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

from owe_hip import lens as L
from pyoracle import LensOracle


def thick_lens_efl(n, R1, R2, t):
    return 1 / ((n - 1) * (1 / R1 - 1 / R2 + (n - 1) * t / (n * R1 * R2)))


KEPLER = L.Prescription([L.Surface(200.0, 6.0, "N-BK7", 20), L.Surface(-200.0, 250.0, "air", 20, solve=True),
                         L.Surface(25.0, 5.0, "N-BK7", 8), L.Surface(-25.0, 30.0, "air", 8)], name="kepler")


def test_keplerian_telescope_is_afocal_with_magnification_minus_fo_over_fe(oracle):
    p = L.solve_afocal(KEPLER, L.LAMBDA_D)
    n = L.index("N-BK7", L.LAMBDA_D) / L.ciddor_air(L.LAMBDA_D)
    fo, fe = thick_lens_efl(n, 0.2, -0.2, 0.006), thick_lens_efl(n, 0.025, -0.025, 0.005)
    gap = p.surfaces[1].t * 1e-3
    assert 0.9 * (fo + fe) < gap < 1.1 * (fo + fe)
    _, px = LensOracle(oracle, [p.with_rings(4)], field_deg=0.0).moments(0, L.LAMBDA_D)
    assert px["afocal"]
    assert px["angular_magnification"] == pytest.approx(-fo / fe, rel=1e-9)


def gregorian():
    return L.solve_afocal(L.GREGORIAN, 587.5618)


def test_gregorian_magnifies_54_times(oracle):
    g = gregorian().with_rings(30)
    lo = LensOracle(oracle, [g], field_deg=0.05, reference=587.5618)
    m, px = lo.moments(0, 587.5618)
    s = L.afocal_spot(m)
    M = px["angular_magnification"]
    print(f"Gregorian: M {M:.4f}, apparent angle {s['apparent_angle']:.6f} rad, angular RMS {s['angular_rms']:.3e}")
    assert px["afocal"] and M == pytest.approx(54, rel=0.01)
    assert s["apparent_angle"] == pytest.approx(M * math.radians(0.05), rel=5e-3)  # plus 0.2 % distortion


def test_mirror_pair_images_without_spherical_aberration_and_loses_the_hole(oracle):
    P, S = L.GREGORIAN.surfaces[:2]
    pair = L.Prescription([L.Surface(**{**P.__dict__}), L.Surface(**{**S.__dict__, "t": 800.0, "solve": False})],
                          name="gregorian mirrors").with_rings(40)
    lo = LensOracle(oracle, [pair], field_deg=0.0)
    m, px = lo.moments(0, L.LAMBDA_D)
    s = L.spot(m, px["last_vertex_z"] + px["bfl"], px["last_vertex_z"])
    # The pupil points outside the 13 mm hole (ring k at radius 0.999 · 54 mm · k / 40).
    r = [0.999 * 54.0 * k / 40 for k in range(41) for _ in range(1 if k == 0 else 6 * k)]
    outside = sum(1 for v in r if v >= 13.0)
    print(f"mirror pair: rms {s['rms_best']:.3e} m, arrived {s['arrived']} / {pair.rays}, outside the hole {outside}")
    assert s["rms_best"] < 1e-9
    assert s["arrived"] == outside
    assert s["best_focus_z"] == pytest.approx(px["bfl"], rel=1e-6)


@pytest.mark.parametrize("field", [L.PARAM_CURVATURE, L.PARAM_CONIC, L.PARAM_THICKNESS])
def test_perturbed_traces_converge(oracle, field):
    g = gregorian().with_rings(10)
    lo = LensOracle(oracle, [g], field_deg=0.05, reference=587.5618)
    idx = np.arange(g.rays, dtype=np.uint64)
    base = lo.trace_perturbed(0, 587.5618, [(1, field)], [0.0], idx)
    ok = base[:, 6] > 0
    slopes = []
    for h in (1e-5, 1e-6):
        plus = lo.trace_perturbed(0, 587.5618, [(1, field)], [h], idx)[ok]
        minus = lo.trace_perturbed(0, 587.5618, [(1, field)], [-h], idx)[ok]
        slopes.append((plus[:, 3:5] - minus[:, 3:5]) / (2 * h))
    assert np.abs(slopes[0] - slopes[1]).max() <= 1e-3 * np.abs(slopes[1]).max()
    assert np.abs(slopes[1]).max() > 0
