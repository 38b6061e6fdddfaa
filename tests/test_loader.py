"""Scene ingest: the `.owe` language and the loader, against the reference's own scene texts and tests.

* the language: numbers, units on numbers and tuples, one-element tuples, lists, calls with keyword
  arguments, nested blocks, both comment forms, identifiers with '-' and '.', errors that name the
  line; edits ``Block.key=value`` override the file (named blocks, unnamed ones by keyword, render
  created when absent);
* the loader against the reference's tests/test_references.cpp: the prism of 278-320 sends each
  wavelength out where a two-face Snell trace does (1e-9 rad) and deviates the d line by the
  classical 66.4° (geometry, units, placement and the N-SF11 entry pinned together); the lens,
  water and mirror scenes of 77-175, loaded from their scene text, meet the same power budgets in
  light tracing through the oracle within 1 % (lens design from a focal length, box media, the
  mirror body and sensors);
* the reference's benchmark and regression scene files load (when the reference tree is present),
  and what the hip backend cannot render is refused with the line it is on;
* helpers: ripple fields draw from PCG32 exactly as the oracle's generator; the catalogue's Schott
  glasses give their catalogue n_d; rgb(1, 1, 1) is the flat reflectance and the 6504 K illuminant.
"""
from __future__ import annotations

import math
import re
from pathlib import Path

import numpy as np
import pytest

from owe_hip import bodies as B
from owe_hip import loader as LD
from owe_hip import output as O
from owe_hip import owe_format as F
from owe_hip import scene as S
from pyoracle import World

REF = Path("/root/reference")


# ---------------------------------------------------------------- the language


def test_values_units_and_blocks():
    doc = F.parse("""
        units = mm   // top-level setting
        # a comment
        body A-1.x { type = sheet  size = (2, 3) cm  r = 48 mm; angle = 12deg  t = 5000K
                     list = [N-BK7, F2]  f = sphere(R = -120, 1e-3)  one = (7)  s = "a b"  big = inf
                     assembly Inner { position = (0, 0, 1) m } }
    """, "t.owe")
    assert [k for k, _ in doc.named] == ["units"]
    b = doc.items[0]
    assert (b.str, b.name) == ("body", "A-1.x")
    size = b.get("size")
    assert size.kind == "tuple" and [x.num for x in size.items] == pytest.approx([0.02, 0.03])
    assert all(x.unit == F.LENGTH for x in size.items)
    assert b.get("r").num == pytest.approx(0.048) and b.get("angle").num == pytest.approx(math.radians(12))
    assert b.get("t").unit == F.TEMPERATURE and b.get("t").num == 5000
    assert [x.str for x in b.get("list").items] == ["N-BK7", "F2"]
    f = b.get("f")
    assert f.kind == "call" and f.str == "sphere" and f.get("R").num == -120 and f.items[0].num == 1e-3
    assert b.get("one").kind == "number" and b.get("one").num == 7
    assert b.get("s").kind == "string" and b.get("s").str == "a b"
    assert math.isinf(b.get("big").num)
    inner = b.items[0]
    assert inner.str == "assembly" and inner.name == "Inner" and inner.get("position").items[2].num == 1.0
    neg = F.parse("x = -y").get("x")
    assert neg.kind == "ident" and neg.str == "-y"


@pytest.mark.parametrize("text,line,msg", [("a = (1, 2", 1, "expected ')'"), ("\n\nbody X { a = 1", 3, "end of file"),
                                           ("}", 1, "unmatched"), ("a = 2 mm mm", 1, "expected '{'"), ("= 1", 1, "expected a keyword"),
                                           ("a = \"x\n\"", 1, "unterminated string"), ("a = 1 ?", 1, "unexpected")])
def test_errors_name_the_line(text, line, msg):
    with pytest.raises(F.SceneError, match=f":{line}: .*{re.escape(msg)}"):
        F.parse(text, "e.owe")


def test_edits_override_the_file():
    doc = F.parse("world { sun { elevation = 10deg } }\nbody Lens { focal = 100 }\nbody Other { focal = 5 }")
    F.apply_edit(doc, "Lens.focal=80")
    F.apply_edit(doc, "sun.elevation=30deg")
    F.apply_edit(doc, "render.spp=64")
    assert doc.items[1].get("focal").num == 80 and doc.items[2].get("focal").num == 5
    assert doc.items[0].items[0].get("elevation").num == pytest.approx(math.radians(30))
    assert doc.items[-1].str == "render" and doc.items[-1].get("spp").num == 64
    with pytest.raises(F.SceneError, match="no block named"):
        F.apply_edit(doc, "Nope.x=1")
    with pytest.raises(F.SceneError, match="Block.key=value"):
        F.apply_edit(doc, "novalue")


# ---------------------------------------------------------------- the reference's scenes


def _refract(d, n, eta):
    """Snell: d and the normal n facing the incoming ray, eta = n_in / n_out."""
    c = -np.dot(n, d)
    k = 1 - eta * eta * (1 - c * c)
    assert k > 0
    return eta * d + (eta * c - math.sqrt(k)) * n


def _cross_line(o, d, a, b):
    """Where the 2-D ray o + t d meets the line through a, b (z ignored)."""
    e = b - a
    M = np.array([[d[0], -e[0]], [d[1], -e[1]]])
    t, _ = np.linalg.solve(M, (a - o)[:2])
    return o + t * d


def test_prism_deviates_each_wavelength_as_snell_predicts(oracle):
    ls = LD.load_scene_text("""
        units = mm
        body Prism { type = prism  apex = 60deg  side = 50  length = 40  medium = N-SF11  position = (0, 0, 0) }
        observer Eye { position = (0, -300, 0)  look_at = (0, 0, 0) }
    """)
    w = World(oracle, ls.scene)
    apex, side = math.radians(60), 0.05
    hx, hy = side * math.sin(apex / 2), side * math.cos(apex / 2)
    A, Bv, C = np.array([0, 2 * hy / 3, 0]), np.array([-hx, -hy / 3, 0]), np.array([hx, -hy / 3, 0])
    d0 = np.array([math.cos(math.radians(33.2)), math.sin(math.radians(33.2)), 0.0])
    origin = (A + Bv) / 2 - 0.2 * d0
    glass, air = LD.catalog_medium("N-SF11"), LD.catalog_medium("air")
    dev = []
    for lam in (404.7, 486.1, 546.1, 587.6, 656.3, 706.5):
        n = O.medium_index(glass, lam) / O.medium_index(air, lam)
        # The loaded world: the entry face and the exit face as the oracle's walk finds them.
        h1 = w.intersect(origin[None], d0[None])[0]
        p1, n1 = h1[1:4], h1[4:7]
        n1 = n1 if np.dot(n1, d0) < 0 else -n1
        d1 = _refract(d0, n1, 1 / n)
        h2 = w.intersect(p1[None], d1[None], np.zeros(1, np.uint32))[0]
        p2, n2 = h2[1:4], h2[4:7]
        n2 = n2 if np.dot(n2, d1) < 0 else -n2
        d2 = _refract(d1, n2, n)
        # The analytic two-face trace of the same triangle.
        q1 = _cross_line(origin, d0, A, Bv)
        e1 = _refract(d0, -np.array([-(Bv - A)[1], (Bv - A)[0], 0]) / np.linalg.norm(Bv - A), 1 / n)
        q2 = _cross_line(q1, e1, A, C)
        e2 = _refract(e1, np.array([-(C - A)[1], (C - A)[0], 0]) / np.linalg.norm(C - A) * -1, n)
        assert np.abs(p1 - q1).max() < 1e-12 and np.abs(p2 - q2).max() < 1e-12
        measured = math.atan2(np.cross(d0, d2)[2], np.dot(d0, d2))
        expected = math.atan2(np.cross(d0, e2)[2], np.dot(d0, e2))
        assert measured == pytest.approx(expected, abs=1e-9)
        dev.append(-math.degrees(measured))
    assert all(dev[i] < dev[i - 1] for i in range(1, len(dev)))
    assert dev[3] == pytest.approx(66.4, abs=0.2)


LIGHT = S.settings(seed=1, max_depth=32, rr_depth=6, integrator=S.INTEGRATOR_LIGHT)

SCENES = {
    "lens": ("""
        units = mm
        world { sun { elevation = 90deg  azimuth = 0deg  radiance = 1 } }
        body Lens { type = lens  medium = N-BK7  focal = 100  form = bi  diameter = 40  thickness = 8  rim = black
                    position = (0, 0, 100)  axis = (0, 0, 1) }
        sensor Focus { size = (24, 24)  resolution = (48, 48)  position = (0, 0, -1.5)  axis = (0, 0, 1)  up = (0, 1, 0) }
        render { detector = Focus }
    """, 200, (0.024, 0.024), B.lens_budget),
    "water": ("""
        units = m
        world { sun { elevation = 40deg  azimuth = 0deg  radiance = 1 } }
        body Water { type = box  size = (1.2, 1.2, 0.1)  position = (0, 0, 0.3)  medium = water }
        sensor Floor { size = (0.2, 0.2)  resolution = (16, 16)  position = (0, 0, 0)  axis = (0, 0, 1)  up = (0, 1, 0) }
        render { detector = Floor }
    """, 8000, (0.2, 0.2), lambda: B.water_budget() * 0.04),
    "mirror": ("""
        units = mm
        world { sun { elevation = 90deg  azimuth = 0deg  radiance = 1 } }
        material coating { type = mirror  reflectance = 0.9 }
        body Mirror { type = mirror  surface = sphere(R = 400)  diameter = 100  position = (0, 0, 0)  axis = (0, 0, 1)
                      material = coating }
        sensor Focus { size = (10, 10)  resolution = (20, 20)  position = (0, 0, 199)  axis = (0, 0, -1)  up = (0, 1, 0) }
        render { detector = Focus }
    """, 1000, (0.010, 0.010), B.mirror_budget),
}


@pytest.mark.parametrize("case", sorted(SCENES))
def test_loaded_reference_scenes_meet_the_power_budgets(oracle, case):
    text, ppp, size, budget = SCENES[case]
    ls = LD.load_scene_text(text)
    assert ls.active == {"lens": "Focus", "water": "Floor", "mirror": "Focus"}[case]
    img, stats = World(oracle, ls.scene).render_particles(LIGHT, ppp)
    P = img[..., 1].sum() * size[0] * size[1] / (img.shape[0] * img.shape[1])
    print(f"{case}: power {P:.6e}, budget {budget():.6e}, ratio {P / budget():.4f}")
    assert P / budget() == pytest.approx(1.0, abs=0.01)
    assert int(stats[2]) == 0


def test_lens_from_a_focal_length_matches_the_lens_makers_equation():
    ls = LD.load_scene_text(SCENES["lens"][0])
    sags = [b for b in ls.scene.boundaries if b["shape"] == S.SHAPE_SAG]
    assert len(sags) == 2
    n = B.relative_index(S.N_BK7, 587.5618)
    R1, R2 = B.design_simple_lens(0.100, n, 0.008)
    assert sags[0]["params"][0] == pytest.approx(1 / R1, rel=1e-12)
    assert sags[1]["params"][0] == pytest.approx(1 / R2, rel=1e-12)
    assert sags[0]["a"] == pytest.approx(0.020) and sags[1]["xf"][2, 3] == pytest.approx(0.108)


def test_units_placement_and_edits_reach_the_world():
    text = """
        units = cm
        body Ball { type = sphere  radius = 2  position = (1, 2, 3)  medium = N-BK7 }
        body Slab { type = box  size = (10, 10, 1) mm  position = (0, 0, 50)  rotate = (90, 0, 0)  material = grey }
        observer Eye { position = (0, -1, 0) m  look_at = (0, 0, 0)  fov = 30deg  pupil = 4 mm  focus = 0.5 m }
    """
    ls = LD.load_scene_text(text, edits=("Ball.radius=3", "Eye.focus=inf"))
    ball = ls.scene.boundaries[0]
    assert ball["shape"] == S.SHAPE_SPHERE and ball["a"] == pytest.approx(0.03)
    assert ball["xf"][:, 3] == pytest.approx([0.01, 0.02, 0.03])
    slab = ls.scene.boundaries[1]
    assert slab["xf"][:, :3] == pytest.approx(S.rotation((1, 0, 0), math.pi / 2))
    obs = ls.scene.observer
    assert obs["pupil_radius"] == pytest.approx(0.002) and math.isinf(obs["focus_distance"])
    assert obs["fov_y"] == pytest.approx(math.radians(30)) and obs["position"] == pytest.approx((0, -1, 0))


@pytest.mark.parametrize("text,msg", [
    ("material m { type = diffuse  texture = checker(a = 1, b = 0) }", "line 1: 'texture' is not supported"),
    ("world { sky = gradient(zenith = 1, horizon = 1) }", "gradient skies are not supported"),
    ("\nbody V { type = vessel }", "line 2: body type 'vessel' is not supported"),
    ("body X { type = sphere  medium = unobtainium }", "unknown medium 'unobtainium'"),
    ("camera C { lens = \"x.lens\" }", "physical cameras are not supported"),
    ("body S { type = sheet size = (1, 1) }", "defines no observer"),
])
def test_what_the_backend_cannot_render_is_refused(text, msg):
    with pytest.raises(F.SceneError, match=msg):
        LD.load_scene_text(text)


def test_integrators_of_other_estimators_are_named():
    ls = LD.load_scene_text("observer E { position = (0, -1, 0)  look_at = (0, 0, 0) }\n"
                            "render { integrator = mlt  seed = 9  max_depth = 12 }")
    with pytest.raises(F.SceneError, match="integrator 'mlt'"):
        ls.settings()
    st = ls.settings(integrator="path")
    assert st.seed == 9 and st.max_depth == 12 and st.integrator == S.INTEGRATOR_PATH
    st = LD.load_scene_text("observer E { position = (0, -1, 0)  look_at = (0, 0, 0) }\n"
                            "render { integrator = hybrid }").settings()
    assert st.integrator == S.INTEGRATOR_HYBRID


def test_vcm_and_bdpt_settings_come_from_the_render_block():
    ls = LD.load_scene_text("observer E { position = (0, -1, 0)  look_at = (0, 0, 0) }\n"
                            "render { integrator = vcm  photon_radius = 3  photons = 2  vcm_groups = 2 }")
    st = ls.settings()
    assert st.integrator == S.INTEGRATOR_VCM
    assert (st.photon_radius, st.photons_per_pixel, st.photon_groups) == (3.0, 2.0, 2)
    st = LD.load_scene_text("observer E { position = (0, -1, 0)  look_at = (0, 0, 0) }\n"
                            "render { integrator = bdpt }").settings()
    assert st.integrator == S.INTEGRATOR_BDPT and st.photon_groups == 1


def test_sppm_settings_come_from_the_render_block():
    ls = LD.load_scene_text("observer E { position = (0, -1, 0)  look_at = (0, 0, 0) }\n"
                            "render { integrator = sppm  photon_radius = 3  photon_alpha = 0.5  photons = 2 }")
    st = ls.settings()
    assert st.integrator == S.INTEGRATOR_SPPM
    assert (st.photon_radius, st.photon_alpha, st.photons_per_pixel, st.photon_groups) == (3.0, 0.5, 2.0, 4)
    assert LD.load_scene_text("observer E { position = (0, -1, 0)  look_at = (0, 0, 0) }\n"
                              "render { integrator = sppm }").settings().photon_alpha == pytest.approx(2 / 3)
    bad = LD.load_scene_text("observer E { position = (0, -1, 0)  look_at = (0, 0, 0) }\n"
                             "render { integrator = sppm  photon_alpha = 1.5 }")
    with pytest.raises(F.SceneError, match="photon_alpha in"):
        bad.settings()


BENCHMARKS = ["caustic_glass_sphere", "caustic_seen_in_mirror", "caustic_seen_through_glass", "caustic_water_surface",
              "concave_mirror_image", "dispersion_glass_sphere", "dispersion_prism", "double_refraction_slabs",
              "physical_lens_free_observer", "telescope_exit_pupil_walk"]


@pytest.mark.skipif(not REF.exists(), reason="the reference tree is not present")
@pytest.mark.parametrize("name", BENCHMARKS)
def test_reference_benchmark_scenes_load(name):
    ls = LD.load_scene(str(REF / "benchmarks" / f"{name}.owe"))
    sc = ls.scene
    c = sc.c()
    assert c.n_boundaries == len(sc.boundaries) > 0
    d = ls.detectors[ls.detector_index()]
    assert (sc.observer["width"], sc.observer["height"]) == tuple(d["resolution"])
    if name == "telescope_exit_pupil_walk":
        # The observer sits at the afocal refractor's exit pupil: 2.0388 m behind the objective
        # along −z (the scope points at −z), a 0.59 mm pupil for a ×126 telescope.
        px = ls.instruments["Scope"]["paraxial"]
        assert px["afocal"] and px["angular_magnification"] == pytest.approx(-126.28, abs=0.01)
        assert sc.observer["position"] == pytest.approx((0, 0, -px["exit_pupil_z"]))
        assert px["exit_pupil_radius"] == pytest.approx(150e-3 / 2 / 126.28, rel=1e-3)


@pytest.mark.skipif(not REF.exists(), reason="the reference tree is not present")
def test_reference_scenes_beyond_the_backend_are_refused_not_misread():
    for f in sorted((REF / "scenes").glob("*.owe")) + [REF / "benchmarks" / "glossy_reflection.owe"]:
        try:
            LD.load_scene(str(f))
        except F.SceneError as e:
            assert "not supported by the hip backend" in str(e), str(e)


# ---------------------------------------------------------------- helpers


def test_ripple_field_draws_from_pcg32(oracle):
    step, _ = LD.pcg32_stream(7, 0x7269707063)
    got = np.array([step() for _ in range(16)], dtype=np.uint32)
    assert (got == oracle.pcg(7, 0x7269707063, 16)).all()
    waves = LD.ripple_field(24, 1, 0.01, 0.05, 0.1, 0.0, math.radians(90))
    assert len(waves) == 24
    ks = np.array([math.hypot(w[1], w[2]) for w in waves])
    assert (2 * math.pi / ks >= 0.01 - 1e-12).all() and (2 * math.pi / ks <= 0.05 + 1e-12).all()
    assert sum((w[0] * k) ** 2 for w, k in zip(waves, ks)) == pytest.approx(0.1 ** 2)   # Σ (a k)² = slope²


@pytest.mark.parametrize("name,nd", [("N-BK7", 1.5168), ("F2", 1.62004), ("N-SF11", 1.78472), ("N-SF6", 1.80518),
                                     ("N-SK16", 1.62041), ("N-BAF10", 1.67003)])
def test_catalogue_glasses_give_their_d_line_index(name, nd):
    m, air = LD.catalog_medium(name), LD.catalog_medium("air")
    assert O.medium_index(m, 587.5618) / O.medium_index(air, 587.5618) == pytest.approx(nd, abs=2e-5)


def test_white_rgb_is_flat_and_colours_reproduce():
    assert LD.rgb_weights(1, 1, 1) == pytest.approx([1, 1, 1], abs=1e-12)
    s = LD.rgb_spectrum(1, 1, 1, False, 0.5)
    assert s.kind == S.SPEC_CONSTANT and s.value * s.scale == 0.5
    e = LD.rgb_spectrum(1, 1, 1, True, 3.0)
    assert e.kind == S.SPEC_BLACKBODY and e.kelvin == 6504.0
    # A saturated red reflectance is the red basis weight clamped to [0, 1]: high in the red, low in the blue.
    red = LD.rgb_spectrum(0.8, 0.1, 0.1, False)
    lam, val = zip(*red.table)
    assert np.interp(650, lam, val) > 0.5 and np.interp(450, lam, val) < 0.2


def test_suite_specs_and_table():
    from owe_hip import suite

    assert suite.split_spec("scenes/the_study.owe:Cam") == ("scenes/the_study.owe", "Cam")
    assert suite.split_spec("a/b.owe") == ("a/b.owe", "")
    assert suite.split_spec("C:/x/b.owe") == ("C:/x/b.owe", "")
    rows = [dict(scene="x/p.owe", detector="Eye", resolution="8x6", spp=2, mpaths_per_s=1.5, msegs_per_s=3.0,
                 seconds=0.01, mean_Y=0.25, digest="0123456789abcdef")]
    t = suite.format_table(dict(rows=rows, total_seconds=0.01, mpaths_per_s=1.5))
    assert "p.owe" in t and "0123456789abcdef" in t and t.splitlines()[-1].startswith("total: 0.010 s")
    for path, _ in suite.DEFAULT_CASES:
        assert LD.load_scene(path).scene.c() is not None
