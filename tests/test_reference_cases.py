"""The reference's benchmark scenes and the closed forms they are built around, on the CPU.

* every scene file of the scene-file benchmark loads, with the detector and render block it names;
* the slab scene: ``beam_shift`` (Snell's law at every face the central ray crosses) reduces to the
  textbook plate shift t·sin(θ − θt)/cos θt for each slab alone, is zero at n = 1 and leaves the
  beam's direction unchanged; the fp64 oracle, scanning a pinhole eye across the beam, finds the
  beam where that shift puts it (the reference's benchmarks/double_refraction_slabs.owe);
* the exit-pupil walk: ``disc_overlap`` against its special cases and a Monte Carlo area, and the
  oracle's image means as the eye walks off the refractor's exit pupil against the overlap — the
  check of tests/test_navigation.cpp:220-275, whose own run measured 0.693, 0.386, 0.139, 0 against
  0.692, 0.395, 0.146, 0.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

from owe_hip import reference_cases as RC
from owe_hip import suite
from pyoracle import World

N_BK7_D = 1.5168  # N-BK7 at 587.6 nm (the catalogue's nd)


@pytest.mark.parametrize("name", RC.NAMES)
def test_benchmark_scenes_load(name):
    ls = RC.load(name)
    d = ls.detectors[ls.detector_index()]
    want = {"dispersion_prism": ("Eye", (640, 400), "sppm"), "caustic_seen_through_glass": ("Eye", (320, 320), "sppm"),
            "caustic_seen_in_mirror": ("Eye", (320, 320), "sppm"),
            "double_refraction_slabs": ("Eye", (480, 320), "sppm"),
            "telescope_exit_pupil_walk": ("Centred", (64, 64), "path")}[name]
    assert (d["name"], tuple(d["resolution"]), ls.render["integrator"]) == want
    sc = ls.use(resolution=(16, 12))
    assert len(sc.boundaries) > 2
    assert (str(RC.path(name)), "") in suite.DEFAULT_CASES


def test_slab_variants_drop_what_they_say():
    full = RC.load_text(RC.double_refraction_slabs()).use(resolution=(8, 8))
    bare = RC.load_text(RC.double_refraction_slabs(slabs=False)).use(resolution=(8, 8))
    assert len(full.boundaries) - len(bare.boundaries) == 2  # two analytic boxes
    ls = RC.load_text(RC.double_refraction_slabs(beam_eye=(0.0, 8)))
    assert [d["name"] for d in ls.detectors] == ["Beam"]
    assert ls.render["detector"] == "Beam"
    with pytest.raises(ValueError):
        RC._drop("body A { x = 1 }\n", "body B")


def test_plate_shift_closed_form():
    assert RC.slab_shift(0.0, 0.012, 1.5) == 0.0
    assert RC.slab_shift(0.7, 0.012, 1.0) == pytest.approx(0.0, abs=1e-15)
    # Small angles: t·θ·(1 − 1/n).
    assert RC.slab_shift(1e-4, 0.012, 1.5) == pytest.approx(0.012 * 1e-4 * (1 - 1 / 1.5), rel=1e-6)
    # Grazing incidence: the shift approaches t·cos θt / cos θt · ... → t (sin(90° − θt)/cos θt = 1).
    assert RC.slab_shift(math.pi / 2 - 1e-9, 0.012, 1.5) == pytest.approx(0.012, rel=1e-6)


def _through_one(centre, normal, n):
    o, d = np.array(RC.STOP_CENTRE, float), np.array(RC.BEAM)
    o2, d2 = RC._through_box(o, d, centre, normal, n)
    shift = (o2 - o) - d * np.dot(o2 - o, d)
    return shift, d2


def test_each_slab_alone_shifts_as_a_plate():
    # Slab B: the beam crosses its two big faces (12 mm apart) at θ = acos |d·n|.
    d = np.array(RC.BEAM)
    m = np.array(RC.SLAB_NORMALS[1]) / np.linalg.norm(RC.SLAB_NORMALS[1])
    shift, d2 = _through_one(RC.SLAB_CENTRES[1], RC.SLAB_NORMALS[1], N_BK7_D)
    assert np.linalg.norm(shift) == pytest.approx(RC.slab_shift(math.acos(abs(d @ m)), RC.SLAB_T, N_BK7_D), rel=1e-9)
    assert np.allclose(d2, d, atol=1e-12)
    # Slab A: at 85° to its big faces the beam enters and leaves by the 12 mm edge faces, 60 mm apart.
    u = RC._box_axes(RC.SLAB_NORMALS[0])[1]
    assert abs(d @ u) > 0.99
    shift, d2 = _through_one(RC.SLAB_CENTRES[0], RC.SLAB_NORMALS[0], N_BK7_D)
    assert np.linalg.norm(shift) == pytest.approx(RC.slab_shift(math.acos(abs(d @ u)), 0.06, N_BK7_D), rel=1e-9)
    assert np.allclose(d2, d, atol=1e-12)


def test_beam_shift_properties():
    assert np.allclose(RC.beam_shift(1.0), 0.0, atol=1e-12)
    s = RC.beam_shift(N_BK7_D)
    assert abs(s @ np.array(RC.BEAM)) < 1e-12 and abs(s[1]) < 1e-15
    assert RC.beam_shift_across(N_BK7_D) == pytest.approx(1.0677e-3, abs=1e-6)
    # More index, more shift; the two slabs' shifts add with their own signs.
    assert RC.beam_shift_across(1.53) > RC.beam_shift_across(1.50) > 0


def _half_max_midpoint(x, p):
    """The midpoint of the profile's two half-maximum crossings (linear interpolation)."""
    p = np.asarray(p) / np.max(p)
    i = int(np.argmax(p >= 0.5))
    j = len(p) - 1 - int(np.argmax(p[::-1] >= 0.5))
    left = x[i - 1] + (0.5 - p[i - 1]) / (p[i] - p[i - 1]) * (x[i] - x[i - 1])
    right = x[j] + (p[j] - 0.5) / (p[j] - p[j + 1]) * (x[j + 1] - x[j])
    return 0.5 * (left + right)


OFFSETS = np.arange(-5e-3, 7.01e-3, 0.25e-3)


def slab_profile(render_mean, slabs: bool):
    """The sun's image mean as the beam eye moves across the beam, at each of OFFSETS."""
    out = []
    for off in OFFSETS:
        ls = RC.load_text(RC.double_refraction_slabs(slabs=slabs, beam_eye=(float(off), 16)))
        out.append(render_mean(ls.use(), ls.settings(seed=1)))
    return np.array(out)


def test_oracle_finds_the_beam_where_the_slabs_move_it(oracle):
    def mean(sc, st):
        img, _ = World(oracle, sc).render(st, 16, layout=1, threads=8)
        return img[..., 1].mean()

    bare, moved = slab_profile(mean, False), slab_profile(mean, True)
    c0, c1 = _half_max_midpoint(OFFSETS, bare), _half_max_midpoint(OFFSETS, moved)
    want = RC.beam_shift_across(N_BK7_D)
    print(f"beam centre without slabs {c0 * 1e3:.3f} mm, with {c1 * 1e3:.3f} mm; central ray {want * 1e3:.3f} mm; "
          f"transmitted {moved.sum() / bare.sum():.3f}")
    assert abs(c0) < 0.05e-3
    assert c1 - c0 == pytest.approx(want, abs=0.12e-3)
    # Four faces near normal incidence (R ≈ 0.043 each) and the sliver of the beam that slab A's big
    # faces reflect away inside the glass.
    assert 0.78 < moved.sum() / bare.sum() < (1 - 0.042) ** 4


def test_disc_overlap_special_cases():
    assert RC.disc_overlap(1.0, 0.5, 0.0) == pytest.approx(math.pi * 0.25)
    assert RC.disc_overlap(1.0, 0.5, 0.4) == pytest.approx(math.pi * 0.25)
    assert RC.disc_overlap(1.0, 0.5, 1.5) == 0.0
    r = 0.7  # equal discs a radius apart: r²(2π/3 − √3/2)
    assert RC.disc_overlap(r, r, r) == pytest.approx(r * r * (2 * math.pi / 3 - math.sqrt(3) / 2), rel=1e-12)
    rng = np.random.default_rng(3)
    p = rng.uniform(-1.0, 1.8, size=(2_000_000, 2))
    inside = (np.hypot(p[:, 0], p[:, 1]) < 1.0) & (np.hypot(p[:, 0] - 0.8, p[:, 1]) < 0.6)
    assert RC.disc_overlap(1.0, 0.6, 0.8) == pytest.approx(inside.mean() * 2.8 * 2.8, rel=0.01)


def exit_pupil_walk(render_mean, resolution=32):
    """[(f, seen / centred, expected)] for the eye f·(rx + re) off the exit pupil, the walk's
    centred mean, the mean 1.6·(rx + re) off, and the centred mean after walking back."""
    ls = RC.load("telescope_exit_pupil_walk")
    rx = ls.instruments["Scope"]["paraxial"]["exit_pupil_radius"]
    re = ls.detectors[0]["pupil_radius"]
    assert 0.3 * re < rx < 3 * re

    def at(dx):
        return render_mean(RC.exit_pupil_walk(ls, dx, (resolution, resolution)), ls.settings(seed=1))

    centred = at(0.0)
    rows = []
    for f in (0.25, 0.5, 0.75, 1.0, 1.25):
        dx = f * (rx + re)
        rows.append((f, at(dx) / centred, RC.disc_overlap(rx, re, dx) / RC.disc_overlap(rx, re, 0.0)))
    return rows, centred, at(1.6 * (rx + re)), at(0.0)


def check_walk(rows, centred, lost, back, tight):
    previous = 1.0
    for f, seen, expected in rows:
        print(f"eye {f:.2f}·(rx + re) off the exit pupil: {seen:.4f} of the centred image, disc overlap {expected:.4f}")
        assert seen <= previous * 1.05 + 2e-3  # never brighter beyond noise (stray light at the 1e-3 level once lost)
        assert abs(seen - expected) <= 0.12 + 0.1 * expected  # the reference test's own bound
        assert abs(seen - expected) <= tight
        previous = seen
    assert centred > 0 and lost < 0.02 * centred
    assert back == centred


def test_oracle_exit_pupil_walk_follows_the_disc_overlap(oracle):
    def mean(sc, st):
        img, _ = World(oracle, sc).render(st, 64, layout=1, threads=8)
        return img[..., 1].mean()

    check_walk(*exit_pupil_walk(mean, resolution=32), tight=0.02)
