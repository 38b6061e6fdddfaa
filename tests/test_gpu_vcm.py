"""Vertex connection and merging (vcm) and bidirectional path tracing (bdpt) against the closed forms
and cases the reference checks its own implementation with.

* the white furnace for every integrator, L = Le / (1 − ρ) within 2 % (tests/test_backends.cpp:540-574);
* the lens caustic seen through a 5 mm N-BK7 plate: through / direct within 0.015 of the plate's
  ȳ-weighted (1 − R)²/(1 − R²) at 48×48 and 2048 iterations (tests/test_vcm.cpp:311-351 — the eye
  looks through glass at a point lit through glass: merging's case);
* both draw the path tracer's image of a scene every strategy reaches (tests/test_vcm.cpp:297-309:
  mean within 0.6 % there; here within 1.5 % of the path tracer at the sample counts below);
* light tracing, connections and merging all contribute: vcm's image equals bdpt's in the mean
  where merging is redundant.
"""
from __future__ import annotations

import numpy as np
import pytest

from owe_hip import caustic_scenes as C
from owe_hip import render
from owe_hip import scene as S

pytestmark = pytest.mark.gpu


def _render(sc, st, passes, spp):
    r = render.Renderer(sc, st)
    for _ in range(passes):
        r.run_pass(spp)
    return r


@pytest.mark.parametrize("integrator", [S.INTEGRATOR_VCM, S.INTEGRATOR_BDPT])
@pytest.mark.parametrize("rho", [0.0, 0.5])
def test_white_furnace(cuda, integrator, rho):
    name = "vcm" if integrator == S.INTEGRATOR_VCM else "bdpt"
    print(f"{name} furnace rho={rho}: 8x8, 2048 iterations", flush=True)
    r = _render(S.white_furnace_scene(rho, size=8), S.settings(seed=1, integrator=integrator), 1, 2048)
    y = r.resolve()[..., 1].mean()
    s = r.stats()
    print(f"{name} furnace rho={rho}: {y:.5f} (expected {1 / (1 - rho):.5f}); {s}", flush=True)
    assert y == pytest.approx(1 / (1 - rho), rel=0.02)
    assert s["leaks"] == 0 and s["inconsistencies"] == 0


def test_vcm_sees_a_caustic_through_a_glass_plate(cuda):
    cases = C.caustic_cases(resolution=48)
    means = {}
    for name in ("direct", "plate"):
        ls = cases[name]
        print(f"vcm caustic {name}: 48x48, 2048 iterations", flush=True)
        r = _render(ls.use(), ls.settings(integrator="vcm"), 1, 2048)
        means[name] = r.resolve()[..., 1].mean()
        s = r.stats()
        print(f"  mean Y {means[name]:.6g}; {s}", flush=True)
        assert s["leaks"] == 0
    ratio = means["plate"] / means["direct"]
    print(f"vcm through the plate {ratio:.4f} (Fresnel {C.plate_transmission():.4f})", flush=True)
    assert ratio == pytest.approx(C.plate_transmission(), abs=0.015)


def _sky_prism():
    sc = S.dispersion_prism_scene(32, 32, sun=False)
    sc.set_uniform_sky(S.Spectrum.constant(0.3))
    return sc


@pytest.mark.parametrize("integrator", [S.INTEGRATOR_VCM, S.INTEGRATOR_BDPT])
def test_draws_the_path_tracers_image(cuda, integrator):
    name = "vcm" if integrator == S.INTEGRATOR_VCM else "bdpt"
    sc = _sky_prism()
    path = _render(sc, S.settings(seed=2, max_depth=16, rr_depth=6), 4, 1024).resolve()[..., 1]
    print(f"{name} against path on the sky-lit prism, 32x32: path 4096 spp, {name} 4096 iterations", flush=True)
    img = _render(sc, S.settings(seed=3, max_depth=16, rr_depth=6, integrator=integrator), 4, 1024).resolve()[..., 1]
    # 4×4-pixel blocks: the relative error of each block's mean.
    blocks = lambda a: a.reshape(8, 4, 8, 4).mean(axis=(1, 3))  # noqa: E731
    rel = np.abs(blocks(img) - blocks(path)) / np.maximum(blocks(path), 1e-3 * path.mean())
    print(f"{name}: mean {img.mean():.6g} path {path.mean():.6g} ratio {img.mean() / path.mean():.4f}; "
          f"worst block {rel.max():.4f}", flush=True)
    assert img.mean() / path.mean() == pytest.approx(1.0, abs=0.015)
    assert np.median(rel) < 0.05


def test_unsupported_cases_are_refused(cuda):
    from owe_hip import _lib

    with pytest.raises(_lib.OweHipError):
        render.Renderer(S.submerged_sensor_scene(8), S.settings(integrator=S.INTEGRATOR_BDPT))


@pytest.mark.parametrize("integrator", [S.INTEGRATOR_VCM, S.INTEGRATOR_BDPT])
def test_scattering_furnace(cuda, oracle, integrator):
    """A medium that scatters: the bidirectional paths take it as extinction and a camera pass adds
    the light that reached the eye through a medium scattering vertex (gpu_renderer.cpp:43-47). The
    furnace stays at 1/(1 − ρ); after two iterations every pixel equals the oracle's sum of the same
    two estimates from the same streams."""
    from pyoracle import World

    name = "vcm" if integrator == S.INTEGRATOR_VCM else "bdpt"
    rho = 0.5
    sc = S.scattering_furnace_scene(1.5, 0.0, 0.5, rho, size=8)
    st = S.settings(seed=8, integrator=integrator, max_depth=200)
    print(f"{name} scattering furnace: 8x8, 2048 iterations", flush=True)
    r = _render(sc, st, 1, 2048)
    y = r.resolve()[..., 1].mean()
    print(f"  {y:.5f} (expected 2); {r.stats()}", flush=True)
    assert y == pytest.approx(1 / (1 - rho), rel=0.02)
    g = _render(sc, st, 1, 2).resolve()[..., 1]
    w = World(oracle, sc)
    o = w.vcm(st, 2, merging=integrator == S.INTEGRATOR_VCM)["image"][..., 1] + w.render(st, 2, layout=1)[0][..., 1]
    close = np.abs(g - o) <= 1e-3 * np.maximum(np.abs(g), np.abs(o)) + 1e-7 * float(np.abs(o).max())
    print(f"  after 2 iterations {close.mean():.4f} of pixels equal the oracle's", flush=True)
    assert close.mean() > 0.9


# The caustic through the plate (every strategy: light tracing, next-event estimation, connections
# through glass refused, merging), the reference's benchmark scene of a lens caustic seen through a
# tilted plate, and the pool (absorbing rippled water, contacts); bdpt on the plate; three
# iterations of the plate (the shrinking radius, the rotation of the wavelengths).
@pytest.mark.parametrize("case,integrator,iterations", [("plate", "vcm", 1), ("plate", "bdpt", 1),
                                                        ("caustic_seen_through_glass", "vcm", 1), ("pool", "vcm", 1),
                                                        ("plate", "vcm", 3)])
def test_matches_the_oracle_pixel_by_pixel(cuda, oracle, case, integrator, iterations):
    """The cache is filled in the same order on both (light subpath by light subpath), so the camera
    path of a pixel draws the same light vertex to connect to and merges with the same vertices:
    after a few iterations the images agree pixel by pixel but for float decisions (a roulette or
    Fresnel draw at the threshold, a vertex at the merging radius)."""
    from pathlib import Path

    from owe_hip import loader as LD
    from pyoracle import World

    if case in ("plate",):
        ls = C.caustic_cases(resolution=32)[case]
        sc = ls.use()
    else:
        ls = LD.load_scene(str(Path(LD.__file__).parent / "scenes" / f"{case}.owe"))
        sc = ls.use(resolution=(32, 24))
    st = ls.settings(seed=5, integrator=integrator)
    print(f"{integrator} {case}: {sc.observer['width']}x{sc.observer['height']}, {iterations} iterations", flush=True)
    r = _render(sc, st, 1, iterations)
    g = r.resolve()[..., 1]
    o = World(oracle, sc).vcm(st, iterations, merging=integrator == "vcm")
    y = o["image"][..., 1]
    close = np.abs(g - y) <= 1e-3 * np.maximum(np.abs(g), np.abs(y)) + 1e-7 * max(float(np.abs(y).max()), 1e-30)
    lit = y > 0
    print(f"  agrees on {close.mean():.4f} of pixels ({lit.mean():.3f} lit); mean gpu {g.mean():.6g} oracle "
          f"{y.mean():.6g}; cache {list(o['cached'])}", flush=True)
    assert lit.mean() > 0.2
    assert close.mean() > 0.95
    s = r.stats()
    assert s["paths"] == o["stats"][0]
    assert s["leaks"] == o["stats"][3] == 0


@pytest.mark.parametrize("integrator", ["vcm", "bdpt"])
def test_iterations_in_flight_give_the_same_image(cuda, integrator):
    """Iterations run together, each in its own slot of the cache and the merging hash, give the
    image of iterations run one at a time bit for bit (each iteration's cache is filled in subpath
    order and its merges are summed in integers), and a second renderer repeats it exactly."""
    ls = C.caustic_cases(resolution=48)["plate"]
    sc, st = ls.use(), ls.settings(seed=9, integrator=integrator)
    images = []
    for inflight in (1, 0, 3):
        r = render.Renderer(sc, st)
        r.set_vcm_in_flight(inflight)
        n, most = r.vcm_in_flight()
        print(f"{integrator}: {n} iterations in flight (at most {most}), 2 passes of 7", flush=True)
        r.run_pass(7)
        r.run_pass(7)
        images.append(r.resolve())
        r.close()
    assert most > 1
    assert np.array_equal(images[0], images[1]) and np.array_equal(images[0], images[2])
    r = render.Renderer(sc, st)
    with pytest.raises(Exception):
        r.set_vcm_in_flight(most + 1)
    r.close()


def _sky_plate(glass: bool):
    from owe_hip import loader as LD

    return LD.load_scene_text("""
units = m
world { sky = uniform(rgb(0.7, 0.3, 1, luminance = 1)) }
""" + ("""
body Plate { type = box  size = (20, 20, 0.01)  position = (0, 0, 0)  medium = N-BK7 }
""" if glass else "") + """
observer Eye { position = (0, 0, 1)  look_at = (0, 0, 0)  up = (0, 1, 0)  fov = 20deg  pupil = 1mm  focus = 1
               resolution = (8, 8) }
render { detector = Eye  max_depth = 16 }
""")


def test_dispersive_plate_keeps_the_sky_colour(cuda):
    """A lossless dispersive plate in an isotropic coloured sky cannot change its radiance; the glass
    terminates the secondary wavelengths, so each channel of vcm's and bdpt's image through it must
    equal the path tracer's without it (tests/test_vcm.cpp:260-295: within 1.5 %)."""
    ref_ls = _sky_plate(False)
    ref = _render(ref_ls.use(), ref_ls.settings(integrator="path", seed=11), 1, 4096).resolve()
    ls = _sky_plate(True)
    for integrator in ("bdpt", "vcm"):
        img = _render(ls.use(), ls.settings(integrator=integrator, seed=11), 1, 4096).resolve()
        ratio = img.sum(axis=(0, 1)) / ref.sum(axis=(0, 1))
        print(f"{integrator} through the dispersive plate, XYZ ratios {ratio}", flush=True)
        assert (ref.sum(axis=(0, 1)) > 0).all()
        assert np.abs(ratio - 1).max() < 0.015
