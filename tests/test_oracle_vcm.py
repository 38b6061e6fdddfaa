"""The oracle's vertex connection and merging (vcm) and bidirectional path tracing (bdpt) against
the closed forms and cases the reference checks its own implementation with (Georgiev et al.
2012's recursive MIS in the balance heuristic, src/owe/backends/gpu/shaders/vcm.slang):

* the white furnace, L = Le / (1 − ρ), for both (tests/test_backends.cpp:540-574): every strategy —
  light tracing, next-event estimation, emission met, connections, merging — contributes there, and
  only weights that sum to one over the strategies give the furnace value;
* the lens caustic seen through a 5 mm N-BK7 plate: through / direct equals the plate's ȳ-weighted
  (1 − R)²/(1 − R²) (tests/test_vcm.cpp:311-351), a case only merging and light tracing reach;
* the sky-lit prism: vcm and bdpt draw the path tracer's image (tests/test_vcm.cpp:297-309), and
  the caustic seen directly, which light tracing and connections reach, is the same image for both;
* the cache holds one vertex per non-specular scattering event of a light subpath.
"""
from __future__ import annotations

import numpy as np
import pytest

from owe_hip import caustic_scenes as C
from owe_hip import scene as S
from pyoracle import World


@pytest.mark.parametrize("integrator", [S.INTEGRATOR_VCM, S.INTEGRATOR_BDPT])
@pytest.mark.parametrize("rho", [0.0, 0.5])
def test_white_furnace(oracle, integrator, rho):
    w = World(oracle, S.white_furnace_scene(rho, size=8))
    r = w.vcm(S.settings(seed=1, integrator=integrator), 512, merging=integrator == S.INTEGRATOR_VCM)
    assert r["image"][..., 1].mean() == pytest.approx(1 / (1 - rho), rel=0.01)
    assert r["stats"][2] == 0 and r["stats"][3] == 0
    # Light tracing (weighted) is a part of the furnace, not all of it; a black wall reflects nothing.
    light = r["light"][..., 1].mean() / (r["paths"] * 512)
    assert (0.005 < light < 0.5) if rho > 0 else light == 0


def test_caustic_through_a_plate(oracle):
    cases = C.caustic_cases(resolution=32)
    means = {}
    for name in ("direct", "plate"):
        ls = cases[name]
        r = World(oracle, ls.use()).vcm(ls.settings(integrator="vcm"), 2048, merging=True)
        means[name] = r["image"][..., 1].mean()
        assert r["stats"][3] == 0
    assert means["plate"] / means["direct"] == pytest.approx(C.plate_transmission(), abs=0.006)


@pytest.mark.parametrize("integrator", [S.INTEGRATOR_VCM, S.INTEGRATOR_BDPT])
def test_draws_the_path_tracers_image(oracle, integrator):
    sc = S.dispersion_prism_scene(32, 32, sun=False)
    sc.set_uniform_sky(S.Spectrum.constant(0.3))
    w = World(oracle, sc)
    path = w.render(S.settings(seed=2, max_depth=16, rr_depth=6), 1024, layout=0)[0][..., 1]
    r = w.vcm(S.settings(seed=3, max_depth=16, rr_depth=6, integrator=integrator), 1024,
              merging=integrator == S.INTEGRATOR_VCM)
    img = r["image"][..., 1]
    blocks = lambda a: a.reshape(8, 4, 8, 4).mean(axis=(1, 3))  # noqa: E731
    rel = np.abs(blocks(img) - blocks(path)) / np.maximum(blocks(path), 1e-3 * path.mean())
    assert img.mean() / path.mean() == pytest.approx(1.0, abs=0.006)
    assert np.median(rel) < 0.01


def test_bdpt_and_vcm_agree_on_a_caustic_seen_directly(oracle):
    """The lens caustic seen directly: light tracing and connections reach it without merging, so
    bdpt and vcm estimate the same image (through the plate only merging reaches it with low
    variance: bdpt's estimate there rests on camera paths that happen to hit the sun's disc)."""
    ls = C.caustic_cases(resolution=32)["direct"]
    w = World(oracle, ls.use())
    vcm = w.vcm(ls.settings(integrator="vcm"), 512, merging=True)["image"][..., 1].mean()
    bdpt = w.vcm(ls.settings(integrator="bdpt"), 512, merging=False)["image"][..., 1].mean()
    assert bdpt == pytest.approx(vcm, rel=0.03)


def test_cache_holds_one_vertex_per_non_specular_event(oracle):
    """In the black furnace (ρ = 0) every light subpath reaches the wall once and stops: one cached
    vertex per subpath whose first segment hits the wall (all of them)."""
    w = World(oracle, S.white_furnace_scene(0.0, size=8))
    r = w.vcm(S.settings(seed=5, integrator=S.INTEGRATOR_VCM), 3, merging=True)
    assert list(r["cached"]) == [r["paths"]] * 3


@pytest.mark.parametrize("integrator", [S.INTEGRATOR_VCM, S.INTEGRATOR_BDPT])
def test_scattering_furnace(oracle, integrator):
    """With a medium that scatters, vcm and bdpt take media as extinction and a camera pass counts
    only light that reached the eye through a medium scattering vertex (gpu_renderer.cpp:43-47):
    together they keep the scattering furnace at 1/(1 − ρ); each alone does not."""
    rho = 0.5
    w = World(oracle, S.scattering_furnace_scene(1.5, 0.0, 0.5, rho, size=8))
    st = S.settings(seed=8, integrator=integrator, max_depth=200)
    its = 512
    bidir = w.vcm(st, its, merging=integrator == S.INTEGRATOR_VCM)["image"][..., 1].mean()
    medium, mst = w.render(st, its, layout=1)  # vcm settings: the medium-only camera pass
    total = bidir + medium[..., 1].mean()
    print(f"bidirectional {bidir:.4f} + medium pass {medium[..., 1].mean():.4f} = {total:.4f}")
    assert total == pytest.approx(1 / (1 - rho), rel=0.015)
    assert 0.1 < medium[..., 1].mean() < 0.9 * total
