"""The wavefront camera-path renderer on the GPU against the fp64 oracle and closed forms.

* closed forms of the reference's tests (tests/test_transport.cpp:83-142): white furnace
  1/(1 − ρ) within 1 %, an invisible dielectric sphere within 0.4 %;
* per-path parity: the same (pixel, sample, pass) streams traced by the float32 kernels and the
  fp64 oracle give the same XYZ up to float rounding, except the few paths where a float decision
  (Russian roulette, a Fresnel branch, the sun disc's rim) falls the other way;
* image parity: block-z consistency with the oracle (src/owe/backends/compare.cpp:64-119) on the
  dispersion-prism scene, and relative RMSE against a high-sample fp64 image;
* determinism: the integer film makes two identical renders bit-identical.
"""
from __future__ import annotations

import numpy as np
import pytest

from owe_hip import compare, render
from owe_hip import scene as S
from pyoracle import World

pytestmark = pytest.mark.gpu

PRISM = dict(max_depth=16, rr_depth=4, spectral=S.SPECTRAL_UNIFORM1, lambda_min=380.0, lambda_max=780.0)


@pytest.mark.parametrize("rho", [0.5, 0.8])
def test_white_furnace(cuda, rho):
    r = render.Renderer(S.white_furnace_scene(rho), S.settings(seed=1))
    r.run_pass(256)
    img = r.resolve()
    st = r.stats()
    print(f"furnace rho={rho}: mean Y {img[..., 1].mean():.5f} (expect {1 / (1 - rho):.5f}), {st}", flush=True)
    assert img[..., 1].mean() == pytest.approx(1 / (1 - rho), abs=0.01 / (1 - rho))
    assert st["inconsistencies"] == 0 and st["leaks"] == 0


def test_dielectric_sphere_is_invisible(cuda):
    r = render.Renderer(S.dielectric_sphere_scene(), S.settings(seed=1))
    r.run_pass(256)
    img = r.resolve()
    st = r.stats()
    print(f"sphere: mean Y {img[..., 1].mean():.5f}, disk {img[6:18, 6:18, 1].mean():.5f}, {st}", flush=True)
    assert img[..., 1].mean() == pytest.approx(1.0, abs=0.004)
    assert img[6:18, 6:18, 1].mean() == pytest.approx(1.0, abs=0.01)
    assert st["inconsistencies"] == 0 and st["leaks"] == 0


def test_prism_paths_match_oracle_path_by_path(cuda, oracle):
    sc = S.dispersion_prism_scene(64, 64)
    st = S.settings(seed=7, **PRISM)
    spp = 16
    n = 64 * 64 * spp
    pixel = np.repeat(np.arange(64 * 64, dtype=np.uint32), spp)
    sample = np.tile(np.arange(spp, dtype=np.uint32), 64 * 64)
    gpu = render.Renderer(sc, st).camera_paths(spp, pixel, sample)
    ora, segs = World(oracle, sc).camera_paths(st, spp, pixel, sample)
    gy, oy = gpu[:, 1], ora[:, 1]
    scale = np.abs(oy).max()
    close = np.abs(gy - oy) <= 1e-3 * np.abs(oy) + 1e-5 * scale
    frac = close.mean()
    print(f"per-path: {close.sum()}/{n} within 1e-3 rel ({frac:.5f}); nonzero gpu {np.count_nonzero(gy)}, "
          f"oracle {np.count_nonzero(oy)}; mean Y gpu {gy.mean():.5f} oracle {oy.mean():.5f}", flush=True)
    assert frac > 0.995
    assert abs(gy.mean() - oy.mean()) < 0.02 * abs(oy.mean())


def test_prism_image_matches_oracle_with_the_same_streams(cuda, oracle):
    """Paired comparison: the GPU pass and the oracle pass draw the same streams, so their images agree
    pixel by pixel up to float rounding — rare sun-caustic paths included."""
    sc = S.dispersion_prism_scene(64, 64)
    st = S.settings(seed=11, **PRISM)
    r = render.Renderer(sc, st)
    r.run_pass(64)
    gpu = r.resolve()
    ora, _ = World(oracle, sc).render(st, 64, layout=1)
    d = np.abs(gpu[..., 1] - ora[..., 1]) / (np.abs(ora[..., 1]) + 1e-3)
    print(f"same streams: mean Y gpu {gpu[..., 1].mean():.6f} oracle {ora[..., 1].mean():.6f}, "
          f"max rel pixel diff {d.max():.2e}, pixels over 1e-3: {(d > 1e-3).sum()}", flush=True)
    assert (d > 1e-3).mean() < 1e-3
    assert gpu[..., 1].mean() == pytest.approx(ora[..., 1].mean(), rel=1e-3)


# (seed, pass, pixel, sample, Y) of sun-caustic paths of the 64×64, 256-spp prism scene found by the
# oracle in 32 M paths: the light leaves the floor, refracts twice through the prism, and meets the
# sun's disc with no next-event estimation possible (a specular chain).
CAUSTIC_PATHS = [(2001, 0, 2455, 254, 297748.4), (2004, 0, 2009, 138, 303091.7), (2008, 1, 2388, 231, 55559.3)]


def test_gpu_reproduces_sun_caustic_paths(cuda, oracle):
    sc = S.dispersion_prism_scene(64, 64)
    w = World(oracle, sc)
    for seed, p, pixel, sample, y in CAUSTIC_PATHS:
        st = S.settings(seed=seed, **PRISM)
        ora, _ = w.camera_paths(st, 256, [pixel], [sample], pass_index=p)
        gpu = render.Renderer(sc, st).camera_paths(256, [pixel], [sample], pass_index=p)
        print(f"caustic path seed {seed} pixel {pixel} sample {sample}: oracle Y {ora[0, 1]:.1f}, gpu {gpu[0, 1]:.1f}",
              flush=True)
        assert ora[0, 1] == pytest.approx(y, rel=1e-6)
        assert gpu[0, 1] == pytest.approx(ora[0, 1], rel=1e-3)


def test_sky_lit_prism_consistent_with_oracle_independent_seeds(cuda, oracle):
    """Block-z agreement with independent seeds (compare.cpp:64-119) on the prism lit by a uniform
    sky: the same geometry, glass and floor, with an estimator of bounded variance."""
    sc = S.dispersion_prism_scene(64, 64, sun=False, sky=1.0)
    w = World(oracle, sc)

    def gpu(seed):
        r = render.Renderer(sc, S.settings(seed=seed, **PRISM))
        r.run_pass(64)
        return r.resolve()

    def ora(seed):
        return w.render(S.settings(seed=seed, **PRISM), 64, layout=0)[0]

    rep = compare.compare_renderers(ora, gpu, runs=4, B=8)
    print("block-z:", rep.text(), flush=True)
    assert rep.consistent(), rep.text()


def test_sky_lit_prism_relative_rmse_against_high_sample_fp64(cuda, oracle):
    sc = S.dispersion_prism_scene(64, 64, sun=False, sky=1.0)
    ref, _ = World(oracle, sc).render(S.settings(seed=99, **PRISM), 2048, layout=1)
    r = render.Renderer(sc, S.settings(seed=5, **PRISM))
    r.run_pass(256)
    img = r.resolve()
    same, _ = World(oracle, sc).render(S.settings(seed=5, **PRISM), 256, layout=1)
    e_gpu, e_cpu = compare.rel_rmse(img, ref), compare.rel_rmse(same, ref)
    print(f"rel RMSE vs 2048-spp fp64: gpu {e_gpu:.4f}, fp64 with the same streams {e_cpu:.4f}", flush=True)
    assert e_gpu == pytest.approx(e_cpu, rel=0.02)


def test_render_is_deterministic_and_resets(cuda):
    sc = S.dispersion_prism_scene(48, 48)
    a = render.Renderer(sc, S.settings(seed=3, **PRISM))
    a.run_pass(8)
    img1 = a.resolve()
    b = render.Renderer(sc, S.settings(seed=3, **PRISM))
    b.run_pass(8)
    assert np.array_equal(img1, b.resolve())
    a.reset()
    assert a.progress() == (0, 0)
    a.run_pass(8)
    assert np.array_equal(img1, a.resolve())
    a.run_pass(8)  # a second pass draws new streams (pass index 1)
    assert not np.array_equal(img1, a.resolve())
    assert a.progress() == (2, 16)


def test_unsupported_integrator_is_refused(cuda):
    from owe_hip import _lib

    with pytest.raises(_lib.OweHipError) as e:
        render.Renderer(S.white_furnace_scene(0.5), S.settings(integrator=7))
    assert e.value.code == _lib.ERR_UNSUPPORTED
