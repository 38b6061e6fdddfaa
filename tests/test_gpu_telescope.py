"""Mirrors, afocal spots and lens-parameter derivatives in the batched lens kernel on the GPU.

* the reference's Gregorian telescope (two conic mirrors, a holed primary, a Huygenian ocular),
  solved afocal: ray by ray against the fp64 oracle (the same rays lost in the hole and stopped by
  the primary's body; points and directions to float rounding), and its afocal spot (apparent angle
  and angular RMS) from the summed angles;
* d(spot metric)/dθ from the forward-mode dual numbers of the kernel against central differences of
  the fp64 oracle over the same rays — the derivative of the metric over a fixed ray set, which is
  what the duals compute: the Gregorian's angular RMS with respect to both mirrors' curvatures and
  conics, both of its gaps and two ocular curvatures, and the reference doublet's RMS spot radius
  at best focus with respect to its curvatures, gaps, a conic and an A4 term; each within 1 % (or
  1e-3 of the largest), the vector within 1 %;
* 64 perturbed Gregorians (the differentiable-telescope workload, reduced) run in one batch with
  their gradients, and every variant's derivatives equal those it gets alone.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

from owe_hip import lens as L
from pyoracle import LensOracle

pytestmark = pytest.mark.gpu

LAM = 587.5618


def gregorian(rings=60):
    return L.solve_afocal(L.GREGORIAN, LAM).with_rings(rings)


def one_lambda(field=0.05):
    return L.c_settings(field_deg=field, wavelengths=(LAM,), pixels=0, reference=LAM)


@pytest.mark.parametrize("field", [0.0, 0.08])
def test_gregorian_rays_and_afocal_spot_match_oracle(cuda, oracle, field):
    g = gregorian()
    b = L.LensBatch([g], one_lambda(field))
    b.run()
    per, mom, _ = b.results()
    lo = LensOracle(oracle, [g], field_deg=field, reference=LAM)
    idx = np.arange(g.rays, dtype=np.uint64)
    gt, ct = b.trace(0, 0, idx), lo.trace(0, LAM, idx)
    okg, okc = gt[:, 6] > 0, ct[:, 6] > 0
    both = okg & okc
    flips = int((okg != okc).sum())
    dp = np.abs(gt[both, :3] - ct[both, :3]).max()
    dd = np.abs(gt[both, 3:6] - ct[both, 3:6]).max()
    m, px = lo.moments(0, LAM)
    s = L.afocal_spot(m)
    r = per[0]
    print(f"field {field}: {both.sum()} arrive, {flips} flips, |Δp| {dp:.2e} |Δd| {dd:.2e}; apparent "
          f"{r['apparent_angle']:.7f} / {s['apparent_angle']:.7f}, angular RMS {r['angular_rms']:.5e} / "
          f"{s['angular_rms']:.5e}, M {r['angular_magnification']:.6f}", flush=True)
    assert r["afocal"] == 1 and r["angular_magnification"] == pytest.approx(px["angular_magnification"], rel=1e-12)
    # Float rounding over 1.3 m of folded path: points to 1e-7 m; directions to a few 1e-6, which
    # the 54-fold magnification makes a few 1e-3 of the angular RMS.
    assert flips <= 2 and dp < 1e-7 and dd < 1e-5
    assert r["arrived"] == s["arrived"]
    assert r["apparent_angle"] == pytest.approx(s["apparent_angle"], abs=1e-6)
    assert r["angular_rms"] == pytest.approx(s["angular_rms"], rel=5e-3)
    b.close()


def metric_fd(lo, lens, params, k, h, lams, afocal, idx):
    """Central difference of the spot metric over the rays arriving at θ − h, θ and θ + h."""
    np_ = len(params)

    def traces(delta):
        return [lo.trace_perturbed(lens, lam, params, delta, idx) for lam in lams]

    zero = np.zeros(np_)
    e = np.zeros(np_)
    e[k] = h
    T0, Tp, Tm = traces(zero), traces(e), traces(-e)
    keep = [(a[:, 6] > 0) & (p[:, 6] > 0) & (m[:, 6] > 0) for a, p, m in zip(T0, Tp, Tm)]

    def metric(T):
        rows = np.concatenate([t[kk] for t, kk in zip(T, keep)])
        P, D = rows[:, :3], rows[:, 3:6]
        B = D[:, :2] / D[:, 2:3]
        if afocal:
            A = np.arctan(B)
            c = np.arctan(B.mean(0))
            return math.sqrt(((A - c) ** 2).sum(1).mean())
        A = P[:, :2] - B * P[:, 2:3]
        dA, dB = A - A.mean(0), B - B.mean(0)
        va, vb, cv = (dA ** 2).sum(1).mean(), (dB ** 2).sum(1).mean(), (dA * dB).sum(1).mean()
        return math.sqrt(va - cv * cv / vb)

    return (metric(Tp) - metric(Tm)) / (2 * h)


CASES = {
    "gregorian": (gregorian, [(0, L.PARAM_CURVATURE), (0, L.PARAM_CONIC), (1, L.PARAM_CURVATURE),
                              (1, L.PARAM_CONIC), (0, L.PARAM_THICKNESS), (1, L.PARAM_THICKNESS),
                              (2, L.PARAM_CURVATURE), (4, L.PARAM_CURVATURE)],
                  [1e-6, 1e-6, 1e-6, 1e-6, 1e-8, 1e-8, 1e-5, 1e-5], 0.05, True),
    "doublet": (lambda: L.DOUBLET.with_rings(60),
                [(0, L.PARAM_CURVATURE), (1, L.PARAM_CURVATURE), (2, L.PARAM_CURVATURE), (0, L.PARAM_THICKNESS),
                 (1, L.PARAM_THICKNESS), (0, L.PARAM_CONIC), (2, L.PARAM_A4), (2, L.PARAM_CONIC)],
                [1e-5, 1e-5, 1e-6, 1e-8, 1e-8, 1e-5, 1e-3, 1e-5], 0.5, False),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_dual_number_gradients_match_fp64_differences(cuda, oracle, case):
    make, params, steps, field, afocal = CASES[case]
    p = make()
    lams = (LAM,) if afocal else (486.13, 587.56, 656.27)
    st = L.c_settings(field_deg=field, wavelengths=lams, pixels=0, reference=LAM if afocal else L.LAMBDA_D)
    b = L.LensBatch([p], st)
    b.set_params(params)
    b.run()
    g = b.gradients()[0]
    lo = LensOracle(oracle, [p], field_deg=field, reference=LAM if afocal else L.LAMBDA_D)
    idx = np.arange(p.rays, dtype=np.uint64)
    fd = np.array([metric_fd(lo, 0, params, k, steps[k], lams, afocal, idx) for k in range(len(params))])
    rel = np.abs(g - fd) / np.maximum(np.abs(fd), 1e-3 * np.abs(fd).max())
    print(f"{case}: gpu {np.array2string(g, precision=5)}\n fd {np.array2string(fd, precision=5)}\n rel "
          f"{np.array2string(rel, precision=2)}", flush=True)
    # Each within 1 %, or within 1e-3 of the largest derivative: the float sums of a corrected
    # system's small derivatives cancel to that level.
    assert (np.abs(g - fd) <= 0.01 * np.abs(fd) + 1e-3 * np.abs(fd).max()).all()
    assert np.linalg.norm(g - fd) <= 0.01 * np.linalg.norm(fd)
    b.close()


def test_batch_of_perturbed_telescopes_keeps_each_variants_gradient(cuda):
    base = gregorian(rings=40)
    rng = np.random.default_rng(5)
    variants = []
    for _ in range(64):
        sf = [L.Surface(**{**s.__dict__}) for s in base.surfaces]
        sf[0].R *= 1 + rng.uniform(-2e-3, 2e-3)
        sf[1].R *= 1 + rng.uniform(-2e-3, 2e-3)
        sf[1].k += rng.uniform(-2e-3, 2e-3)
        sf[2].R *= 1 + rng.uniform(-5e-3, 5e-3)
        # Each variant's ocular gap re-solved for an afocal system, as the scene loader does for an
        # instrument marked afocal (loader_bodies.cpp:531).
        variants.append(L.solve_afocal(L.Prescription(sf, base.ambient, base.rays, base.name, base.rings), LAM))
    params = CASES["gregorian"][1]
    b = L.LensBatch(variants, one_lambda())
    b.set_params(params)
    b.run()
    g = b.gradients()
    per, _, _ = b.results()
    for i in (0, 17, 63):
        one = L.LensBatch([variants[i]], one_lambda())
        one.set_params(params)
        one.run()
        gi = one.gradients()[0]
        assert np.allclose(g[i], gi, rtol=1e-6, atol=1e-9 * np.abs(gi).max())
        assert one.results()[0][0]["angular_rms"] == pytest.approx(per[i]["angular_rms"], rel=1e-12)
        one.close()
    assert np.isfinite(g).all()
    b.close()


def test_telescope_in_a_room_matches_oracle_path_by_path(cuda, oracle):
    # The instrument as matter (mirrors with bore and spider, ocular lenses with rims) in a room with
    # a 2,880-triangle statue; camera paths from the eye point through the ocular.
    from owe_hip import bodies as B
    from owe_hip import render
    from owe_hip import scene as S
    from pyoracle import World

    sc = B.telescope_room_scene(gregorian(), size=32, statue_frequency=12)
    st = S.settings(seed=3, max_depth=32, rr_depth=4)
    spp = 8
    pixel = np.repeat(np.arange(32 * 32, dtype=np.uint32), spp)
    sample = np.tile(np.arange(spp, dtype=np.uint32), 32 * 32)
    r = render.Renderer(sc, st)
    gpu = r.camera_paths(spp, pixel, sample)
    ora, _ = World(oracle, sc).camera_paths(st, spp, pixel, sample)
    gy, oy = gpu[:, 1], ora[:, 1]
    close = np.abs(gy - oy) <= 1e-3 * np.maximum(np.abs(oy), np.abs(gy)) + 1e-6
    print(f"telescope room: {close.mean():.5f} of {len(gy)} paths agree; mean Y gpu {gy.mean():.5f} oracle "
          f"{oy.mean():.5f}", flush=True)
    assert close.mean() > 0.99
    assert abs(gy.mean() - oy.mean()) < 0.03 * abs(oy.mean())
    r.run_pass(64)
    assert r.stats()["inconsistencies"] <= 1e-4 * r.stats()["paths"]
