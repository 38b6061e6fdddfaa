"""The fp64 oracle's core formulas against the reference's own outputs and against closed forms.

Golden values in tests/golden/formulas.json come from the reference engine's sources
(tests/golden/make_golden.py); the closed forms follow the reference's own unit tests
(tests/test_optics.cpp:10-284) and textbook identities, chosen so a dropped term, a wrong sign or a
swapped operand fails at least one check.
"""
from __future__ import annotations

import ctypes
import json
import math

import numpy as np
import pytest

from conftest import GOLDEN

G = json.loads((GOLDEN / "formulas.json").read_text())


def vp(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(ctypes.c_void_p)


# ---------------------------------------------------------------- PCG32 and hashing


def test_pcg32_matches_oneill_published_demo(oracle):
    # pcg32-demo (pcg-c-basic): seed 42, stream 54 prints these six words first.
    assert [hex(x) for x in oracle.pcg(42, 54, 6)] == [
        "0xa15c02b7", "0x7b47f409", "0xba1d3330", "0x83d2f293", "0xbfa4784b", "0xcbed606e"]


def test_pcg32_and_uniform_match_reference_streams(oracle):
    for case in G["pcg"]:
        assert list(oracle.pcg(case["seed"], case["stream"], 16)) == case["u32"]
        assert list(oracle.uniforms(case["seed"], case["stream"], 8)) == case["uniform"]


def test_uniform_is_53_bit_composition_of_two_words(oracle):
    w = oracle.pcg(9, 4, 20).astype(np.uint64)
    u = oracle.uniforms(9, 4, 10)
    expect = ((w[0::2] >> np.uint64(5)).astype(np.float64) * 67108864.0
              + (w[1::2] >> np.uint64(6)).astype(np.float64)) / 9007199254740992.0
    assert np.array_equal(u, expect)
    assert (u >= 0).all() and (u < 1).all()


def test_mix64_and_hash_pair_match_reference(oracle):
    for k, v in G["mix64"]:
        assert oracle.mix64(int(k)) == int(v)
    for a, b, v in G["hash_pair"]:
        assert oracle.hash_pair(int(a), int(b)) == int(v)


# ---------------------------------------------------------------- Fresnel, Snell, conductors


def test_fresnel_dielectric_matches_reference(oracle):
    for c, e, v in G["fresnel_dielectric"]:
        assert oracle.fresnel_dielectric(c, e) == pytest.approx(v, rel=1e-15, abs=1e-300)


def test_fresnel_closed_forms(oracle):
    # Normal incidence: ((n-1)/(n+1))² = 0.04 at n = 1.5 (test_optics.cpp:41-92).
    assert oracle.fresnel_dielectric(1.0, 1.5) == pytest.approx(0.04, rel=1e-12)
    # Symmetric from inside: same value at normal incidence.
    assert oracle.fresnel_dielectric(-1.0, 1.5) == pytest.approx(0.04, rel=1e-12)
    # Brewster's angle: Rp = 0, so the unpolarised R = Rs / 2 with Rs from the s formula.
    n = 1.5
    tb = math.atan(n)
    ci, ct = math.cos(tb), math.cos(math.asin(math.sin(tb) / n))
    rs = ((ci - n * ct) / (ci + n * ct)) ** 2
    assert oracle.fresnel_dielectric(ci, n) == pytest.approx(rs / 2, rel=1e-12)
    # Total internal reflection beyond the critical angle from inside glass.
    crit = math.asin(1 / n)
    assert oracle.fresnel_dielectric(-math.cos(crit + 0.01), n) == 1.0
    assert oracle.fresnel_dielectric(-math.cos(crit - 0.01), n) < 1.0
    # Index matched: no reflection at any angle.
    for c in (1.0, 0.3, -0.7):
        assert oracle.fresnel_dielectric(c, 1.0) < 1e-30


def test_fresnel_conductor_matches_reference_and_normal_incidence(oracle):
    for c, n, k, v in G["fresnel_conductor"]:
        assert oracle.fresnel_conductor(c, n, k) == pytest.approx(v, rel=1e-14)
    for n, k in ((0.96, 6.69), (0.06, 3.59), (1.5, 0.0), (2.0, 1.0)):
        expect = ((n - 1) ** 2 + k ** 2) / ((n + 1) ** 2 + k ** 2)
        assert oracle.fresnel_conductor(1.0, n, k) == pytest.approx(expect, rel=1e-12)
    # k = 0 reduces to the dielectric at every angle; grazing incidence reflects everything.
    for c in (0.9, 0.5, 0.1):
        assert oracle.fresnel_conductor(c, 1.5, 0.0) == pytest.approx(oracle.fresnel_dielectric(c, 1.5), rel=1e-12)
    assert oracle.fresnel_conductor(0.0, 0.96, 6.69) == pytest.approx(1.0, rel=1e-12)


# ---------------------------------------------------------------- spectra


def test_spectral_formulas_match_reference(oracle):
    for u, v in G["sample_visible"]:
        assert oracle.sample_visible(u) == pytest.approx(v, rel=1e-15)
    for lam, v in G["visible_pdf"]:
        assert oracle.visible_pdf(lam) == pytest.approx(v, rel=1e-15, abs=0)
    for lam, x, y, z in G["cie"]:
        assert np.allclose(oracle.cie_xyz(lam), [x, y, z], rtol=1e-15, atol=0)
    assert oracle.cie_y_integral() == pytest.approx(G["cie_y_integral"], rel=1e-15)
    for lam, T, v in G["planck"]:
        assert oracle.planck(lam, T) == pytest.approx(v, rel=1e-15)


def test_cie_y_integral_and_wavelength_pdf(oracle):
    # ∫ȳ ≈ 106.86 ± 0.5 for the multi-lobe fit (test_optics.cpp:122).
    assert abs(oracle.cie_y_integral() - 106.86) < 0.5
    # The sampling pdf integrates to 1 over [360, 830] and inverts the sampler: F(λ(u)) = u.
    lam = np.linspace(360.0, 830.0, 200001)
    pdf = np.array([oracle.visible_pdf(l) for l in lam[::50]])
    integral = np.trapezoid(pdf, lam[::50])
    assert integral == pytest.approx(1.0, abs=2e-4)
    for u in (0.1, 0.3, 0.5, 0.7, 0.9):
        l_u = oracle.sample_visible(u)
        grid = np.linspace(360.0, l_u, 20001)
        cdf = np.trapezoid([oracle.visible_pdf(x) for x in grid], grid)
        assert cdf == pytest.approx(u, abs=3e-4)


def test_planck_closed_form(oracle):
    h, c, k = 6.62607015e-34, 299792458.0, 1.380649e-23
    for lam in (400.0, 550.0, 700.0):
        for T in (3000.0, 5778.0):
            l = lam * 1e-9
            expect = 2 * h * c * c / l ** 5 / (math.exp(h * c / (l * k * T)) - 1)
            assert oracle.planck(lam, T) == pytest.approx(expect, rel=2e-6)


def test_ciddor_air_index(oracle):
    # n − 1 = 2.772e-4 ± 2e-6 at the d line (test_optics.cpp:112).
    assert oracle.ciddor_air(587.5618) - 1 == pytest.approx(2.772e-4, abs=2e-6)


# ---------------------------------------------------------------- phase function and GGX


def test_hg_matches_reference(oracle):
    for c, g, v in G["hg_phase"]:
        assert oracle.hg_phase(c, g) == pytest.approx(v, rel=1e-15)
    for case in G["sample_hg"]:
        wo, wop = vp(case["wo"])
        wi = np.empty(3)
        p = oracle.sample_hg(wop, case["g"], case["u1"], case["u2"], wi.ctypes.data_as(ctypes.c_void_p))
        assert np.allclose(wi, case["wi"], rtol=0, atol=1e-15)
        assert p == pytest.approx(case["pdf"], rel=1e-13)


@pytest.mark.parametrize("g", [-0.6, 0.0, 0.45, 0.85])
def test_hg_normalisation_and_mean_cosine(oracle, g):
    # ∫ p dω = 1 and ∫ cosθ p dω = −g: cosθ is taken between wo and wi, which both point away from
    # the scattering point, so forward scattering (g > 0) favours wi ≈ −wo.
    mu = np.linspace(-1, 1, 400001)
    p = np.array([oracle.hg_phase(m, g) for m in mu[::20]])
    m = mu[::20]
    assert 2 * math.pi * np.trapezoid(p, m) == pytest.approx(1.0, abs=2e-4)
    assert 2 * math.pi * np.trapezoid(p * m, m) == pytest.approx(-g, abs=5e-4)


def test_hg_sampling_density_matches_phase(oracle):
    # Sampled cosines follow the phase: histogram of dot(wo, wi) against 2π p(μ).
    g = 0.5
    rng = np.random.default_rng(3)
    wo, wop = vp([0.0, 0.0, 1.0])
    wi = np.empty(3)
    wip = wi.ctypes.data_as(ctypes.c_void_p)
    mus = []
    for u1, u2 in rng.uniform(size=(40000, 2)):
        pdf = oracle.sample_hg(wop, g, u1, u2, wip)
        mus.append(wi[2])
        assert pdf == pytest.approx(oracle.hg_phase(wi[2], g), rel=1e-9)
    hist, edges = np.histogram(mus, bins=20, range=(-1, 1), density=True)
    centers = (edges[:-1] + edges[1:]) / 2
    expect = np.array([2 * math.pi * oracle.hg_phase(c, g) for c in centers])
    assert np.allclose(hist, expect, rtol=0.08, atol=0.02)


def test_ggx_matches_reference(oracle):
    for case in G["ggx"]:
        w, wp = vp(case["w"])
        m = np.empty(3)
        oracle.ggx_sample_wm(case["alpha"], wp, case["u1"], case["u2"], m.ctypes.data_as(ctypes.c_void_p))
        assert np.allclose(m, case["wm"], rtol=0, atol=1e-15)
        _, mp = vp(m)
        assert oracle.ggx_d(case["alpha"], mp) == pytest.approx(case["D"], rel=1e-13)
        assert oracle.ggx_pdf(case["alpha"], wp, mp) == pytest.approx(case["pdf"], rel=1e-13)


@pytest.mark.parametrize("alpha", [0.1, 0.3, 0.7])
def test_ggx_projected_area_normalisation(oracle, alpha):
    # ∫ D(m) cos θm dω = 1 (test_optics.cpp:148-170).
    th = np.linspace(0, math.pi / 2, 200001)[:-1]
    vals = np.array([oracle.ggx_d(alpha, vp([math.sin(t), 0.0, math.cos(t)])[1]) for t in th[::10]])
    integral = 2 * math.pi * np.trapezoid(vals * np.cos(th[::10]) * np.sin(th[::10]), th[::10])
    assert integral == pytest.approx(1.0, abs=2e-3)


def test_rough_lobes_match_reference(oracle):
    for c in G["rough_dielectric"]:
        wo, wop = vp(c["wo"])
        out = np.empty(6)
        ok = oracle.sample_rough_dielectric(c["alpha"], wop, c["eta"], c["uc"], c["u1"], c["u2"], 1,
                                            out.ctypes.data_as(ctypes.c_void_p))
        assert ok == c["ok"]
        if ok:
            assert np.allclose(out, c["sample"], rtol=1e-12, atol=1e-14)
            wi, wip = vp(out[:3])
            pdf = np.zeros(1)
            f = oracle.eval_rough_dielectric(c["alpha"], wop, wip, c["eta"], 1, pdf.ctypes.data_as(ctypes.c_void_p))
            assert f == pytest.approx(c["f"], rel=1e-10)
            assert pdf[0] == pytest.approx(c["pdf"], rel=1e-10)
            # Sample and evaluation agree on the density (test_optics.cpp:197-230).
            assert pdf[0] == pytest.approx(out[4], rel=1e-6)
    for c in G["rough_conductor"]:
        w, wp = vp(c["w"])
        out = np.empty(5)
        ok = oracle.sample_rough_conductor(c["alpha"], wp, 0.96, 6.69, 0.9, c["u1"], c["u2"],
                                           out.ctypes.data_as(ctypes.c_void_p))
        assert ok == c["ok"]
        if ok:
            assert np.allclose(out, c["sample"], rtol=1e-12, atol=1e-14)
            wi, wip = vp(out[:3])
            pdf = np.zeros(1)
            f = oracle.eval_rough_conductor(c["alpha"], wp, wip, 0.96, 6.69, 0.9, pdf.ctypes.data_as(ctypes.c_void_p))
            assert f == pytest.approx(c["f"], rel=1e-10)
            assert pdf[0] == pytest.approx(out[4], rel=1e-6)


def test_rough_dielectric_reciprocity(oracle):
    # Reflection lobe is symmetric in (wo, wi) (test_optics.cpp:231-252).
    rng = np.random.default_rng(5)
    for _ in range(50):
        a = rng.uniform(0.1, 0.5)
        wo = rng.standard_normal(3); wo[2] = abs(wo[2]) + 0.1; wo /= np.linalg.norm(wo)
        wi = rng.standard_normal(3); wi[2] = abs(wi[2]) + 0.1; wi /= np.linalg.norm(wi)
        _, pa = vp(wo)
        _, pb = vp(wi)
        p = np.zeros(1)
        pp = p.ctypes.data_as(ctypes.c_void_p)
        f1 = oracle.eval_rough_dielectric(a, pa, pb, 1.5, 0, pp)
        f2 = oracle.eval_rough_dielectric(a, pb, pa, 1.5, 0, pp)
        assert f1 == pytest.approx(f2, rel=1e-10)
