"""Output and measurement formats: the raw PFM, the benchmark runner's relMSE, the JSON render record
and the FNV-1a digests, pinned by published vectors and the reference's own tests.

* FNV-1a 64: the published test vectors of the algorithm ("", "a", "foobar");
* PFM: byte layout (header, bottom-to-top rows, little-endian float32), D65 white maps to sRGB
  (1, 1, 1) and the inverse matrix's Y row is Rec. 709's luminance weights (spectrum.hpp:23) — a
  transposed or mistyped matrix fails both; round trip within float32 rounding;
* the runner's reader and errors: tools/test_run_benchmarks.py's cases (relMSE 1/1.01 and mean ratio
  2 for a doubled image, |scale| 2 applied in both byte orders, a dark reference, non-finite and
  mismatched inputs rejected);
* the record: the reference's display reader finds its keys by the text it searches for
  (record.cpp:126-133), media describe as IndexModel::describe prints, N-BK7's n_d is the catalogue's
  1.5168 relative to air and Ciddor air at the d line is 1.000277.
"""
from __future__ import annotations

import json
import math

import numpy as np
import pytest

from owe_hip import output as O
from owe_hip import scene as S


def test_fnv1a64_published_vectors():
    assert O.fnv1a64(b"") == 0xcbf29ce484222325
    assert O.fnv1a64(b"a") == 0xaf63dc4c8601ec8c
    assert O.fnv1a64("foobar") == 0x85944171f73967e8


def test_image_digest_sees_one_bit():
    img = np.random.default_rng(1).random((4, 5, 3))
    d0 = O.image_digest(img)
    bits = img.view(np.uint64).copy()
    bits[2, 3, 1] ^= np.uint64(1)
    assert O.image_digest(bits.view(np.float64)) != d0
    assert O.image_digest(img.copy()) == d0
    assert len(d0) == 16


def test_pfm_layout(tmp_path):
    xyz = np.zeros((2, 3, 3))
    xyz[0, :, 1] = 1.0                                   # top row: Y = 1
    xyz[1, 2] = (0.95047, 1.0, 1.08883)                  # bottom-right: D65 white
    p = str(tmp_path / "a.pfm")
    O.write_pfm(p, xyz)
    raw = open(p, "rb").read()
    head = b"PF\n3 2\n-1.0\n"
    assert raw.startswith(head) and len(raw) == len(head) + 2 * 3 * 3 * 4
    vals = np.frombuffer(raw[len(head):], "<f4").reshape(2, 3, 3)
    # The first stored row is the image's bottom row; D65 white is sRGB (1, 1, 1).
    assert np.abs(vals[0, 2] - 1.0).max() < 1e-4
    assert np.abs(vals[0, :2]).max() == 0
    assert vals[1, 0] == pytest.approx([-1.5371385, 1.8760108, -0.2040259], abs=1e-7)


def test_inverse_matrix_y_row_is_rec709_luminance():
    assert O.LINEAR_SRGB_TO_XYZ[1] == pytest.approx([0.2126, 0.7152, 0.0722], abs=1e-4)
    assert O.LINEAR_SRGB_TO_XYZ @ np.ones(3) == pytest.approx([0.95047, 1.0, 1.08883], abs=1e-4)


def test_pfm_round_trip(tmp_path):
    xyz = np.random.default_rng(2).random((7, 5, 3)) * 3
    p = str(tmp_path / "b.pfm")
    O.write_pfm(p, xyz)
    back = O.read_pfm(p)
    assert back.shape == xyz.shape
    assert np.abs(back - xyz).max() < 1e-6 * 3 * 10


def _write_runner_pfm(path, values, scale=-1):
    # The form tools/test_run_benchmarks.py:15-19 writes: any scale, either byte order.
    values = np.asarray(values, dtype="<f4" if scale < 0 else ">f4")
    with open(path, "wb") as f:
        f.write(f"PF\n{values.shape[1]} {values.shape[0]}\n{scale}\n".encode())
        f.write(values.tobytes())


def test_runner_reader_scale_and_endian(tmp_path):
    p = str(tmp_path / "c.pfm")
    for scale in (-2, 2):
        _write_runner_pfm(p, np.ones((2, 3, 3)), scale)
        np.testing.assert_array_equal(O.read_pfm_raw(p), 2 * np.ones((2, 3, 3)))
        # The XYZ reader honours the byte order and, as output.cpp:34-66 does, not the magnitude.
        assert O.read_pfm(p) @ O.XYZ_TO_LINEAR_SRGB.T == pytest.approx(np.ones((2, 3, 3)))
    with open(p, "wb") as f:
        f.write(b"PF\n2 2\n-1.0\n" + b"\0" * 20)
    with pytest.raises(ValueError, match="truncated"):
        O.read_pfm_raw(p)
    with open(p, "wb") as f:
        f.write(b"Pf\n2 2\n-1.0\n" + b"\0" * 16)
    with pytest.raises(ValueError, match="not a colour PFM"):
        O.read_pfm_raw(p)


def test_runner_errors_match_its_tests():
    ref = np.ones((2, 2, 3))
    got = O.benchmark_errors(2 * ref, ref)
    assert got["rel_mse"] == pytest.approx(1 / 1.01)
    assert got["mean_ratio"] == 2
    dark = np.zeros((1, 1, 3))
    r = O.benchmark_errors(dark, dark)
    assert r["rel_mse"] == 0 and r["mean_ratio"] is None
    assert O.benchmark_errors(np.ones_like(dark), dark)["rel_mse"] is None
    with pytest.raises(ValueError, match="nonfinite"):
        O.benchmark_errors(np.full((1, 1, 3), np.nan), np.ones((1, 1, 3)))
    with pytest.raises(ValueError, match="dimensions"):
        O.benchmark_errors(np.ones((1, 1, 3)), np.ones((2, 2, 3)))


def _glint_mask_by_loops(ref):
    """The glint rule of docs/BENCHMARKS.md:97-99, pixel by pixel (small images only)."""
    m = ref.max(axis=2)
    h, w = m.shape
    glint = np.zeros((h, w), bool)
    for y in range(h):
        for x in range(w):
            win = [m[min(max(y + dy, 0), h - 1), min(max(x + dx, 0), w - 1)] for dy in range(-4, 5) for dx in range(-4, 5)]
            glint[y, x] = m[y, x] > 20 * sorted(win)[40] and m[y, x] > 10 * m.mean()
    mask = np.zeros((h, w), bool)
    for y, x in zip(*np.nonzero(glint)):
        mask[max(y - 4, 0):y + 5, max(x - 4, 0):x + 5] = True
    return mask if m[mask].sum() < 0.5 * m.sum() else np.zeros_like(mask)


def test_glint_mask_matches_the_rule_pixel_by_pixel():
    rng = np.random.default_rng(3)
    for trial in range(6):
        img = rng.gamma(2.0, 1.0, size=(23, 31, 3))
        for _ in range(trial):
            img[rng.integers(0, 23), rng.integers(0, 31)] = rng.uniform(100, 400, 3)
        img[0, 0] = 300.0 if trial % 2 else img[0, 0]  # a glint at the corner: edge extension
        assert np.array_equal(O.glint_mask(img), _glint_mask_by_loops(img)), trial


def test_glint_mask():
    img = np.full((40, 40, 3), 1.0)
    img[10, 20] = 500.0
    m = O.glint_mask(img)
    assert m.sum() == 81 and m[6:15, 16:25].all()
    # Masked relMSE leaves the glint out.
    a = img.copy()
    a[10, 20] = 900.0
    e = O.benchmark_errors(a, img, m)
    assert e["rel_mse"] > 0 and e["rel_mse_no_glints"] == 0
    # A glint holding half the energy or more is not a glint to mask.
    img[10, 20] = 1e6
    assert O.glint_mask(img).sum() == 0


def test_index_description_and_d_line():
    bk7 = S.Medium.sellmeier(**S.N_BK7)
    assert O.describe_index(bk7) == ("sellmeier(B1=1.03961 C1=0.0060007, B2=0.231792 C2=0.0200179, "
                                     "B3=1.01047 C3=103.561)")
    air = S.Medium.air()
    assert O.medium_index(air, O.LAMBDA_D) == pytest.approx(1.000277, abs=2e-6)
    assert O.medium_index(bk7, O.LAMBDA_D) / O.medium_index(air, O.LAMBDA_D) == pytest.approx(1.5168, abs=1e-4)
    assert O.describe_index(S.Medium.cauchy(1.5046, 0.0042)) == "cauchy(A=1.5046, B=0.0042, C=0)"
    assert O.describe_index(S.Medium.constant(1.333)) == "constant(1.333)"
    assert O.medium_index(S.Medium.cauchy(1.5, 0.01), 500.0) == pytest.approx(1.5 + 0.01 / 0.25)


def test_record_round_trip_and_display_reader(tmp_path):
    sc = S.dispersion_prism_scene(8, 6)
    st = S.settings(seed=7, max_depth=16, rr_depth=4)
    img = np.random.default_rng(3).random((6, 8, 3))
    text = "units m\n"
    rec = O.render_record(sc, backend="hip (test)", stats=dict(paths=48, segments=100, inconsistencies=0, leaks=0),
                          image=img, seconds=0.5, samples_per_pixel=1, passes=1, settings=st,
                          scene_path="prism.owe", scene_text=text, exposure_ev=1.5, tone="filmic")
    prefix = str(tmp_path / "r")
    O.write_render_files(prefix, rec, img)
    body = open(prefix + ".json").read()
    back = json.loads(body)
    assert back["seed"] == 7 and back["max_depth"] == 16 and back["samples_per_pixel"] == 1
    assert back["scene"]["fnv1a64"] == f"{O.fnv1a64(text):016x}"
    assert back["mean_Y"] == pytest.approx(img[..., 1].mean())
    assert back["statistics"] == {"paths": 48, "segments": 100, "region_inconsistencies": 0, "leaks": 0}
    assert back["observer_pose"]["width"] == 8 and back["observer_pose"]["focus_distance_m"] is None
    assert len(back["world"]["media"]) == len(sc.media)
    # The reference's reader searches for `"key": ` and reads to the next , } or newline.
    for key, want in (("exposure_ev", "1.5"), ("white_balance_kelvin", "6504.0"), ("auto_exposure", "false")):
        i = body.index(f'"{key}": ') + len(key) + 4
        j = min(body.find(c, i) for c in ",}\n" if body.find(c, i) >= 0)
        assert body[i:j] == want
    assert O.read_display_settings(prefix + ".json") == {"exposure_ev": 1.5, "white_balance_kelvin": 6504.0,
                                                         "auto_exposure": False, "tone": "filmic"}
    assert O.read_display_settings(str(tmp_path / "none.json")) is None
    assert np.abs(O.read_pfm(prefix + ".pfm") - img).max() < 1e-5
    assert math.isfinite(back["render_seconds"])
