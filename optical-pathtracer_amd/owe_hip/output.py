"""Output and measurement formats, so that this backend's results can be diffed against the reference's
captures.

* ``write_pfm`` / ``read_pfm``: the raw image as the reference writes it (src/owe/render/output.cpp:18-70)
  — a colour PFM ("PF", width, height, scale −1 for little-endian), rows stored bottom to top, each
  pixel the linear sRGB (D65) of the XYZ estimate in float32; reading inverts that matrix exactly and
  honours a big-endian file (positive scale).
* ``read_pfm_raw``: the benchmark runner's reading of a PFM (tools/run_benchmarks.py:39-48) — raw
  linear values times |scale|, in file row order;
* ``benchmark_errors`` / ``glint_mask``: the benchmark metric as docs/BENCHMARKS.md:94-101 defines
  it — full-image relMSE, mean ratio, and the relMSE outside the sun-glint mask.
* ``render_record`` / ``write_render_files`` / ``read_display_settings``: the JSON render record
  (src/owe/render/record.cpp:23-141), with the same keys, and the reader of its display block
  (record.cpp:118-139).
* ``fnv1a64`` / ``image_digest``: the FNV-1a 64-bit hash of the scene text (record.cpp:15-19) and the
  ``owe bench`` digest of the raw XYZ accumulation, its doubles hashed as bytes (apps/owe.cpp:588-650).
"""
from __future__ import annotations

import json
import math
import os

import numpy as np

from .scene import DETECTOR_OBSERVER, INDEX_CAUCHY, INDEX_CIDDOR, INDEX_CONSTANT, INDEX_SELLMEIER, Medium, Scene

# CIE XYZ → linear sRGB (D65), the IEC 61966-2-1 matrix (src/owe/core/spectrum.hpp:21-22).
XYZ_TO_LINEAR_SRGB = np.array([[3.2404542, -1.5371385, -0.4985314],
                               [-0.9692660, 1.8760108, 0.0415560],
                               [0.0556434, -0.2040259, 1.0572252]])
LINEAR_SRGB_TO_XYZ = np.linalg.inv(XYZ_TO_LINEAR_SRGB)

LAMBDA_D = 587.5618  # helium d line, nm (src/owe/core/medium.hpp:62)
ENGINE_VERSION = "0.1.0"


# ---------------------------------------------------------------- PFM


def write_pfm(path: str, xyz: np.ndarray) -> None:
    """Writes an XYZ image [H, W, 3] as the reference's raw PFM (output.cpp:18-31)."""
    xyz = np.asarray(xyz, dtype=np.float64)
    if xyz.ndim != 3 or xyz.shape[2] != 3:
        raise ValueError(f"expected an XYZ image [H, W, 3], got shape {xyz.shape}")
    h, w = xyz.shape[:2]
    rgb = (xyz @ XYZ_TO_LINEAR_SRGB.T).astype("<f4")
    with open(path, "wb") as f:
        f.write(f"PF\n{w} {h}\n-1.0\n".encode())
        f.write(rgb[::-1].tobytes())


def _parse_pfm(path: str):
    with open(path, "rb") as f:
        data = f.read()
    tokens, pos = [], 0
    while len(tokens) < 4:  # magic, width, height, scale; each ends at whitespace
        while pos < len(data) and data[pos:pos + 1].isspace():
            pos += 1
        end = pos
        while end < len(data) and not data[end:end + 1].isspace():
            end += 1
        if end == pos:
            raise ValueError(f"{path}: not a colour PFM")
        tokens.append(data[pos:end])
        pos = end
    pos += 1  # the single whitespace before the data
    try:
        magic, w, h, scale = tokens[0], int(tokens[1]), int(tokens[2]), float(tokens[3])
    except ValueError:
        raise ValueError(f"{path}: not a colour PFM") from None
    if magic != b"PF" or w <= 0 or h <= 0 or scale == 0:
        raise ValueError(f"{path}: not a colour PFM")
    n = w * h * 3
    if len(data) - pos < 4 * n:
        raise ValueError(f"{path}: truncated")
    vals = np.frombuffer(data, dtype="<f4" if scale < 0 else ">f4", count=n, offset=pos)
    return vals.reshape(h, w, 3).astype(np.float64), scale


def read_pfm(path: str) -> np.ndarray:
    """The XYZ image [H, W, 3] a PFM of ``write_pfm`` holds (output.cpp:34-66): rows back in top-to-bottom
    order, the exact inverse of the XYZ → linear sRGB matrix applied. The scale's sign gives the byte
    order; its magnitude is not applied, as the reference's reader does not apply it."""
    rgb, _ = _parse_pfm(path)
    return rgb[::-1] @ LINEAR_SRGB_TO_XYZ.T


def read_pfm_raw(path: str) -> np.ndarray:
    """The benchmark runner's view (tools/run_benchmarks.py:39-48): raw values in file row order times
    |scale|. Raises ValueError for anything but a complete RGB PFM."""
    rgb, scale = _parse_pfm(path)
    return rgb * abs(scale)


GLINT_WINDOW = 9          # the median window (pixels) and the mask's dilation (its half, 4)
GLINT_OVER_MEDIAN = 20.0  # a glint exceeds its window median by this factor ...
GLINT_OVER_MEAN = 10.0    # ... and the image mean by this one
GLINT_ENERGY_CAP = 0.5    # a mask holding this share of the light or more is not a glint mask


def glint_mask(ref: np.ndarray) -> np.ndarray:
    """The benchmark metric's diagnostic glint mask (docs/BENCHMARKS.md:95-99), over the reference
    image's per-pixel channel maximum m: a pixel is a glint when m exceeds 20× the median of m over
    the 9×9 window centred on it (edges extended by their nearest value) and 10× the mean of m; the
    mask is every pixel within 4 (Chebyshev distance) of a glint. Dropped — all false — when the
    masked pixels carry half or more of the image's Σm."""
    from scipy import ndimage

    peak = np.asarray(ref, np.float64).max(axis=2)
    local = ndimage.median_filter(peak, size=GLINT_WINDOW, mode="nearest")
    glints = (peak > GLINT_OVER_MEDIAN * local) & (peak > GLINT_OVER_MEAN * peak.mean())
    near = ndimage.maximum_filter(glints.astype(np.uint8), size=GLINT_WINDOW, mode="constant", cval=0) > 0
    if peak[near].sum() >= GLINT_ENERGY_CAP * peak.sum():
        return np.zeros(peak.shape, dtype=bool)
    return near


def benchmark_errors(img: np.ndarray, ref: np.ndarray, mask: np.ndarray | None = None) -> dict:
    """The benchmark metric (docs/BENCHMARKS.md:94-101) of a raw image against a reference:

    * ``rel_mse``: the mean over every pixel and channel of (a − r)² / (r² + 0.01 · r̄²), r̄ the
      reference's mean over pixels and channels;
    * ``rel_mse_no_glints``: the same mean over the pixels outside ``mask`` (None if none are);
    * ``mean_ratio``: ā / r̄ (None for a black reference).

    Where a term's denominator is zero (r = 0 and r̄ = 0: a black reference) it counts as 0 when
    a = r too, and makes the relative error unavailable (every value None) when it does not."""
    a = np.asarray(img, dtype=np.float64)
    r = np.asarray(ref, dtype=np.float64)
    keep = np.ones(r.shape[:2], bool) if mask is None else ~np.asarray(mask, dtype=bool)
    if a.shape != r.shape or r.ndim != 3 or keep.shape != r.shape[:2]:
        raise ValueError(f"image {a.shape} and reference {r.shape} (mask {keep.shape}) dimensions differ")
    if not (np.isfinite(a).all() and np.isfinite(r).all()):
        raise ValueError("nonfinite value in the image or the reference")
    r_bar = r.mean()
    weight = r * r + 0.01 * r_bar * r_bar
    err2 = (a - r) ** 2
    dark = weight == 0
    if (dark & (err2 > 0)).any():
        return {"rel_mse": None, "rel_mse_no_glints": None, "mean_ratio": None}
    terms = np.zeros_like(err2)
    np.divide(err2, weight, out=terms, where=~dark)
    outside = terms[keep]
    return {"rel_mse": float(terms.mean()),
            "rel_mse_no_glints": float(outside.mean()) if outside.size else None,
            "mean_ratio": float(a.mean() / r_bar) if r_bar != 0 else None}


# ---------------------------------------------------------------- hashes


def fnv1a64(data: bytes | str) -> int:
    """FNV-1a, 64 bits (offset basis 14695981039346656037, prime 1099511628211)."""
    if isinstance(data, str):
        data = data.encode()
    h = 0xcbf29ce484222325
    for c in data:
        h = ((h ^ c) * 0x100000001b3) & 0xFFFFFFFFFFFFFFFF
    return h


def image_digest(xyz: np.ndarray) -> str:
    """The ``owe bench`` digest: FNV-1a 64 over the raw XYZ doubles of the image in row-major [H, W, 3]
    order, native byte order, as 16 hex digits (apps/owe.cpp:633-636). One changed bit changes it."""
    a = np.ascontiguousarray(xyz, dtype=np.float64)
    return f"{fnv1a64(a.tobytes()):016x}"


# ---------------------------------------------------------------- the JSON record


def describe_index(m: Medium) -> str:
    """IndexModel::describe (src/owe/core/medium.cpp:68-83), numbers as C++ streams print them (%g)."""
    g = lambda v: f"{v:g}"
    if m.index_kind == INDEX_CONSTANT:
        return f"constant({g(m.a[0])})"
    if m.index_kind == INDEX_CAUCHY:
        a = tuple(m.a) + (0.0,) * (3 - len(m.a))
        return f"cauchy(A={g(a[0])}, B={g(a[1])}, C={g(a[2])})"
    if m.index_kind == INDEX_SELLMEIER:
        return "sellmeier(" + ", ".join(f"B{i + 1}={g(b)} C{i + 1}={g(c)}" for i, (b, c) in
                                        enumerate(zip(m.a, m.b))) + ")"
    if m.index_kind == INDEX_CIDDOR:
        return "ciddor-air"
    return "tabulated"


def medium_index(m: Medium, lam: float) -> float:
    """n(λ) of a medium (medium.cpp:46-66): Ciddor air, Cauchy A + B/λ² + C/λ⁴ and Sellmeier in μm,
    times n_air where the catalogue glass is given relative to air."""
    l2 = (lam * 1e-3) ** 2
    air = 1 + 0.05792105 / (238.0185 - 1 / l2) + 0.00167917 / (57.362 - 1 / l2)
    if m.index_kind == INDEX_CONSTANT:
        v = m.a[0]
    elif m.index_kind == INDEX_CAUCHY:
        a = tuple(m.a) + (0.0,) * (3 - len(m.a))
        v = a[0] + a[1] / l2 + a[2] / (l2 * l2)
    elif m.index_kind == INDEX_SELLMEIER:
        v = math.sqrt(1 + sum(b * l2 / (l2 - c) for b, c in zip(m.a, m.b)))
    else:
        v = air
    return v * air if m.rel_air else v


def render_record(scene: Scene, *, backend: str, stats: dict, image: np.ndarray, seconds: float,
                  samples_per_pixel: int, passes: int, settings=None, scene_path: str = "", scene_text: str = "",
                  edits=(), detector_name: str = "Eye", integrator: str = "path", exposure_ev: float = 0.0,
                  auto_exposure: bool = False, white_balance_kelvin: float = 6504.0, tone: str = "aces",
                  media_names=None) -> dict:
    """The render record's content (record.cpp:23-110) as a dict, with the reference's keys: engine,
    backend, scene (path, FNV-1a of its text, edits), detector, observer_pose (for an observer),
    integrator, sampling, seed, samples_per_pixel, passes, max_depth, display, render_seconds,
    statistics, mean_Y and world. Display metering is not part of transport: ``auto_gain_ev`` is
    recorded as null when auto exposure is on."""
    obs = scene.observer or {}
    rec = {"engine": f"optical-world-engine {ENGINE_VERSION}", "backend": backend,
           "scene": {"path": scene_path, "fnv1a64": f"{fnv1a64(scene_text):016x}", "edits": list(edits)}}
    sigma = float(obs.get("pixel_sigma", 0.0))
    radiance = obs.get("kind", DETECTOR_OBSERVER) == DETECTOR_OBSERVER
    rec["detector"] = {"name": detector_name,
                       "description": ("ideal observer" if radiance else "surface sensor") +
                                      f" {obs.get('width', 0)}x{obs.get('height', 0)}",
                       "quantity": "radiance [W m^-2 sr^-1] as CIE XYZ" if radiance else
                                   "irradiance [W m^-2] as CIE XYZ",
                       "pixel_response": "square" + (f" * gaussian(sigma {sigma:f} px)" if sigma > 0 else "")}
    if radiance and obs:
        fd = float(obs["focus_distance"])
        rec["observer_pose"] = {"position_m": [float(v) for v in obs["position"]],
                                "look_at_m": [float(v) for v in obs["look_at"]], "up": [float(v) for v in obs["up"]],
                                "fov_deg": math.degrees(obs["fov_y"]), "pupil_radius_m": float(obs["pupil_radius"]),
                                "focus_distance_m": fd if math.isfinite(fd) else None,
                                "width": int(obs["width"]), "height": int(obs["height"])}
    rec["integrator"] = integrator
    rr = int(settings.rr_depth) if settings is not None else 0
    floor = float(settings.fresnel_floor) if settings is not None else 0.0
    lo = float(settings.lambda_min) if settings is not None else 380.0
    hi = float(settings.lambda_max) if settings is not None else 780.0
    rec["sampling"] = {"wavelengths": f"4 per path (hero + 3 stratified; secondaries terminated at dispersive "
                                      f"interfaces), visible-importance pdf on [{lo:g}, {hi:g}] nm",
                       "fresnel": "stochastic branch selection", "nee": True, "mis": "power heuristic",
                       "russian_roulette_depth": rr, "fresnel_floor": floor}
    rec["seed"] = int(settings.seed) if settings is not None else 0
    rec["samples_per_pixel"] = int(samples_per_pixel)
    rec["passes"] = int(passes)
    rec["max_depth"] = int(settings.max_depth) if settings is not None else 0
    rec["display"] = {"exposure_ev": float(exposure_ev), "auto_exposure": bool(auto_exposure),
                      "auto_gain_ev": None if auto_exposure else 0.0,
                      "white_balance_kelvin": float(white_balance_kelvin), "tone": tone}
    rec["render_seconds"] = float(seconds)
    rec["statistics"] = {"paths": int(stats["paths"]), "segments": int(stats["segments"]),
                         "region_inconsistencies": int(stats["inconsistencies"]), "leaks": int(stats["leaks"])}
    rec["mean_Y"] = float(np.asarray(image)[..., 1].mean())
    names = list(media_names) if media_names is not None else [f"medium{i}" for i in range(len(scene.media))]
    rec["world"] = {"media": [{"name": names[i], "index": describe_index(m), "n_d": medium_index(m, LAMBDA_D)}
                              for i, m in enumerate(scene.media)],
                    "regions": len(scene.region_medium), "boundaries": len(scene.boundaries)}
    return rec


def write_render_files(prefix: str, record: dict, image: np.ndarray) -> None:
    """``prefix.pfm`` and ``prefix.json`` as writeRenderFiles writes them (record.cpp:112-117); the PNG
    display image is a display function and is not written here."""
    write_pfm(prefix + ".pfm", image)
    with open(prefix + ".json", "w") as f:
        json.dump(record, f, indent=2)
        f.write("\n")


def read_display_settings(path: str) -> dict | None:
    """The display block of a record (record.cpp:118-139): exposure_ev, white_balance_kelvin,
    auto_exposure and tone, each where present; None when the file cannot be read."""
    if not os.path.exists(path):
        return None
    with open(path) as f:
        rec = json.load(f)
    d = rec.get("display", {})
    out = {}
    for k in ("exposure_ev", "white_balance_kelvin", "auto_exposure", "tone"):
        if k in d:
            out[k] = d[k]
    return out
