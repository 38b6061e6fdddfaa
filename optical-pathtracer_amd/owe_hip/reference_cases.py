"""The reference's benchmark scenes (package files scenes/<name>.owe, restated from the reference's
benchmarks/<name>.owe), with the closed-form expectations each is built around.

* ``dispersion_prism``: sunlight at 11° through a 12 mm gap onto an upright 60° N-SF11 prism near
  minimum deviation, its spectrum on a white screen; each wavelength deviates as the two-face Snell
  trace predicts (tests/test_references.cpp:278-320).
* ``caustic_seen_through_glass`` / ``caustic_seen_in_mirror``: the f = 100 mm lens caustic seen
  through a 5 mm N-BK7 plate, or in a 0.9 flat mirror; the images are the direct view
  × (1 − R)²/(1 − R²) and × 0.9 (tests/test_references.cpp:212-250).
* ``double_refraction_slabs``: sunlight through a 6 mm stop and two tilted 12 mm N-BK7 slabs, which
  move the beam sideways without turning it (``beam_shift``, by Snell's law at every face the
  central ray crosses; ``slab_shift``, the textbook plate formula).
* ``telescope_exit_pupil_walk`` (the case of tests/test_navigation.cpp:220-275): the 150 mm f/15
  refractor with a 16 mm Plössl solved afocal, aimed at a distant disc; an eye of 1.2 mm pupil
  walking sideways across the exit pupil sees the image dim as the overlap of its pupil with the
  emerging beam (``disc_overlap``).

The refractor's prescription is scenes/refractor_150mm.lens (the reference's
lenses/refractor_150mm.lens). The scene-file benchmark (owe_hip.suite) runs the five with the
package's rippled pool.
"""
from __future__ import annotations

import math
import re
from pathlib import Path

import numpy as np

from . import loader as LD

SCENES = Path(__file__).parent / "scenes"
NAMES = ("dispersion_prism", "caustic_seen_through_glass", "caustic_seen_in_mirror", "double_refraction_slabs",
         "telescope_exit_pupil_walk")

# double_refraction_slabs.owe's geometry: the stop's axis is the sunlight's direction of travel (the
# sun at 30° elevation, azimuth 270°: light travels along (cos 30°, 0, −sin 30°)).
SLAB_T = 0.012
SLAB_NORMALS = ((0.574, 0.0, 0.819), (0.94, 0.0, -0.34))
SLAB_CENTRES = ((-0.05, 0.0, 0.214), (0.05, 0.0, 0.156))
STOP_CENTRE = (-0.2, 0.0, 0.30)
BEAM = (math.cos(math.radians(30)), 0.0, -math.sin(math.radians(30)))
ACROSS = (BEAM[2], 0.0, -BEAM[0])  # unit, perpendicular to the beam in the slabs' plane of incidence


def path(name: str) -> Path:
    return SCENES / f"{name}.owe"


def text(name: str) -> str:
    return path(name).read_text()


def load(name: str, edits=()) -> LD.LoadedScene:
    return LD.load_scene(str(path(name)), edits)


def load_text(source: str) -> LD.LoadedScene:
    return LD.load_scene_text(source, "<reference case>", str(SCENES))


def _v(p) -> str:
    return "(" + ", ".join(f"{float(x):.9g}" for x in p) + ")"


def _drop(source: str, head: str) -> str:
    """The scene text without the statement that starts with `head` (one brace level deep)."""
    out, n = re.subn(r"^" + re.escape(head) + r"\b[^{]*\{[^{}]*\}[ \t]*\n", "", source, flags=re.M)
    if n != 1:
        raise ValueError(f"no single statement '{head}' in the scene text")
    return out


def double_refraction_slabs(slabs: bool = True, beam_eye=None) -> str:
    """The slab scene's text; without `slabs` the same scene with both slabs taken out. With
    `beam_eye` = (offset, resolution) the wall and the eye make way for a pinhole eye on the beam's
    axis 0.12 m beyond the second slab, moved `offset` metres along ACROSS (the way the slabs move
    the beam), looking back at the sun through the slabs and the stop with a 1.2° field: it sees
    the sun only where the beam reaches it."""
    src = text("double_refraction_slabs")
    if not slabs:
        src = _drop(_drop(src, "body SlabA"), "body SlabB")
    if beam_eye is not None:
        offset, res = beam_eye
        src = _drop(_drop(_drop(src, "body Wall"), "observer Eye"), "render")
        stop, beam = np.array(STOP_CENTRE), np.array(BEAM)
        c = stop + beam * (np.dot(np.array(SLAB_CENTRES[1]) - stop, beam) + 0.12) + offset * np.array(ACROSS)
        src += (f"observer Beam {{ position = {_v(c)}  look_at = {_v(c - beam)}  up = (0, 1, 0)  fov = 1.2deg  "
                f"resolution = ({res}, {res}) }}\n"
                f"render {{ detector = Beam  spp = 16  integrator = path  max_depth = 16 }}\n")
    return src


def slab_shift(theta: float, t: float, n: float) -> float:
    """The sideways shift of a ray crossing a plane-parallel plate of thickness t and relative index
    n at incidence θ: t·sin(θ − θt)/cos θt, sin θt = sin θ / n."""
    tt = math.asin(math.sin(theta) / n)
    return t * math.sin(theta - tt) / math.cos(tt)


def _box_axes(normal) -> np.ndarray:
    """The slab's local axes as rows (a box placed with `axis` and the default up (0, 0, 1)): x,
    along world y; y, the up direction made perpendicular to the normal (across the face, in the
    x–z plane); z, the normal."""
    z = np.array(normal, float) / np.linalg.norm(normal)
    y = np.array((0.0, 0.0, 1.0)) - z[2] * z
    y /= np.linalg.norm(y)
    return np.array((np.cross(y, z), y, z))


def _refract(d, m, eta):
    """d refracted through a face of unit normal m (m·d < 0) at relative index eta; None on total
    internal reflection."""
    ci = -np.dot(d, m)
    k = 1 - eta * eta * (1 - ci * ci)
    if k < 0:
        return None
    d = eta * d + (eta * ci - math.sqrt(k)) * m
    return d / np.linalg.norm(d)


def _through_box(o, d, centre, normal, n):
    """The ray (o, d) after it crosses the slab box (0.06 × 0.06 × SLAB_T, refractive index n):
    refracted into whichever face it meets first and out of whichever face it then reaches (an edge
    face where the light comes in at grazing incidence on the big faces), straight on if it misses."""
    axes = _box_axes(normal)
    half = np.array((0.03, 0.03, SLAB_T / 2))
    lo, ld = axes @ (o - np.array(centre)), axes @ d
    with np.errstate(divide="ignore"):
        t0, t1 = (-half - lo) / ld, (half - lo) / ld
    near, far = np.minimum(t0, t1), np.maximum(t0, t1)
    if near.max() > far.min() or far.min() < 0:
        return o, d
    k = int(np.argmax(near))
    m = np.zeros(3)
    m[k] = -math.copysign(1.0, ld[k])
    p = lo + ld * near.max()
    ld = _refract(ld, m, 1.0 / n)
    with np.errstate(divide="ignore"):
        t = np.where(ld > 0, (half - p) / ld, np.where(ld < 0, (-half - p) / ld, np.inf))
    k = int(np.argmin(t))
    p = p + ld * t[k]
    m = np.zeros(3)
    m[k] = -math.copysign(1.0, ld[k])  # the exit face's normal, turned to face the ray inside
    out = _refract(ld, m, n)
    if out is None:
        raise ValueError("the central ray is totally reflected inside a slab")
    return axes.T @ p + np.array(centre), axes.T @ out


def beam_shift(n: float) -> np.ndarray:
    """The displacement of the beam's central ray after both slabs at refractive index n, by
    Snell's law at every face it crosses: a vector perpendicular to the beam (neither slab deviates
    it). At this geometry the beam meets slab A's big face at 85° and so enters and leaves through
    its 12 mm edge faces, crossing 60 mm of glass nearly head-on; slab B it crosses face to face."""
    o, d = np.array(STOP_CENTRE, float), np.array(BEAM)
    for c, nrm in zip(SLAB_CENTRES, SLAB_NORMALS):
        o, d = _through_box(o, d, c, nrm, n)
    start = np.array(STOP_CENTRE, float)
    straight = start + np.array(BEAM) * np.dot(o - start, np.array(BEAM))
    return o - straight


def beam_shift_across(n: float) -> float:
    """beam_shift's signed length along ACROSS."""
    return float(np.dot(beam_shift(n), np.array(ACROSS)))


def exit_pupil_walk(ls: LD.LoadedScene, dx: float, resolution=None):
    """The walk's scene with the eye moved dx metres sideways (along x) from the exit pupil,
    looking the same way; returns the scene."""
    d = ls.detectors[ls.detector_index()]
    base = d.setdefault("_centred", (tuple(d["position"]), tuple(d["look_at"])))
    d["position"] = tuple(np.add(base[0], (dx, 0.0, 0.0)))
    d["look_at"] = tuple(np.add(base[1], (dx, 0.0, 0.0)))
    return ls.use(resolution=resolution)


def disc_overlap(r1: float, r2: float, d: float) -> float:
    """The area common to two discs of radii r1 and r2 whose centres are d apart."""
    if d >= r1 + r2:
        return 0.0
    if d <= abs(r1 - r2):
        return math.pi * min(r1, r2) ** 2
    a = r1 * r1 * math.acos((d * d + r1 * r1 - r2 * r2) / (2 * d * r1))
    b = r2 * r2 * math.acos((d * d + r2 * r2 - r1 * r1) / (2 * d * r2))
    return a + b - 0.5 * math.sqrt((-d + r1 + r2) * (d + r1 - r2) * (d - r1 + r2) * (d + r1 + r2))
