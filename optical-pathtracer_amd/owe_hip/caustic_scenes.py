"""The lens-caustic scenes of the reference's SPPM case (tests/test_references.cpp:183-250), as scene
text for the loader: used by the SPPM tests and by bench.py's SPPM workload.

The caustic scene: a sun overhead, an f = 100 mm N-BK7 lens 100 mm above a diffuse floor, an eye
335 mm from the caustic looking down at it — optionally through a 5 mm N-BK7 plate across the line
of sight, or into a flat mirror. The caustic seen through the plate is the direct view attenuated by
exactly what the plate passes, (1 − R)²/(1 − R²) for incoherent reflections at normal incidence;
the mirror image is the direct view from the mirror image of the eye, times the mirror's 0.9. Neither
can be rendered by path tracing (a floor point lit through the lens by a 0.5° sun) nor by particles
connected to the eye (the eye looks through glass or a mirror).
"""
from __future__ import annotations

import numpy as np

from . import loader as LD

EYE = np.array([268.3, 0.0, 150.0])
SPOT = np.zeros(3)
MIRROR = ("material coating { type = mirror  reflectance = 0.9 }\n"
          "body Mirror { type = flat_mirror  size = (80, 80)  position = (250, 0, 125)  axis = (-1, 0, 0)  "
          "up = (0, 0, 1)  material = coating }")
VIA_MIRROR_EYE, MIRRORED_SPOT, VIRTUAL_EYE = (200.0, 0.0, 150.0), (500.0, 0.0, 0.0), (300.0, 0.0, 150.0)


def _v(p) -> str:
    return "(" + ", ".join(f"{float(x):.9g}" for x in p) + ")"


def caustic_text(extra: str = "", eye=EYE, target=SPOT, resolution: int = 48) -> str:
    return f"""
units = mm
world {{ sun {{ elevation = 90deg  azimuth = 0deg  radiance = 1 }} }}
material floor {{ type = diffuse  reflectance = 0.5 }}
body Floor {{ type = sheet  size = (300, 300)  position = (0, 0, 0)  material = floor }}
body Lens {{ type = lens  medium = N-BK7  focal = 100  form = bi  diameter = 40  thickness = 8  rim = black
            position = (0, 0, 100)  axis = (0, 0, 1) }}
{extra}
observer Eye {{ position = {_v(eye)}  look_at = {_v(target)}  up = (0, 0, 1)  fov = 7deg  pupil = 2
               focus = 335  resolution = ({resolution}, {resolution}) }}
render {{ detector = Eye  integrator = sppm }}
"""


def plate_text() -> str:
    d = (SPOT - EYE) / np.linalg.norm(SPOT - EYE)
    p = EYE + d * 120
    return f"body Plate {{ type = box  size = (60, 60, 5)  position = {_v(p)}  axis = {_v(d)}  medium = N-BK7 }}"


def caustic_cases(resolution: int = 48) -> dict:
    """name → loaded scene: the direct view, through the plate, in the mirror, and from the virtual eye."""
    texts = dict(direct=caustic_text(resolution=resolution), plate=caustic_text(plate_text(), resolution=resolution),
                 mirror=caustic_text(MIRROR, VIA_MIRROR_EYE, MIRRORED_SPOT, resolution),
                 virtual=caustic_text("", VIRTUAL_EYE, SPOT, resolution))
    return {k: LD.load_scene_text(t) for k, t in texts.items()}


def plate_transmission() -> float:
    """(1 − R)²/(1 − R²) at normal incidence, weighted by ȳ over 360–830 nm, with N-BK7's Sellmeier
    index relative to air."""
    m = LD.catalog_medium("N-BK7")
    lam = np.linspace(360.0, 830.0, 941)
    um2 = (lam * 1e-3) ** 2
    n = np.sqrt(1.0 + sum(b * um2 / (um2 - c) for b, c in zip(m.a, m.b)))
    R = ((n - 1) / (n + 1)) ** 2
    T = (1 - R) ** 2 / (1 - R * R)
    y = np.array([LD.cie_xyz(float(x))[1] for x in lam])
    return float((T * y).sum() / y.sum())
