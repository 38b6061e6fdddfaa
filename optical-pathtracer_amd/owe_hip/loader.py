"""Scene ingest: `.owe` scene files as scenes for the hip backend (src/owe/loader/scene_loader.cpp,
loader_bodies.cpp, loader_materials.cpp, loader_detectors.cpp).

``load_scene(path, edits=())`` parses a file (owe_hip.owe_format), applies ``Block.key=value`` edits,
and builds an ``owe_hip.scene.Scene`` the way the reference's Loader builds its World: units; media
(the reference catalogue, ``medium`` blocks with a base, an index model, absorption, internal
transmittance, scattering and g); materials (the built-in ones and ``material`` blocks: diffuse,
mirror, conductor with the catalogue metals, dielectric, absorber, null; roughness, tint, back =
black); the world (ambient medium, up, a uniform sky, the sun); assemblies; bodies — sheet, screen,
disk (annulus), sphere, box, cylinder, prism, lens (two surfaces, a cemented list, or a focal length),
mirror, flat_mirror, stop/iris/aperture, tube, waves (wave lists and ripple fields, with or without a
basin), prescription (``.lens`` files, afocal solve, tube with mounts), mesh (OBJ) — with placement,
``repeat``, ``in`` and emission; observers and sensors; the render block.

The result, a ``LoadedScene``, holds every detector; ``use(detector, resolution)`` makes one the
scene's detector. Everything the hip backend does not render is refused with the line it is on:
textures, relief and image maps, stained glass, gradient and image skies, tabulated index models,
cameras, and the body types lathe, torus, column, dome, round_wall, starfield, vessel, cup, terrain,
forest and fractal_statue. The sampling hints of the reference that change noise only (an observer's
exit-pupil guide, a detector's sun share, an indirect guide) are recorded and not applied; shading
normals of meshes are ignored, as the reference's CPU backend ignores them.
"""
from __future__ import annotations

import math
import os
from dataclasses import dataclass, field

import numpy as np

from . import bodies as B
from . import lens as L
from . import owe_format as F
from . import scene as S
from .owe_format import SceneError, Value

LAMBDA_MIN, LAMBDA_MAX = 360.0, 830.0


# ---------------------------------------------------------------- catalogues (medium.cpp:103-175)

def _sellmeier(B_, C_, rel_air=False) -> S.Medium:
    return S.Medium.sellmeier(B_, C_, rel_air)


def _squares(v):
    return tuple(x * x for x in v)


def catalog_medium(name: str) -> S.Medium | None:
    """The reference's catalogue media (src/owe/core/medium.cpp:103-152): vacuum, Ciddor air, opaque,
    Schott glasses (Sellmeier, relative to air), fused silica, CaF2, sapphire (ordinary ray), diamond,
    PMMA, water (Daimon & Masumura with Pope & Fry absorption), and the Cauchy crown and flint."""
    if name == "vacuum":
        return S.Medium.constant(1.0)
    if name == "air":
        return S.Medium.air()
    if name == "opaque":
        return S.Medium(S.INDEX_CONSTANT, (1.0,), opaque=True)
    if name in L.CATALOG and L.CATALOG[name] is not None and L.CATALOG[name][0] != "cauchy":
        Bc, Cc = L.CATALOG[name]
        return _sellmeier(Bc, Cc, True)
    table = {
        "fused-silica": ((0.6961663, 0.4079426, 0.8974794), _squares((0.0684043, 0.1162414, 9.896161))),
        "CaF2": ((0.5675888, 0.4710914, 3.8484723), _squares((0.050263605, 0.1003909, 34.649040))),
        "sapphire": ((1.4313493, 0.65054713, 5.3414021), _squares((0.0726631, 0.1193242, 18.028251))),
        "diamond": ((0.3306, 4.3356), _squares((0.1750, 0.1060))),
        "PMMA": ((1.1819,), (0.011313,)),
    }
    if name in table:
        return _sellmeier(*table[name])
    if name == "water":
        return B.water_medium()
    if name == "crown":
        return S.Medium.cauchy(1.5046, 0.00420)
    if name == "flint":
        return S.Medium.cauchy(1.6200, 0.01020)
    return None


# Approximate room-temperature optical constants (n, k) of the catalogue metals (medium.cpp:155-175).
CONDUCTORS = {
    "aluminium": ((400, 450, 500, 550, 600, 650, 700, 750, 800),
                  (0.49, 0.62, 0.77, 0.96, 1.20, 1.47, 1.83, 2.20, 2.80),
                  (4.86, 5.47, 6.08, 6.69, 7.26, 7.79, 8.31, 8.60, 8.45)),
    "silver": ((400, 450, 500, 550, 600, 650, 700, 800), (0.05, 0.04, 0.05, 0.06, 0.06, 0.05, 0.04, 0.04),
               (2.10, 2.65, 3.09, 3.59, 4.01, 4.45, 4.84, 5.60)),
    "gold": ((400, 450, 500, 550, 600, 650, 700, 800), (1.47, 1.40, 0.97, 0.43, 0.25, 0.17, 0.16, 0.15),
             (1.95, 1.88, 1.87, 2.45, 3.00, 3.55, 3.95, 4.90)),
    "copper": ((400, 450, 500, 550, 600, 650, 700, 800), (1.18, 1.17, 1.12, 1.00, 0.27, 0.21, 0.21, 0.23),
               (2.21, 2.40, 2.60, 2.58, 3.40, 3.70, 4.20, 4.90)),
}
CONDUCTORS["aluminum"] = CONDUCTORS["aluminium"]


# ---------------------------------------------------------------- RGB spectra (spectrum.cpp:60-125)

def cie_xyz(l: float):
    """CIE 1931 2° colour matching functions, the Wyman, Sloan & Shirley (2013) multi-lobe fit."""
    def lobe(x, mu, s1, s2):
        t = (x - mu) / (s1 if x < mu else s2)
        return math.exp(-0.5 * t * t)
    return (1.056 * lobe(l, 599.8, 37.9, 31.0) + 0.362 * lobe(l, 442.0, 16.0, 26.7) - 0.065 * lobe(l, 501.1, 20.4, 26.2),
            0.821 * lobe(l, 568.8, 46.9, 40.5) + 0.286 * lobe(l, 530.9, 16.3, 31.1),
            1.217 * lobe(l, 437.0, 11.8, 36.0) + 0.681 * lobe(l, 459.0, 26.0, 13.8))


def rgb_basis(l: float):
    """The partition of unity of the RGB spectra: logistic edges of width 11 nm at 492 and 590 nm."""
    hi = 1 / (1 + math.exp(-(l - 590) / 11))
    lo = 1 / (1 + math.exp(-(l - 492) / 11))
    return hi, lo - hi, 1 - lo


_RGB_CAL = None


def rgb_weights(r: float, g: float, b: float):
    """Basis weights reproducing a linear sRGB colour, white-balanced so that (1, 1, 1) gives equal
    weights (a flat spectrum): M[i][k] is the sRGB of basis k times a 6504 K Planck spectrum."""
    global _RGB_CAL
    if _RGB_CAL is None:
        from .output import XYZ_TO_LINEAR_SRGB
        from .scene import planck
        M = np.zeros((3, 3))
        l = LAMBDA_MIN
        while l <= LAMBDA_MAX:
            w = np.array(rgb_basis(l)) * float(planck(l, 6504.0)) * 0.5
            M += np.outer(XYZ_TO_LINEAR_SRGB @ np.array(cie_xyz(l)), w)
            l += 0.5
        _RGB_CAL = (np.linalg.inv(M), M.sum(axis=1))
    inv, white = _RGB_CAL
    return inv @ (np.array([r, g, b]) * white)


def rgb_spectrum(r: float, g: float, b: float, emission: bool, scale: float = 1.0) -> S.Spectrum:
    """rgb(r, g, b): a reflectance (weights clamped to [0, 1]) or an illuminant on a 6504 K Planck base
    normalised to luminance 1 (weights clamped at 0), times scale. White is exactly a constant or a
    6504 K blackbody; other colours are tabulated every 5 nm over 360–830 nm."""
    w = rgb_weights(r, g, b)
    if abs(r - 1) + abs(g - 1) + abs(b - 1) == 0:
        return S.Spectrum.blackbody(6504.0, scale) if emission else S.Spectrum.constant(scale)
    w = np.maximum(w, 0.0) if emission else np.clip(w, 0.0, 1.0)
    lams = np.arange(LAMBDA_MIN, LAMBDA_MAX + 0.1, 5.0)
    vals = np.array([float(np.dot(w, rgb_basis(l))) for l in lams])
    if emission:
        from .scene import planck
        base = np.array([float(planck(l, 6504.0)) for l in lams])
        ref = S.Spectrum.blackbody(6504.0, 1.0)
        vals = vals * base * ref.scale
    return S.Spectrum.tabulated(lams, vals * scale)


# ---------------------------------------------------------------- the loaded scene


@dataclass
class LoadedScene:
    scene: S.Scene
    detectors: list                 # dicts: name, kind ("observer" | "sensor"), and the detector's keys
    render: dict                    # the render block's settings
    display: dict = field(default_factory=dict)     # per-detector display settings
    source_path: str = ""
    source_text: str = ""
    edits: tuple = ()
    media_names: list = field(default_factory=list)
    body_names: list = field(default_factory=list)
    notes: list = field(default_factory=list)       # what was recorded and not applied

    def detector_index(self, name: str | None = None) -> int:
        if name is None:
            name = self.render.get("detector") or ""
            if not name:
                return 0
        for i, d in enumerate(self.detectors):
            if d["name"] == name:
                return i
        raise SceneError(f"{self.source_path}: no detector named '{name}'")

    def use(self, detector: str | int | None = None, resolution=None) -> S.Scene:
        """Makes a detector (by name or index; default: the render block's, else the first) the scene's,
        at its own resolution or the one given, and returns the scene."""
        i = detector if isinstance(detector, int) else self.detector_index(detector)
        d = dict(self.detectors[i])
        w, h = resolution if resolution is not None else d["resolution"]
        if d["kind"] == "observer":
            self.scene.set_observer(w, h, d["position"], d["look_at"], d["up"], d["fov_y"], d["pupil_radius"],
                                    d["focus_distance"], d["pixel_sigma"], d["region"])
        else:
            self.scene.set_sensor(d["boundary"], w, h, d["half"][0], d["half"][1], pixel_sigma=d["pixel_sigma"],
                                  aim_center=d.get("aim_center"), aim_normal=d.get("aim_normal", (0.0, 0.0, 1.0)),
                                  aim_radius=d.get("aim_radius", 0.0), aim_share=d.get("aim_share", 1.0))
        self.active = d["name"]
        return self.scene

    def settings(self, integrator: str | None = None, seed: int | None = None, max_depth: int | None = None,
                 rr_depth: int = 6, spectral: int = S.SPECTRAL_HERO4) -> S.CRenderSettings:
        """The render block's settings for this backend. The integrator is the block's (or the one
        given): path, light, hybrid, sppm, vcm or bdpt (the reference's six, scene.hpp:19; the photon
        estimators with the block's photon_radius, photon_alpha and photons, scene_loader.cpp:110-121,
        and vcm's wavelength groups)."""
        name = integrator or self.render.get("integrator", "path")
        kinds = {"path": S.INTEGRATOR_PATH, "light": S.INTEGRATOR_LIGHT, "sppm": S.INTEGRATOR_SPPM,
                 "vcm": S.INTEGRATOR_VCM, "bdpt": S.INTEGRATOR_BDPT, "hybrid": S.INTEGRATOR_HYBRID}
        if name not in kinds:
            raise SceneError(f"{self.source_path}: integrator '{name}' is not available in the hip backend "
                             "(path, light, hybrid, sppm, vcm or bdpt)")
        radius = float(self.render.get("photon_radius", 2.0))
        alpha = float(self.render.get("photon_alpha", 2.0 / 3.0))
        photons = float(self.render.get("photons", 1.0))
        if not (radius > 0 and 0 < alpha <= 1 and photons > 0):
            raise SceneError(f"{self.source_path}: photon_radius and photons must be positive, photon_alpha in (0, 1]")
        return S.settings(seed=int(self.render.get("seed", 1) if seed is None else seed),
                          max_depth=int(self.render.get("max_depth", 64) if max_depth is None else max_depth),
                          rr_depth=rr_depth, spectral=spectral, integrator=kinds[name],
                          fresnel_floor=float(self.render.get("fresnel_floor", 0.0)),
                          photon_radius=radius, photon_alpha=alpha, photons_per_pixel=photons,
                          photon_groups=int(self.render.get("vcm_groups", 1)) if name in ("vcm", "bdpt") else None)


# ---------------------------------------------------------------- the loader


def _rotation_to(frm, to) -> np.ndarray:
    """Transform::rotateTo: the minimal rotation taking unit vector frm to to (a half turn if opposite)."""
    frm, to = np.asarray(frm, float), np.asarray(to, float)
    axis = np.cross(frm, to)
    if np.linalg.norm(axis) > 1e-12:
        return S.rotation(axis, math.acos(max(-1.0, min(1.0, float(np.dot(frm, to))))))
    if np.dot(frm, to) > 0:
        return np.eye(3)
    a = np.cross(to, [1.0, 0.0, 0.0] if abs(to[0]) < 0.9 else [0.0, 1.0, 0.0])
    return S.rotation(a / np.linalg.norm(a), math.pi)


def _apply(xf, p):
    xf = np.asarray(xf, float)
    return xf[:, :3] @ np.asarray(p, float) + xf[:, 3]


class _Loader:
    def __init__(self, origin: str, base_dir: str):
        self.origin, self.base = origin, base_dir
        self.unit = 1.0
        self.sc = S.Scene()
        self.media: dict[str, int] = {}          # medium name → index into sc.media
        self.defined_media: dict[str, S.Medium] = {}
        self.media_names: list[str] = []
        self.outside = 0                         # the region a body's outside is: ambient, or its host's
        self.body_regions: dict[str, list[int]] = {}
        self.names: set[str] = set()
        self.instruments: dict[str, dict] = {}
        self.assemblies: list[np.ndarray] = []
        self.detectors: list[dict] = []
        self.render: dict = {}
        self.display: dict = {}
        self.notes: list[str] = []
        self.up = (0.0, 0.0, 1.0)
        self.materials: dict[str, int] = {}
        self._dielectric: dict[float, int] = {}
        self._absorber = None
        amb = self.medium("air")
        self.sc.add_region(amb)
        self._register_defaults()

    # ---- errors and values
    def fail(self, v: Value | None, msg: str):
        raise SceneError(f"{self.origin}: line {v.line if v is not None else 0}: {msg}")

    def s(self, v: Value) -> str:
        if v.kind in ("string", "ident"):
            return v.str
        self.fail(v, "expected a name or string")

    def num(self, v: Value) -> float:
        if v.kind != "number":
            self.fail(v, "expected a number")
        return v.num

    def len(self, v: Value) -> float:
        if v.kind != "number":
            self.fail(v, "expected a length")
        if v.unit == F.LENGTH:
            return v.num
        if v.unit != F.NONE:
            self.fail(v, "expected a length unit")
        return v.num * self.unit

    def ang(self, v: Value) -> float:
        if v.kind != "number":
            self.fail(v, "expected an angle")
        if v.unit == F.ANGLE:
            return v.num
        if v.unit != F.NONE:
            self.fail(v, "expected an angle unit")
        return math.radians(v.num)

    def nm(self, v: Value) -> float:
        return v.num * 1e9 if v.unit == F.LENGTH else self.num(v)

    def boolean(self, v: Value) -> bool:
        t = self.s(v)
        if t in ("true", "yes", "on"):
            return True
        if t in ("false", "no", "off"):
            return False
        self.fail(v, "expected true or false")

    def seq(self, v: Value, n: int):
        if v.kind not in ("tuple", "list") or (n and len(v.items) != n):
            self.fail(v, f"expected a tuple of {n} values")
        return v.items

    def vec_num(self, v: Value):
        a = self.seq(v, 3)
        return tuple(self.num(x) for x in a)

    def vec_len(self, v: Value):
        if v.kind == "call" and v.str == "exit_pupil":
            return self.exit_pupil(v)
        if v.kind == "call" and v.str == "sky":
            if len(v.items) != 3:
                self.fail(v, "sky(azimuth, elevation, distance)")
            d = S.direction_from_azimuth_elevation(self.ang(v.items[0]), self.ang(v.items[1]))
            r = self.len(v.items[2])
            return tuple(r * x for x in d)
        if v.kind == "call" and v.str == "on_terrain":
            self.fail(v, "terrain bodies are not supported by the hip backend")
        a = self.seq(v, 3)
        return tuple(self.len(x) for x in a)

    def get(self, b: Value, k: str, conv, default):
        v = b.get(k)
        return conv(v) if v is not None else default

    def need(self, b: Value, k: str) -> Value:
        v = b.get(k)
        if v is None:
            self.fail(b, f"'{b.str} {b.name}' needs '{k}'")
        return v

    def radius(self, b: Value, prefix: str, default: float) -> float:
        v = b.get(prefix + "diameter")
        if v is not None:
            return self.len(v) / 2
        v = b.get(prefix + "radius")
        if v is not None:
            return self.len(v)
        return default

    def half_size(self, v: Value):
        a = self.seq(v, 2)
        return self.len(a[0]) / 2, self.len(a[1]) / 2

    def len_or_inf(self, v: Value) -> float:
        return math.inf if v.kind == "number" and math.isinf(v.num) else self.len(v)

    def up_of(self, b: Value):
        return self.vec_num(b.get("up")) if b.has("up") else self.up

    def path(self, p: str) -> str:
        return p if os.path.isabs(p) else os.path.join(self.base, p)

    # ---- spectra
    def spectrum(self, v: Value, emission: bool) -> S.Spectrum:
        if v.kind == "number":
            return S.Spectrum.constant(v.num)
        if v.kind == "call":
            def arg(i, k):
                x = v.get(k)
                return x if x is not None else (v.items[i] if i < len(v.items) else None)
            scale = self.num(v.get("scale")) if v.has("scale") else 1.0
            if v.str == "rgb":
                if len(v.items) != 3:
                    self.fail(v, "rgb needs three components")
                if v.has("luminance"):
                    scale = self.num(v.get("luminance"))
                r, g, b = (self.num(x) for x in v.items)
                return rgb_spectrum(r, g, b, emission, scale)
            if v.str == "blackbody":
                T = arg(0, "T")
                if T is None:
                    self.fail(v, "blackbody needs a temperature")
                Lv = arg(1, "luminance")
                y = self.num(Lv) if Lv is not None else 1.0
                return S.Spectrum.blackbody(self.num(T), y * scale)
            if v.str == "constant":
                c = arg(0, "value")
                if c is None:
                    self.fail(v, "constant needs a value")
                return S.Spectrum.constant(self.num(c) * scale)
            if v.str == "tabulated":
                lv, vv = arg(0, "lambda"), arg(1, "values")
                if lv is None or vv is None:
                    self.fail(v, "tabulated needs lambda=[...] and values=[...]")
                return S.Spectrum.tabulated([self.nm(x) for x in lv.items], [self.num(x) * scale for x in vv.items])
            if v.str in ("d65", "daylight"):
                Lv = arg(0, "luminance")
                y = self.num(Lv) if Lv is not None else 1.0
                return S.Spectrum.blackbody(6504.0, y * scale)
        self.fail(v, "expected a spectrum: number, rgb(...), blackbody(T, luminance), tabulated(...) or constant(...)")

    # ---- media
    def medium(self, name: str, where: Value | None = None) -> int:
        if name in self.media:
            return self.media[name]
        m = self.defined_media.get(name) or catalog_medium(name)
        if m is None:
            self.fail(where, f"unknown medium '{name}'")
        self.media[name] = self.sc.add_medium(m)
        self.media_names.append(name)
        return self.media[name]

    def index_model(self, v: Value, m: S.Medium) -> None:
        if v.kind == "number":
            m.index_kind, m.a, m.b, m.rel_air = S.INDEX_CONSTANT, (v.num,), (), False
            return
        if v.kind != "call":
            self.fail(v, "expected an index model")

        def lst(k):
            x = v.get(k)
            if x is None:
                self.fail(v, f"missing '{k}'")
            return tuple(self.num(e) for e in x.items)
        if v.str == "constant":
            m.index_kind, m.a, m.b = S.INDEX_CONSTANT, (self.num(v.items[0]),), ()
        elif v.str == "cauchy":
            m.index_kind = S.INDEX_CAUCHY
            m.a = tuple(self.num(v.get(k)) if v.has(k) else d for k, d in (("A", 1.5), ("B", 0.0), ("C", 0.0)))
            m.b = ()
        elif v.str == "sellmeier":
            m.index_kind, m.a, m.b = S.INDEX_SELLMEIER, lst("B"), lst("C")
        elif v.str == "tabulated":
            self.fail(v, "tabulated index models are not supported by the hip backend")
        else:
            self.fail(v, f"unknown index model '{v.str}'")
        m.rel_air = bool(v.has("relative_to_air") and self.boolean(v.get("relative_to_air")))

    def define_medium(self, b: Value) -> None:
        m = S.Medium.constant(1.0)
        if b.has("base"):
            base = catalog_medium(self.s(b.get("base")))
            if base is None:
                self.fail(b.get("base"), "unknown base medium")
            m = base
        if b.has("index"):
            self.index_model(b.get("index"), m)
        if b.has("absorption"):
            m.absorption = self.spectrum(b.get("absorption"), False)
        if b.has("internal_transmittance"):
            t = self.seq(b.get("internal_transmittance"), 2)
            d = self.len(t[1])
            if t[0].kind == "number":
                m.absorption = S.Spectrum.constant(-math.log(self.num(t[0])) / d)
            else:
                tau = self.spectrum(t[0], False)
                lams = np.arange(LAMBDA_MIN, LAMBDA_MAX + 0.1, 5.0)
                vals = [-math.log(max(1e-12, _eval_spectrum(tau, l))) / d for l in lams]
                m.absorption = S.Spectrum.tabulated(lams, vals)
        if b.has("scattering"):
            m.scattering = self.spectrum(b.get("scattering"), False)
        if b.has("g"):
            m.g = self.num(b.get("g"))
        self.defined_media[b.name] = m
        self.media.pop(b.name, None)

    # ---- materials
    def _register_defaults(self):
        add = lambda n, o: self.materials.__setitem__(n, self.sc.add_optics(o))
        add("white", S.Optics(S.SURF_DIFFUSE, reflectance=S.Spectrum.constant(0.8)))
        add("grey", S.Optics(S.SURF_DIFFUSE, reflectance=S.Spectrum.constant(0.18)))
        add("black", S.Optics(S.SURF_ABSORBER))
        add("ideal_mirror", S.Optics(S.SURF_MIRROR, reflectance=S.Spectrum.constant(1.0)))
        for m in ("aluminium", "silver", "gold", "copper"):
            add(m, self._conductor(m))
        self.materials["aluminum"] = self.materials["aluminium"]
        self._absorber = self.materials["black"]

    @staticmethod
    def _conductor(metal: str, roughness: float = 0.0, tint: S.Spectrum | None = None) -> S.Optics:
        lams, n, k = CONDUCTORS[metal]
        return S.Optics(S.SURF_CONDUCTOR, roughness=roughness,
                        reflectance=tint if tint is not None else S.Spectrum.constant(1.0),
                        conductor_n=S.Spectrum.tabulated(lams, n), conductor_k=S.Spectrum.tabulated(lams, k))

    def dielectric(self, roughness: float = 0.0) -> int:
        if roughness not in self._dielectric:
            self._dielectric[roughness] = self.sc.add_optics(S.Optics(S.SURF_DIELECTRIC, roughness=roughness))
        return self._dielectric[roughness]

    def unsupported_surface_keys(self, b: Value):
        for k in ("texture", "relief", "normal_map", "roughness_map"):
            if b.has(k):
                self.fail(b.get(k), f"'{k}' is not supported by the hip backend")

    def roughness(self, b: Value) -> float:
        r = b.get("roughness")
        if r is None:
            return 0.0
        if r.kind == "tuple":
            self.fail(r, "roughness maps need a texture, which the hip backend does not support")
        return self.num(r)

    def define_material(self, b: Value, name: str) -> int:
        t = self.get(b, "type", self.s, "diffuse")
        self.unsupported_surface_keys(b)
        if t in ("diffuse", "lambertian"):
            o = S.Optics(S.SURF_DIFFUSE, reflectance=self.get(b, "reflectance", lambda v: self.spectrum(v, False),
                                                               S.Spectrum.constant(0.8)))
        elif t == "mirror":
            o = S.Optics(S.SURF_MIRROR, reflectance=self.get(b, "reflectance", lambda v: self.spectrum(v, False),
                                                              S.Spectrum.constant(1.0)))
        elif t in ("conductor", "metal"):
            metal = self.get(b, "metal", self.s, "aluminium")
            if metal not in CONDUCTORS:
                self.fail(b, f"unknown metal '{metal}'")
            o = self._conductor(metal, self.roughness(b),
                                self.get(b, "tint", lambda v: self.spectrum(v, False), None))
        elif t == "dielectric":
            o = S.Optics(S.SURF_DIELECTRIC, roughness=self.roughness(b))
        elif t in ("absorber", "black"):
            o = S.Optics(S.SURF_ABSORBER)
        elif t == "null":
            o = S.Optics(S.SURF_NULL)
        elif t == "stained_glass":
            self.fail(b, "stained_glass is not supported by the hip backend")
        else:
            self.fail(b, f"unknown material type '{t}'")
        if b.has("back"):
            o.back_absorbs = self.s(b.get("back")) == "black"
        idx = self.sc.add_optics(o)
        if name:
            self.materials[name] = idx
        return idx

    def material_ref(self, v: Value) -> int:
        if v.kind == "block":
            return self.define_material(v, "")
        n = self.s(v)
        if n not in self.materials:
            self.fail(v, f"unknown material '{n}'")
        return self.materials[n]

    def material_or(self, b: Value, key: str, default: str) -> int:
        return self.material_ref(b.get(key)) if b.has(key) else self.materials[default]

    def solid_material(self, b: Value) -> dict:
        """BodyMaterial (loader_materials.cpp:316-360): a transparent medium with a Fresnel surface (or a
        null one), or an opaque material; optional emission."""
        m = dict(transparent=False, medium=None, optics=None, emission=-1)
        if b.has("medium"):
            name = self.s(b.get("medium"))
            mi = self.medium(name, b.get("medium"))
            if self.sc.media[mi].opaque:
                m["optics"] = self.materials["black"]
            else:
                m["transparent"], m["medium"] = True, name
                self.unsupported_surface_keys(b)
                if b.has("surface"):
                    if self.s(b.get("surface")) != "null":
                        self.fail(b.get("surface"), "a medium body's surface can only be null")
                    m["optics"] = self.sc.add_optics(S.Optics(S.SURF_NULL))
                else:
                    m["optics"] = self.dielectric(self.get(b, "roughness", self.num, 0.0))
        elif b.has("material"):
            m["optics"] = self.material_ref(b.get("material"))
            if self.sc.optics[m["optics"]].type == S.SURF_DIELECTRIC:
                self.fail(b.get("material"), "use 'medium' for transparent bodies")
        else:
            m["optics"] = self.materials["white"]
        if b.has("emission"):
            sides = self.get(b, "emission_sides", self.s, "front")
            e = S.Emission(self.spectrum(b.get("emission"), True), front=sides in ("front", "both"),
                           back=sides in ("back", "both"))
            m["emission"] = self.sc.add_emission(e)
        return m

    def interior(self, name: str, m: dict) -> int:
        med = self.medium(m["medium"]) if m["transparent"] else self.medium("opaque")
        r = self.sc.add_region(med)
        self.body_regions.setdefault(name, []).append(r)
        return r

    # ---- placement
    def placement(self, b: Value) -> np.ndarray:
        pos = self.vec_len(b.get("position")) if b.has("position") else (0.0, 0.0, 0.0)
        R = np.eye(3)
        if b.has("rotate"):
            a = self.seq(b.get("rotate"), 3)
            R = S.rotation((0, 0, 1), self.ang(a[2])) @ S.rotation((0, 1, 0), self.ang(a[1])) @ \
                S.rotation((1, 0, 0), self.ang(a[0]))
        up = self.up_of(b)
        if b.has("axis"):
            R = B.align_z(self.vec_num(b.get("axis")), up)[:, :3] @ R
        elif b.has("look_at"):
            R = B.align_z(np.subtract(self.vec_len(b.get("look_at")), pos), up)[:, :3] @ R
        elif b.has("point_at"):
            R = B.align_z(np.subtract(pos, self.vec_len(b.get("point_at"))), up)[:, :3] @ R
        xf = S.transform(R, pos)
        if b.has("decenter"):
            d = self.seq(b.get("decenter"), 2)
            xf = B.compose(xf, S.transform(None, (self.len(d[0]), self.len(d[1]), 0.0)))
        if b.has("tilt"):
            t = self.seq(b.get("tilt"), 2)
            xf = B.compose(xf, S.transform(S.rotation((1, 0, 0), self.ang(t[0])) @
                                           S.rotation((0, 1, 0), self.ang(t[1]))))
        return xf

    def exit_pupil(self, v: Value):
        if not v.items:
            self.fail(v, "exit_pupil needs an instrument name")
        n = self.s(v.items[0])
        if n not in self.instruments:
            self.fail(v, f"exit_pupil: no instrument named '{n}' (define it first)")
        ins = self.instruments[n]
        off = self.len(v.items[1]) if len(v.items) > 1 else 0.0
        return tuple(_apply(ins["xf"], (0.0, 0.0, ins["paraxial"]["exit_pupil_z"] + off)))

    # ---- statements
    def run(self, doc: Value):
        for k, v in doc.named:
            if k == "units":
                u = self.s(v)
                if u not in F.LENGTH_UNITS:
                    self.fail(v, f"unknown length unit '{u}'")
                self.unit = F.LENGTH_UNITS[u]
            else:
                self.fail(v, f"unknown top-level setting '{k}'")
        for b in doc.items:
            self.statement(b, S.transform())

    def statement(self, b: Value, parent: np.ndarray):
        kw = b.str
        if kw == "medium":
            self.define_medium(b)
        elif kw == "material":
            self.define_material(b, b.name)
        elif kw == "world":
            self.world(b)
        elif kw == "body":
            self.body(b, parent)
        elif kw == "assembly":
            xf = B.compose(parent, self.placement(b))
            for c in b.items:
                self.statement(c, xf)
        elif kw == "observer":
            self.observer(b, parent)
        elif kw == "camera":
            self.fail(b, "physical cameras are not supported by the hip backend")
        elif kw == "sensor":
            self.sensor(b, parent)
        elif kw == "render":
            self.render_block(b)
        else:
            self.fail(b, f"unknown block '{kw}'")

    def world(self, b: Value):
        if b.has("medium"):
            self.sc.region_medium[0] = self.medium(self.s(b.get("medium")), b.get("medium"))
        if b.has("up"):
            u = np.array(self.vec_num(b.get("up")))
            self.up = tuple(u / np.linalg.norm(u))
        if b.has("sky"):
            s = b.get("sky")
            if s.kind == "call" and s.str in ("gradient", "map"):
                self.fail(s, f"{s.str} skies are not supported by the hip backend")
            if s.kind == "call" and s.str == "uniform":
                self.sc.set_uniform_sky(self.spectrum(s.items[0], True))
            elif s.kind == "ident" and s.str == "none":
                self.sc.env["sky"] = S.SKY_NONE
            else:
                self.sc.set_uniform_sky(self.spectrum(s, True))
        for c in b.items:
            if c.str != "sun":
                self.fail(c, f"unknown block in world: {c.str}")
            if c.has("direction"):
                d = np.array(self.vec_num(c.get("direction")))
                d = tuple(d / np.linalg.norm(d))
            else:
                d = S.direction_from_azimuth_elevation(self.get(c, "azimuth", self.ang, 0.0),
                                                       self.get(c, "elevation", self.ang, math.radians(45)))
            if c.has("radiance"):
                rad = self.spectrum(c.get("radiance"), True)
            else:
                rad = S.Spectrum.blackbody(self.get(c, "temperature", self.num, 5778.0),
                                           self.get(c, "luminance", self.num, 1e5))
            self.sc.set_sun(d, self.get(c, "angular_diameter", self.ang, math.radians(0.533)) / 2, rad,
                            self.get(c, "nee_share", self.num, 0.0))

    def render_block(self, b: Value):
        r = self.render
        for k, conv in (("detector", self.s), ("integrator", self.s), ("spp", self.num), ("seed", self.num),
                        ("max_depth", self.num), ("exposure", self.num), ("fresnel_floor", self.num),
                        ("photon_radius", self.num), ("photon_alpha", self.num), ("photons", self.num),
                        ("vcm_groups", self.num), ("indirect_guide", self.s), ("tone", self.s)):
            if b.has(k):
                r[k] = conv(b.get(k))
        if b.has("white_balance"):
            wb = b.get("white_balance")
            r["white_balance"] = 0.0 if wb.kind == "ident" and wb.str == "none" else self.num(wb)
        if b.has("auto_exposure"):
            r["auto_exposure"] = self.boolean(b.get("auto_exposure"))
        if r.get("indirect_guide"):
            self.notes.append(f"indirect_guide = {r['indirect_guide']}: a sampling hint, not applied")

    # ---- bodies
    def body(self, b: Value, parent: np.ndarray):
        if not b.name:
            self.fail(b, "bodies need a name")
        if b.has("repeat"):
            q = self.seq(b.get("repeat"), 2)
            n = int(self.num(q[0]))
            step = self.vec_len(q[1])
            if n < 1:
                self.fail(b.get("repeat"), "repeat count must be positive")
            for i in range(n):
                pre = B.compose(parent, S.transform(None, tuple(i * x for x in step)))
                self.body_impl(b, f"{b.name}_{i}", pre)
            return
        self.body_impl(b, b.name, parent)

    def body_impl(self, b: Value, name: str, pre: np.ndarray):
        t = self.get(b, "type", self.s, "")
        if name in self.names:
            self.fail(b, f"duplicate body name '{name}'")
        self.names.add(name)
        xf = B.compose(pre, self.placement(b))
        self.outside = 0
        if b.has("in"):
            host = self.s(b.get("in"))
            regions = [r for r in self.body_regions.get(host, [])
                       if not self.sc.media[self.sc.region_medium[r]].opaque]
            if not regions:
                self.fail(b.get("in"), f"body '{host}' has no transparent medium to be inside (define it first)")
            self.outside = regions[0]
        builders = {"lens": self.lens_body, "prescription": self.prescription_body, "mirror": self.mirror_body,
                    "flat_mirror": self.flat_mirror_body, "stop": self.stop_body, "iris": self.stop_body,
                    "aperture": self.stop_body, "tube": self.tube_body, "sphere": self.sphere_body,
                    "box": self.box_body, "cylinder": self.cylinder_body, "prism": self.prism_body,
                    "sheet": self.sheet_body, "screen": self.sheet_body, "disk": self.sheet_body,
                    "mesh": self.mesh_body, "waves": self.waves_body}
        unsupported = ("lathe", "torus", "ring", "column", "dome", "round_wall", "starfield", "vessel", "cup",
                       "terrain", "forest", "fractal_statue", "statue")
        if t in unsupported:
            self.fail(b, f"body type '{t}' is not supported by the hip backend")
        if t not in builders:
            self.fail(b, f"unknown body type '{t}'")
        try:
            builders[t](b, name, xf)
        finally:
            self.outside = 0

    def sheet_body(self, b, name, xf):
        m = self.solid_material(b)
        o = self.outside
        if b.has("inner_radius") or b.has("inner_diameter"):
            r, ri = self.radius(b, "", 0.1), self.radius(b, "inner_", 0.0)
            self.sc.add_plane(S.APERTURE_DISK, r, ri, xf, o, o, m["optics"], m["emission"])
        elif self.get(b, "type", self.s, "") == "disk" or b.has("radius") or b.has("diameter"):
            r = self.radius(b, "", 0.1)
            self.sc.add_plane(S.APERTURE_DISK, r, 0.0, xf, o, o, m["optics"], m["emission"])
        else:
            hx, hy = self.half_size(self.need(b, "size"))
            self.sc.add_plane(S.APERTURE_RECT, hx, hy, xf, o, o, m["optics"], m["emission"])

    def sphere_body(self, b, name, xf):
        m = self.solid_material(b)
        self.sc.add_sphere(self.radius(b, "", 0.01), xf, self.outside, self.interior(name, m), m["optics"],
                           m["emission"])

    def _mesh_solid(self, name, pos, tri, xf, m):
        self.sc.add_mesh(pos, tri, xf, self.outside, self.interior(name, m), m["optics"], m["emission"])

    def box_body(self, b, name, xf):
        size = np.array(self.vec_len(self.need(b, "size")))
        pos, tri = S.box_mesh(-0.5 * size, 0.5 * size)
        self._mesh_solid(name, pos, tri, xf, self.solid_material(b))

    def prism_body(self, b, name, xf):
        pos, tri = S.prism_mesh(self.get(b, "apex", self.ang, math.radians(60)), self.len(self.need(b, "side")),
                                self.len(self.need(b, "length")))
        self._mesh_solid(name, np.asarray(pos), np.asarray(tri), xf, self.solid_material(b))

    def cylinder_body(self, b, name, xf):
        m = self.solid_material(b)
        if m["emission"] >= 0:
            self.fail(b.get("emission"), "emitting cylinders are not supported by the hip backend")
        R, h = self.radius(b, "", 0.01), self.len(self.need(b, "height"))
        inside = self.interior(name, m)
        o = self.outside
        self.sc.add_cylinder(R, 0.0, h, xf, o, inside, m["optics"])
        self.sc.add_plane(S.APERTURE_DISK, R, 0.0, B.compose(xf, S.transform(None, (0, 0, h))), o, inside, m["optics"])
        flip = S.transform(S.rotation((1, 0, 0), math.pi))
        self.sc.add_plane(S.APERTURE_DISK, R, 0.0, B.compose(xf, flip), o, inside, m["optics"])

    def stop_body(self, b, name, xf):
        inner = self.radius(b, "aperture_", -1.0)
        if inner < 0:
            inner = self.len(self.need(b, "aperture")) / 2
        outer = self.radius(b, "outer_", inner * 4)
        self.sc.add_plane(S.APERTURE_DISK, outer, inner, xf, self.outside, self.outside, self._absorber)

    def tube_body(self, b, name, xf):
        o = self.material_or(b, "material", "black")
        self.sc.add_cylinder(self.radius(b, "", 0.05), 0.0, self.len(self.need(b, "length")), xf, self.outside,
                             self.outside, o)

    def flat_mirror_body(self, b, name, xf):
        base = self.sc.optics[self.material_or(b, "material", "aluminium")]
        o = self.sc.add_optics(S.Optics(base.type, base.roughness, base.reflectance, base.conductor_n,
                                        base.conductor_k, back_absorbs=True))
        if b.has("ellipse"):
            hx, hy = self.half_size(b.get("ellipse"))
            self.sc.add_plane(S.APERTURE_ELLIPSE, hx, hy, xf, self.outside, self.outside, o)
        else:
            hx, hy = self.half_size(self.need(b, "size"))
            self.sc.add_plane(S.APERTURE_RECT, hx, hy, xf, self.outside, self.outside, o)

    def surface(self, v: Value, sd: float) -> dict:
        """SurfaceSpec (loader_bodies.cpp:56-78): plane/flat, sphere(R), conic(R, k), asphere(R, k, A4…)."""
        s = dict(R=math.inf, k=0.0, A=(), sd=sd)
        if v.kind != "call":
            self.fail(v, "expected a surface: plane(), sphere(R=...), conic(R=..., k=...), asphere(...)")
        if v.str in ("plane", "flat"):
            return s
        R = v.get("R")
        if R is None and v.items:
            R = v.items[0]
        if R is None:
            self.fail(v, "surface needs R")
        s["R"] = self.len_or_inf(R)
        if v.str == "sphere":
            return s
        if v.str in ("conic", "asphere"):
            s["k"] = self.get(v, "k", self.num, 0.0)
            A = []
            for order in range(4, 21, 2):
                a = v.get(f"A{order}")
                if a is not None:
                    slot = (order - 4) // 2
                    A += [0.0] * (slot + 1 - len(A))
                    A[slot] = self.num(a) * self.unit ** (1 - order)
            if len(A) > 4:
                self.fail(v, "asphere terms beyond A10 are not supported by the hip backend")
            s["A"] = tuple(A)
            return s
        self.fail(v, f"unknown surface '{v.str}'")

    def mirror_body(self, b, name, xf):
        sf = self.surface(self.need(b, "surface"), self.radius(b, "", 0.05))
        o = self.material_or(b, "material", "aluminium")
        th = self.get(b, "thickness", self.len, 0.1 * sf["sd"] * 2)
        B.add_mirror(self.sc, sf["R"], sf["k"], sf["sd"], th, o, xf, self.outside, hole=self.radius(b, "hole_", 0.0),
                     asphere=sf["A"])

    def _rim(self, b, default: str) -> int:
        rim = self.get(b, "rim", self.s, default)
        if rim == "black":
            return self._absorber
        if rim == "polished":
            return self.dielectric(0.0)
        if rim != "ground":
            self.fail(b, "rim must be ground, polished or black")
        return self.dielectric(0.35)

    def _medium_object(self, name: str, where=None) -> S.Medium:
        return self.sc.media[self.medium(name, where)]

    def n_of(self, name: str, lam: float) -> float:
        from .output import medium_index
        return medium_index(self._medium_object(name), lam)

    def lens_body(self, b, name, xf):
        r = self.radius(b, "", 0.0125)
        if b.has("focal"):
            glass = self.get(b, "medium", self.s, "N-BK7")
            n = self.n_of(glass, 587.5618) / self.n_of("air", 587.5618)
            t = self.get(b, "thickness", self.len, 0.2 * r)
            try:
                R1, R2 = B.design_simple_lens(self.len(b.get("focal")), n, t, self.get(b, "form", self.s, "bi"))
            except ValueError as e:
                self.fail(b, str(e))
            surfaces = [dict(R=R1, k=0.0, A=(), sd=r), dict(R=R2, k=0.0, A=(), sd=r)]
            media, thick = [glass], [t]
        elif b.has("surfaces"):
            surfaces = [self.surface(s, r) for s in self.need(b, "surfaces").items]
            media = [self.s(m) for m in self.need(b, "media").items]
            thick = [self.len(t) for t in self.need(b, "thicknesses").items]
        else:
            media = [self.s(self.need(b, "medium"))]
            surfaces = [self.surface(self.need(b, "front"), self.radius(b, "front_", r)),
                        self.surface(self.need(b, "back"), self.radius(b, "back_", r))]
            thick = [self.len(self.need(b, "thickness"))]
        if len(surfaces) != len(media) + 1 or len(thick) != len(media) or not media:
            self.fail(b, f"lens '{name}': need n media, n thicknesses and n+1 surfaces")
        for s in surfaces:
            if not s["sd"] > 0:
                self.fail(b, f"lens '{name}': surface without semi-diameter")
        start = len(self.sc.region_medium)
        try:
            B.add_lens(self.sc, [s["R"] for s in surfaces], [s["sd"] for s in surfaces], thick,
                       [self._medium_object(m) for m in media], xf, self.outside, self._rim(b, "ground"),
                       dielectric=self.dielectric(0.0), conics=[s["k"] for s in surfaces],
                       aspheres=[s["A"] for s in surfaces], edge_radius=self.radius(b, "edge_", 0.0))
        except ValueError as e:
            self.fail(b, f"lens '{name}': {e}")
        self.body_regions.setdefault(name, []).extend(range(start, len(self.sc.region_medium)))

    def waves_body(self, b, name, xf):
        hx, hy = self.half_size(self.need(b, "size"))
        depth = self.get(b, "depth", self.len, 0.2)
        margin = self.get(b, "margin", self.len, 0.2 * min(hx, hy))
        medium = self.get(b, "medium", self.s, "water")
        rim = self.get(b, "rim", self.len, 0.25 * max(hx, hy))
        base = self.get(b, "base", self.len, 0.25 * depth)
        bottom = self.material_ref(b.get("bottom")) if b.has("bottom") else -1
        walls = self.material_ref(b.get("walls")) if b.has("walls") else -1
        for o in (bottom, walls):
            if o >= 0 and self.sc.optics[o].type == S.SURF_DIELECTRIC:
                self.fail(b, "bottom and walls must be opaque materials")
        wv = self.need(b, "waves")
        if wv.kind == "call" and wv.str == "ripples":
            rng = self.seq(self.need(wv, "wavelength"), 2)
            waves = ripple_field(int(self.get(wv, "count", self.num, 24)), int(self.get(wv, "seed", self.num, 1)),
                                 self.len(rng[0]), self.len(rng[1]), self.get(wv, "slope", self.num, 0.1),
                                 self.get(wv, "direction", self.ang, 0.0),
                                 self.get(wv, "spread", self.ang, math.radians(90)))
        else:
            waves = []
            for it in wv.items:
                if it.kind != "call" or it.str != "wave":
                    self.fail(wv, "waves must be a list of wave(...) or ripples(...)")
                k = 2 * math.pi / self.len(self.need(it, "wavelength"))
                d = self.get(it, "direction", self.ang, 0.0)
                waves.append((self.len(self.need(it, "amplitude")), k * math.cos(d), k * math.sin(d),
                              self.get(it, "phase", self.ang, 0.0)))
        liquid = self.sc.add_region(self.medium(medium, b))
        self.body_regions.setdefault(name, []).append(liquid)
        add_water_basin(self.sc, waves, hx, hy, margin, depth, xf, self.outside, liquid, self.dielectric(0.0),
                        bottom, walls, rim, base, self.medium("opaque"))

    def prescription_body(self, b, name, xf):
        path = self.path(self.s(self.need(b, "file")))
        try:
            p = L.load_prescription(path)
        except (OSError, ValueError) as e:
            self.fail(b, str(e))
        index_fn = lambda n, lam: self.n_of(n, lam)
        afocal = b.has("afocal") and self.boolean(b.get("afocal"))
        if afocal:
            try:
                p = L.solve_afocal(p, p.wavelength, index_fn)
            except ValueError as e:
                self.fail(b, str(e))
        tube = 0.0
        t = b.get("tube")
        if t is not None and t.kind == "ident":
            if self.boolean(t):
                tube = max(s.sd for s in p.surfaces) * 1e-3 * 1.1 + 2e-3
        elif b.has("tube_radius"):
            tube = self.len(b.get("tube_radius"))
        mirrors = any(s.reflect for s in p.surfaces)
        if tube > 0 and mirrors:
            self.fail(b, "tube = true is not supported for mirror prescriptions: use tube = false and a housing")
        if b.has("eyecup") and self.boolean(b.get("eyecup")):
            self.fail(b.get("eyecup"), "eyecup is not supported by the hip backend")

        def mirror_material(n):
            if n not in self.materials:
                self.fail(b, f"the lens file's mirror material '{n}' is not defined in the scene")
            if self.sc.optics[self.materials[n]].type not in (S.SURF_MIRROR, S.SURF_CONDUCTOR):
                self.fail(b, f"mirror material '{n}' must be a mirror or conductor")
            return self.materials[n]

        start = len(self.sc.region_medium)
        try:
            bi = B.add_prescription(self.sc, p, xf, self.outside, mirror_material, self._rim(b, "black"), tube,
                                    media_of=lambda n: self._medium_object(n))
        except ValueError as e:
            self.fail(b, str(e))
        self.body_regions.setdefault(name, []).extend(range(start, len(self.sc.region_medium)))
        if tube > 0:
            last = bi["vertex_z"][-1]
            z0, z1 = -0.01 * last - 1e-3, last + 1e-3
            self.sc.add_cylinder(tube, z0, z1, xf, self.outside, self.outside, self.materials["black"])
            if b.has("tube_material"):
                self.sc.add_cylinder(tube * 1.03, z0, z1, xf, self.outside, self.outside,
                                     self.material_ref(b.get("tube_material")))
        px = L.paraxial(p, p.wavelength, index_fn)
        self.instruments[name] = dict(xf=xf, paraxial=px, prescription=p)
        if afocal and px["exit_pupil_radius"] > 0:
            self.notes.append(f"{name}: exit pupil guide for observers (a sampling hint) not applied")

    def mesh_body(self, b, name, xf):
        f = self.path(self.s(self.need(b, "file")))
        if b.has("optional") and self.boolean(b.get("optional")) and not os.path.exists(f):
            self.notes.append(f"optional mesh '{name}' not found ({f}); body skipped")
            return
        if not f.lower().endswith(".obj"):
            self.fail(b, "only OBJ meshes are supported by the hip backend")
        if b.has("part"):
            self.fail(b.get("part"), "mesh parts (glTF) are not supported by the hip backend")
        try:
            pos, tri = load_obj(f, self.get(b, "scale", self.num, 1.0) * self.unit)
        except (OSError, ValueError) as e:
            self.fail(b, str(e))
        if b.has("model_up"):
            u = b.get("model_up")
            if u.kind == "tuple":
                up = np.array(self.vec_num(u))
                up /= np.linalg.norm(up)
            else:
                axes = {"x": (1, 0, 0), "-x": (-1, 0, 0), "y": (0, 1, 0), "-y": (0, -1, 0), "z": (0, 0, 1),
                        "-z": (0, 0, -1)}
                if self.s(u) not in axes:
                    self.fail(u, "model_up must be one of x, -x, y, -y, z, -z")
                up = np.array(axes[self.s(u)], float)
            pos = pos @ _rotation_to(up, (0.0, 0.0, 1.0)).T
        if b.has("fit_height") or b.has("ground"):
            lo, hi = pos.min(axis=0), pos.max(axis=0)
            s = self.len(b.get("fit_height")) / (hi[2] - lo[2]) if b.has("fit_height") else 1.0
            foot = np.array([(lo[0] + hi[0]) / 2, (lo[1] + hi[1]) / 2, lo[2]])
            pos = (pos - foot) * s
        m = self.solid_material(b)
        closed = self.boolean(b.get("closed")) if b.has("closed") else m["transparent"]
        if closed:
            self._mesh_solid(name, pos, tri, xf, m)
        else:
            self.sc.add_mesh(pos, tri, xf, self.outside, self.outside, m["optics"], m["emission"])

    # ---- detectors
    def resolution(self, b: Value):
        v = b.get("resolution")
        if v is None:
            return 320, 240
        a = self.seq(v, 2)
        return int(self.num(a[0])), int(self.num(a[1]))

    def detector_display(self, b: Value):
        d = {}
        if b.has("exposure"):
            d["exposure"] = self.num(b.get("exposure"))
        if b.has("white_balance"):
            wb = b.get("white_balance")
            d["white_balance"] = 0.0 if wb.kind == "ident" and wb.str == "none" else self.num(wb)
        if b.has("tone"):
            d["tone"] = self.s(b.get("tone"))
        if b.has("sun_share"):
            d["sun_share"] = self.num(b.get("sun_share"))
            self.notes.append(f"{b.name}: sun_share (a sampling choice) not applied")
        if d:
            self.display[b.name] = d

    def pixel_sigma(self, b: Value, default: float) -> float:
        f = b.get("pixel_filter")
        if f is None:
            return default
        if f.kind == "ident" and f.str == "box":
            s = 0.0
        elif f.kind == "ident" and f.str == "gaussian":
            s = 0.4
        elif f.kind == "call" and f.str == "gaussian" and len(f.items) == 1:
            s = self.num(f.items[0])
        else:
            self.fail(f, "pixel_filter must be box, gaussian or gaussian(sigma in pixels)")
        if not 0 <= s <= 8:
            self.fail(f, "pixel filter sigma must be between 0 and 8 pixels")
        return s

    def observer(self, b: Value, parent: np.ndarray):
        self.detector_display(b)
        pos = self.vec_len(self.need(b, "position"))
        if b.has("look_at"):
            look = self.vec_len(b.get("look_at"))
        elif b.has("direction"):
            d = np.array(self.vec_num(b.get("direction")))
            look = tuple(np.asarray(pos) + d / np.linalg.norm(d))
        else:
            self.fail(b, "observer needs look_at or direction")
        pr = self.radius(b, "pupil_", -1.0)
        if pr < 0:
            pr = self.len(b.get("pupil")) / 2 if b.has("pupil") else 0.0
        focus = self.len_or_inf(b.get("focus")) if b.has("focus") else math.inf
        self.detectors.append(dict(name=b.name, kind="observer", resolution=self.resolution(b), position=pos,
                                   look_at=look, up=self.up_of(b), fov_y=self.get(b, "fov", self.ang, math.radians(40)),
                                   pupil_radius=pr, focus_distance=focus, pixel_sigma=self.pixel_sigma(b, 0.4),
                                   region=0))

    def sensor(self, b: Value, parent: np.ndarray):
        hx, hy = self.half_size(self.need(b, "size"))
        xf = B.compose(parent, self.placement(b))
        if b.has("material"):
            base = self.sc.optics[self.material_ref(b.get("material"))]
            if base.type != S.SURF_DIFFUSE:
                self.fail(b.get("material"), "sensor material must be diffuse (a measurement screen)")
            o = self.sc.add_optics(S.Optics(S.SURF_DIFFUSE, reflectance=base.reflectance,
                                            back_absorbs=base.back_absorbs))
        else:
            o = self.sc.add_optics(S.Optics(S.SURF_DETECTOR))
        d = dict(name=b.name, kind="sensor", resolution=self.resolution(b), half=(hx, hy),
                 pixel_sigma=self.pixel_sigma(b, 0.0))
        if b.has("aim_center"):
            d["aim_center"] = self.vec_len(b.get("aim_center"))
            n = np.array(self.vec_num(self.need(b, "aim_axis")))
            d["aim_normal"] = tuple(n / np.linalg.norm(n))
            d["aim_radius"] = self.radius(b, "aim_", 0.01)
            if self.sc.optics[o].type == S.SURF_DIFFUSE:
                if not d["aim_radius"] > 0:
                    self.fail(b, "screen aim radius must be positive")
                d["aim_share"] = 0.95
        d["boundary"] = self.sc.add_plane(S.APERTURE_RECT, hx, hy, xf, 0, 0, o)
        self.detectors.append(d)


def _eval_spectrum(s: S.Spectrum, l: float) -> float:
    if s.kind == S.SPEC_CONSTANT:
        return s.scale * s.value
    if s.kind == S.SPEC_BLACKBODY:
        return s.scale * float(S.planck(l, s.kelvin))
    lams, vals = zip(*s.table)
    return s.scale * float(np.interp(l, lams, vals))


# ---------------------------------------------------------------- helpers shared with tests


def pcg32_stream(seed: int, stream: int):
    """The reference's Rng (src/owe/core/sampling.hpp:11-31): PCG32 (XSH RR) seeded as state = 0,
    inc = 2·stream + 1, one step, state += seed, one step; uniform() takes 27 + 26 bits of two draws."""
    M = (1 << 64) - 1
    state, inc = 0, ((stream << 1) | 1) & M

    def step():
        nonlocal state
        old = state
        state = (old * 6364136223846793005 + inc) & M
        x = (((old >> 18) ^ old) >> 27) & 0xFFFFFFFF
        rot = old >> 59
        return ((x >> rot) | (x << ((-rot) & 31))) & 0xFFFFFFFF

    step()
    state = (state + seed) & M
    step()

    def uniform():
        a, b = step() >> 5, step() >> 6
        return (a * 67108864.0 + b) * (1.0 / 9007199254740992.0)

    return step, uniform


def ripple_field(count: int, seed: int, min_wavelength: float, max_wavelength: float, slope: float,
                 direction: float, spread: float):
    """rippleField (src/owe/scene/builders.cpp:443-461): count plane waves, wavelengths log-uniform in
    [min, max], directions uniform in direction ± spread/2, random phases, amplitudes slope / (k √count),
    drawn from Rng(seed, stream 0x7269707063)."""
    _, uniform = pcg32_stream(seed, 0x7269707063)
    waves = []
    for _ in range(count):
        u1, u2, u3 = uniform(), uniform(), uniform()
        lam = min_wavelength * (max_wavelength / min_wavelength) ** u1
        k = 2 * math.pi / lam
        th = direction + (u2 - 0.5) * spread
        waves.append((slope / (k * math.sqrt(count)), k * math.cos(th), k * math.sin(th), 2 * math.pi * u3))
    return waves


def add_water_basin(sc: S.Scene, waves, hx, hy, margin, depth, xf, outside, liquid, interface, bottom, walls,
                    rim, base, opaque_medium) -> None:
    """buildWater (src/owe/scene/builders.cpp:402-441): the wave surface over the liquid; the floor and
    four walls facing the liquid — dielectric into the outside without a basin material, else opaque
    against a basin region — and, with a basin, the block it is cut into: a rim at z = 0, four outer
    sides and an underside, all facing the outside."""
    sc.add_wave(waves, hx, hy, margin, xf, outside, liquid, interface)
    basin = sc.add_region(opaque_medium) if bottom >= 0 or walls >= 0 else outside

    def side(o):
        return (basin, o) if o >= 0 else (outside, interface)

    fb, fo = side(bottom)
    sc.add_plane(S.APERTURE_RECT, hx, hy, B.compose(xf, S.transform(None, (0, 0, -depth))), liquid, fb, fo)
    wb, wo = side(walls)
    hz = depth / 2
    for at, inward, half in (((hx, 0, -hz), (-1, 0, 0), hy), ((-hx, 0, -hz), (1, 0, 0), hy),
                             ((0, hy, -hz), (0, -1, 0), hx), ((0, -hy, -hz), (0, 1, 0), hx)):
        w = B.compose(S.transform(None, at), B.align_z(inward, (0, 0, 1)))
        sc.add_plane(S.APERTURE_RECT, half, hz, B.compose(xf, w), liquid, wb, wo)
    if basin != outside:
        ground = walls if walls >= 0 else bottom
        ox, oy, zb = hx + rim, hy + rim, -depth - base
        oz = -zb / 2
        for cx, cy, ex, ey in ((0, (oy + hy) / 2, ox, (oy - hy) / 2), (0, -(oy + hy) / 2, ox, (oy - hy) / 2),
                               ((ox + hx) / 2, 0, (ox - hx) / 2, hy), (-(ox + hx) / 2, 0, (ox - hx) / 2, hy)):
            sc.add_plane(S.APERTURE_RECT, ex, ey, B.compose(xf, S.transform(None, (cx, cy, 0))), outside, basin,
                         ground)
        for at, outward, half in (((ox, 0, zb / 2), (1, 0, 0), oy), ((-ox, 0, zb / 2), (-1, 0, 0), oy),
                                  ((0, oy, zb / 2), (0, 1, 0), ox), ((0, -oy, zb / 2), (0, -1, 0), ox)):
            w = B.compose(S.transform(None, at), B.align_z(outward, (0, 0, 1)))
            sc.add_plane(S.APERTURE_RECT, half, oz, B.compose(xf, w), outside, basin, ground)
        under = B.compose(S.transform(None, (0, 0, zb)), B.align_z((0, 0, -1), (0, 1, 0)))
        sc.add_plane(S.APERTURE_RECT, ox, oy, B.compose(xf, under), outside, basin, ground)


def load_obj(path: str, scale: float = 1.0):
    """Vertices and triangles of a Wavefront OBJ file (``v`` and ``f`` lines; polygons fanned from their
    first corner; negative indices count from the end), vertices times scale."""
    pos, tri = [], []
    with open(path) as f:
        for line in f:
            t = line.split()
            if not t:
                continue
            if t[0] == "v":
                pos.append((float(t[1]), float(t[2]), float(t[3])))
            elif t[0] == "f":
                idx = []
                for c in t[1:]:
                    i = int(c.split("/")[0])
                    idx.append(i - 1 if i > 0 else len(pos) + i)
                for k in range(1, len(idx) - 1):
                    tri.append((idx[0], idx[k], idx[k + 1]))
    if not pos or not tri:
        raise ValueError(f"{path}: no triangles")
    pos = np.array(pos) * scale
    tri = np.array(tri, dtype=np.uint32)
    if tri.max() >= len(pos):
        raise ValueError(f"{path}: face index out of range")
    return pos, tri


def load_scene_text(text: str, origin: str = "<scene>", base_dir: str = ".", edits=()) -> LoadedScene:
    doc = F.parse(text, origin)
    for e in edits:
        F.apply_edit(doc, e)
    ld = _Loader(origin, base_dir)
    ld.run(doc)
    if not ld.detectors:
        raise SceneError(f"{origin}: scene defines no observer, camera or sensor")
    ls = LoadedScene(ld.sc, ld.detectors, ld.render, ld.display, origin, text, tuple(edits), ld.media_names,
                     sorted(ld.names), ld.notes)
    ls.instruments = ld.instruments
    ls.use()
    return ls


def load_scene(path: str, edits=()) -> LoadedScene:
    """Loads a `.owe` file (relative paths inside it resolve against its directory) with edits applied,
    its render block's detector (else the first) made the scene's detector."""
    with open(path) as f:
        text = f.read()
    return load_scene_text(text, path, os.path.dirname(path) or ".", edits)
