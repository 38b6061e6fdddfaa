"""Flattened scenes for the hip backend (include/owe_hip_scene.h) and the measured scene builders.

The reference builds its World from the ``.owe`` language (src/owe/loader) and flattens it for its
float kernels (src/owe/backends/gpu/gpu_scene.cpp); here scenes are built programmatically into the
same plain-struct layout the C ABI takes. Builders mirror the reference's:

* ``buildPrism`` (src/owe/scene/builders.cpp:463-479): a triangular prism mesh, apex at +y, centroid
  at the origin, extruded along z — 8 triangles, outward winding;
* ``buildSheet`` (builders.cpp:481-488): a plane rectangle with the ambient region on both sides;
* ``buildSphere``: one sphere boundary around an interior region;
* the sun: a blackbody of given luminance (``Spectrum::blackbody``, core/spectrum.cpp:37-46) in a
  cone of the given angular radius (loader/scene_loader.cpp:94-101).
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass, field

import numpy as np

# ---------------------------------------------------------------- C layout (owe_hip_scene.h)

SPEC_CONSTANT, SPEC_BLACKBODY, SPEC_TABULATED = 0, 1, 2
INDEX_CONSTANT, INDEX_CAUCHY, INDEX_SELLMEIER, INDEX_CIDDOR = 0, 1, 2, 3
SURF_NULL, SURF_DIELECTRIC, SURF_DIFFUSE, SURF_CONDUCTOR, SURF_MIRROR, SURF_ABSORBER, SURF_DETECTOR = range(7)
SHAPE_MESH, SHAPE_PLANE, SHAPE_SPHERE, SHAPE_CYLINDER, SHAPE_SAG, SHAPE_WAVE = 0, 1, 2, 3, 4, 5
APERTURE_DISK, APERTURE_RECT, APERTURE_ELLIPSE = 0, 1, 2
SKY_NONE, SKY_UNIFORM = 0, 1
INTEGRATOR_PATH, INTEGRATOR_LIGHT, INTEGRATOR_SPPM, INTEGRATOR_VCM, INTEGRATOR_BDPT, INTEGRATOR_HYBRID = 0, 1, 2, 3, 4, 5
SPECTRAL_HERO4, SPECTRAL_UNIFORM1 = 0, 1
DETECTOR_OBSERVER, DETECTOR_SENSOR = 0, 1


class CSpectrum(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("count", ctypes.c_int32), ("offset", ctypes.c_int32),
                ("reserved", ctypes.c_int32), ("value", ctypes.c_double), ("kelvin", ctypes.c_double),
                ("scale", ctypes.c_double), ("reserved2", ctypes.c_double)]


class CMedium(ctypes.Structure):
    _fields_ = [("index_kind", ctypes.c_int32), ("terms", ctypes.c_int32), ("rel_air", ctypes.c_int32),
                ("opaque", ctypes.c_int32), ("coeff_a", ctypes.c_double * 4), ("coeff_b", ctypes.c_double * 4),
                ("absorption", CSpectrum), ("scattering", CSpectrum), ("g", ctypes.c_double),
                ("reserved", ctypes.c_double)]


class COptics(ctypes.Structure):
    _fields_ = [("type", ctypes.c_int32), ("back_absorbs", ctypes.c_int32), ("roughness", ctypes.c_double),
                ("reflectance", CSpectrum), ("conductor_n", CSpectrum), ("conductor_k", CSpectrum)]


class CEmission(ctypes.Structure):
    _fields_ = [("radiance", CSpectrum), ("front", ctypes.c_int32), ("back", ctypes.c_int32),
                ("nee", ctypes.c_int32), ("reserved", ctypes.c_int32)]


class CBoundary(ctypes.Structure):
    _fields_ = [("shape", ctypes.c_int32), ("aperture", ctypes.c_int32), ("front", ctypes.c_uint32),
                ("back", ctypes.c_uint32), ("optics", ctypes.c_int32), ("emission", ctypes.c_int32),
                ("first_triangle", ctypes.c_uint32), ("triangle_count", ctypes.c_uint32), ("a", ctypes.c_double),
                ("b", ctypes.c_double), ("to_world", ctypes.c_double * 12), ("params", ctypes.c_double * 6),
                ("first_wave", ctypes.c_uint32), ("wave_count", ctypes.c_uint32)]


class CPlaneWave(ctypes.Structure):
    _fields_ = [("amplitude", ctypes.c_double), ("kx", ctypes.c_double), ("ky", ctypes.c_double),
                ("phase", ctypes.c_double)]


class CEnvironment(ctypes.Structure):
    _fields_ = [("has_sun", ctypes.c_int32), ("sky", ctypes.c_int32), ("reserved", ctypes.c_int32 * 2),
                ("sun_dir", ctypes.c_double * 3), ("sun_angular_radius", ctypes.c_double),
                ("sun_radiance", CSpectrum), ("sky_radiance", CSpectrum), ("sun_nee_share", ctypes.c_double),
                ("reserved2", ctypes.c_double)]


class CObserver(ctypes.Structure):
    _fields_ = [("width", ctypes.c_int32), ("height", ctypes.c_int32), ("region", ctypes.c_uint32),
                ("kind", ctypes.c_int32), ("position", ctypes.c_double * 3), ("look_at", ctypes.c_double * 3),
                ("up", ctypes.c_double * 3), ("fov_y", ctypes.c_double), ("pupil_radius", ctypes.c_double),
                ("focus_distance", ctypes.c_double), ("pixel_sigma", ctypes.c_double),
                ("sensor_boundary", ctypes.c_uint32), ("flip_x", ctypes.c_int32), ("flip_y", ctypes.c_int32),
                ("has_aim", ctypes.c_int32), ("half_x", ctypes.c_double), ("half_y", ctypes.c_double),
                ("aim_center", ctypes.c_double * 3), ("aim_normal", ctypes.c_double * 3),
                ("aim_radius", ctypes.c_double), ("aim_share", ctypes.c_double)]


class CScene(ctypes.Structure):
    _fields_ = [("n_media", ctypes.c_uint32), ("n_regions", ctypes.c_uint32), ("n_optics", ctypes.c_uint32),
                ("n_emissions", ctypes.c_uint32), ("n_boundaries", ctypes.c_uint32), ("n_vertices", ctypes.c_uint32),
                ("n_triangles", ctypes.c_uint32), ("n_table", ctypes.c_uint32),
                ("media", ctypes.POINTER(CMedium)), ("region_medium", ctypes.POINTER(ctypes.c_uint32)),
                ("optics", ctypes.POINTER(COptics)), ("emissions", ctypes.POINTER(CEmission)),
                ("boundaries", ctypes.POINTER(CBoundary)), ("vertices", ctypes.POINTER(ctypes.c_double)),
                ("triangles", ctypes.POINTER(ctypes.c_uint32)), ("table", ctypes.POINTER(ctypes.c_double)),
                ("env", CEnvironment), ("observer", CObserver), ("waves", ctypes.POINTER(CPlaneWave)),
                ("n_waves", ctypes.c_uint32), ("reserved", ctypes.c_uint32)]


class CRenderSettings(ctypes.Structure):
    _fields_ = [("seed", ctypes.c_uint64), ("max_depth", ctypes.c_int32), ("rr_depth", ctypes.c_int32),
                ("integrator", ctypes.c_int32), ("spectral", ctypes.c_int32), ("lambda_min", ctypes.c_double),
                ("lambda_max", ctypes.c_double), ("fresnel_floor", ctypes.c_double),
                ("photon_radius", ctypes.c_double), ("photon_alpha", ctypes.c_double),
                ("photons_per_pixel", ctypes.c_double), ("photon_groups", ctypes.c_int32), ("reserved", ctypes.c_int32)]


class CTransportStats(ctypes.Structure):
    _fields_ = [("paths", ctypes.c_uint64), ("segments", ctypes.c_uint64), ("inconsistencies", ctypes.c_uint64),
                ("leaks", ctypes.c_uint64)]


# ---------------------------------------------------------------- spectra


def _cie_y(l):
    def lobe(x, mu, s1, s2):
        t = (x - mu) / np.where(x < mu, s1, s2)
        return np.exp(-0.5 * t * t)

    return 0.821 * lobe(l, 568.8, 46.9, 40.5) + 0.286 * lobe(l, 530.9, 16.3, 31.1)


def _simpson_visible(f):
    """Composite Simpson over [360, 830] nm with 0.5 nm steps (core/spectrum.cpp:11-18)."""
    n = int((830.0 - 360.0) * 2)
    h = (830.0 - 360.0) / n
    i = np.arange(n + 1)
    w = np.where((i == 0) | (i == n), 1.0, np.where(i % 2 == 1, 4.0, 2.0))
    return float((f(360.0 + i * h) * w).sum() * h / 3)


def planck(lam_nm, kelvin):
    c1, c2 = 1.191042972e8, 1.438776877e4
    lu = np.asarray(lam_nm, dtype=np.float64) * 1e-3
    return 1e6 * c1 / (lu ** 5 * (np.exp(c2 / (lu * kelvin)) - 1))


def blackbody_scale(kelvin: float, luminance: float) -> float:
    """The scale making Y of the Planck spectrum equal `luminance` (Spectrum::blackbody)."""
    lum = _simpson_visible(lambda l: planck(l, kelvin) * _cie_y(l)) / _simpson_visible(_cie_y)
    return luminance / lum


@dataclass
class Spectrum:
    kind: int = SPEC_CONSTANT
    value: float = 0.0
    kelvin: float = 0.0
    scale: float = 1.0
    table: list = field(default_factory=list)  # [(λ, v), ...]

    @staticmethod
    def constant(v: float) -> "Spectrum":
        return Spectrum(SPEC_CONSTANT, value=float(v))

    @staticmethod
    def blackbody(kelvin: float, luminance: float = 1.0) -> "Spectrum":
        return Spectrum(SPEC_BLACKBODY, kelvin=float(kelvin), scale=blackbody_scale(kelvin, luminance))

    @staticmethod
    def tabulated(lams, vals) -> "Spectrum":
        return Spectrum(SPEC_TABULATED, table=list(zip(map(float, lams), map(float, vals))))


@dataclass
class Medium:
    index_kind: int = INDEX_CONSTANT
    a: tuple = (1.0,)
    b: tuple = ()
    rel_air: bool = False
    absorption: Spectrum = field(default_factory=lambda: Spectrum.constant(0.0))
    scattering: Spectrum = field(default_factory=lambda: Spectrum.constant(0.0))
    g: float = 0.0
    opaque: bool = False

    @staticmethod
    def constant(n: float) -> "Medium":
        return Medium(INDEX_CONSTANT, (float(n),))

    @staticmethod
    def sellmeier(B, C, rel_air: bool = False) -> "Medium":
        return Medium(INDEX_SELLMEIER, tuple(map(float, B)), tuple(map(float, C)), rel_air)

    @staticmethod
    def cauchy(A, B, C=0.0) -> "Medium":
        return Medium(INDEX_CAUCHY, (float(A), float(B), float(C)))

    @staticmethod
    def air() -> "Medium":
        return Medium(INDEX_CIDDOR, ())


# Catalog media used by the measured scenes (src/owe/core/medium.cpp:119-135).
N_BK7 = dict(B=(1.03961212, 0.231792344, 1.01046945), C=(0.00600069867, 0.0200179144, 103.560653), rel_air=True)
PMMA = dict(B=(1.1819,), C=(0.011313,), rel_air=False)


@dataclass
class Optics:
    type: int = SURF_DIELECTRIC
    roughness: float = 0.0
    reflectance: Spectrum = field(default_factory=lambda: Spectrum.constant(1.0))
    conductor_n: Spectrum = field(default_factory=lambda: Spectrum.constant(0.0))
    conductor_k: Spectrum = field(default_factory=lambda: Spectrum.constant(0.0))
    back_absorbs: bool = False


@dataclass
class Emission:
    radiance: Spectrum
    front: bool = True
    back: bool = False
    nee: bool = True


def rotation(axis, angle) -> np.ndarray:
    a = np.asarray(axis, dtype=np.float64)
    a = a / np.linalg.norm(a)
    c, s = math.cos(angle), math.sin(angle)
    t = 1 - c
    x, y, z = a
    return np.array([[t * x * x + c, t * x * y - s * z, t * x * z + s * y],
                     [t * x * y + s * z, t * y * y + c, t * y * z - s * x],
                     [t * x * z - s * y, t * y * z + s * x, t * z * z + c]])


def transform(R=None, t=(0.0, 0.0, 0.0)) -> np.ndarray:
    M = np.zeros((3, 4))
    M[:, :3] = np.eye(3) if R is None else R
    M[:, 3] = t
    return M


class Scene:
    """A scene under construction; ``c()`` returns the C struct (valid while this object lives)."""

    def __init__(self):
        self.media: list[Medium] = []
        self.region_medium: list[int] = []
        self.optics: list[Optics] = []
        self.emissions: list[Emission] = []
        self.boundaries: list[dict] = []
        self.waves: list[tuple] = []
        self.vertices: list[np.ndarray] = []
        self.triangles: list[np.ndarray] = []
        self.n_vertices = 0
        self.n_triangles = 0
        self.env = dict(has_sun=False, sky=SKY_NONE, sun_dir=(0.0, 0.0, 1.0), sun_angular_radius=0.0,
                        sun_radiance=Spectrum.constant(0.0), sky_radiance=Spectrum.constant(0.0), sun_nee_share=0.0)
        self.observer = None
        self._keep = []

    def add_medium(self, m: Medium) -> int:
        self.media.append(m)
        return len(self.media) - 1

    def add_region(self, medium: int) -> int:
        self.region_medium.append(int(medium))
        return len(self.region_medium) - 1

    def add_optics(self, o: Optics) -> int:
        self.optics.append(o)
        return len(self.optics) - 1

    def add_emission(self, e: Emission) -> int:
        self.emissions.append(e)
        return len(self.emissions) - 1

    def add_plane(self, aperture: int, a: float, b: float, xf, front: int, back: int, optics: int,
                  emission: int = -1) -> int:
        self.boundaries.append(dict(shape=SHAPE_PLANE, aperture=aperture, a=a, b=b, xf=np.asarray(xf, float),
                                    front=front, back=back, optics=optics, emission=emission, first=0, count=0))
        return len(self.boundaries) - 1

    def add_sphere(self, radius: float, xf, front: int, back: int, optics: int, emission: int = -1) -> int:
        self.boundaries.append(dict(shape=SHAPE_SPHERE, aperture=0, a=radius, b=0.0, xf=np.asarray(xf, float),
                                    front=front, back=back, optics=optics, emission=emission, first=0, count=0))
        return len(self.boundaries) - 1

    def add_cylinder(self, radius: float, z0: float, z1: float, xf, front: int, back: int, optics: int) -> int:
        """Open cylinder x² + y² = radius², z0 ≤ z ≤ z1, normal outward (CylinderShape)."""
        self.boundaries.append(dict(shape=SHAPE_CYLINDER, aperture=0, a=radius, b=0.0, xf=np.asarray(xf, float),
                                    front=front, back=back, optics=optics, emission=-1, first=0, count=0,
                                    params=(z0, z1)))
        return len(self.boundaries) - 1

    def add_sag(self, curvature: float, conic: float, asphere, r_max: float, r_min: float, xf, front: int, back: int,
                optics: int) -> int:
        """z = sag(r) for r_min ≤ r ≤ r_max, vertex at the origin, normal +z (SagSurface)."""
        A = list(asphere) + [0.0] * (4 - len(asphere))
        self.boundaries.append(dict(shape=SHAPE_SAG, aperture=0, a=r_max, b=r_min, xf=np.asarray(xf, float),
                                    front=front, back=back, optics=optics, emission=-1, first=0, count=0,
                                    params=(curvature, conic, *A[:4])))
        return len(self.boundaries) - 1

    def add_wave(self, waves, half_x: float, half_y: float, margin: float, xf, front: int, back: int, optics: int,
                 spacing: float = -1.0) -> int:
        """z = h(x, y) over the rectangle, a sum of plane waves (amplitude, kx, ky, phase) tapered
        to zero within `margin` of the edges (WaveSurface); spacing > 0 samples it as a bicubic
        Hermite spline, < 0 chooses as the reference does (analytic up to six waves)."""
        first = len(self.waves)
        self.waves.extend(tuple(map(float, w)) for w in waves)
        self.boundaries.append(dict(shape=SHAPE_WAVE, aperture=0, a=half_x, b=half_y, xf=np.asarray(xf, float),
                                    front=front, back=back, optics=optics, emission=-1, first=0, count=0,
                                    params=(margin, spacing), waves=(first, len(waves))))
        return len(self.boundaries) - 1

    def add_mesh(self, positions, triangles, xf, front: int, back: int, optics: int, emission: int = -1) -> int:
        pos = np.asarray(positions, dtype=np.float64).reshape(-1, 3)
        tri = np.asarray(triangles, dtype=np.uint32).reshape(-1, 3) + self.n_vertices
        self.boundaries.append(dict(shape=SHAPE_MESH, aperture=0, a=0.0, b=0.0, xf=np.asarray(xf, float),
                                    front=front, back=back, optics=optics, emission=emission,
                                    first=self.n_triangles, count=len(tri)))
        self.vertices.append(pos)
        self.triangles.append(tri)
        self.n_vertices += len(pos)
        self.n_triangles += len(tri)
        return len(self.boundaries) - 1

    def set_sun(self, direction, angular_radius: float, radiance: Spectrum, nee_share: float = 0.0):
        d = np.asarray(direction, dtype=np.float64)
        self.env.update(has_sun=True, sun_dir=tuple(d / np.linalg.norm(d)), sun_angular_radius=float(angular_radius),
                        sun_radiance=radiance, sun_nee_share=nee_share)

    def set_uniform_sky(self, radiance: Spectrum):
        self.env.update(sky=SKY_UNIFORM, sky_radiance=radiance)

    def set_observer(self, width, height, position, look_at, up=(0.0, 0.0, 1.0), fov_y=math.radians(40),
                     pupil_radius=0.0, focus_distance=math.inf, pixel_sigma=0.4, region=0):
        self.observer = dict(kind=DETECTOR_OBSERVER, width=width, height=height, position=position, look_at=look_at,
                             up=up, fov_y=fov_y, pupil_radius=pupil_radius, focus_distance=focus_distance,
                             pixel_sigma=pixel_sigma, region=region)

    def set_sensor(self, boundary, width, height, half_x, half_y, flip_x=False, flip_y=False, pixel_sigma=0.0,
                   aim_center=None, aim_normal=(0.0, 0.0, 1.0), aim_radius=0.0, aim_share=1.0):
        """The detector is the surface sensor on plane boundary `boundary` (SurfaceSensor,
        src/owe/scene/detector.hpp:125-160): irradiance on its front face over a pixel grid."""
        self.observer = dict(kind=DETECTOR_SENSOR, width=width, height=height, sensor_boundary=boundary,
                             half_x=half_x, half_y=half_y, flip_x=flip_x, flip_y=flip_y, pixel_sigma=pixel_sigma,
                             has_aim=aim_center is not None, aim_center=aim_center or (0.0, 0.0, 0.0),
                             aim_normal=aim_normal, aim_radius=aim_radius, aim_share=aim_share)

    # ---- flattening

    def _spectrum(self, s: Spectrum, table: list) -> CSpectrum:
        c = CSpectrum()
        c.kind, c.value, c.kelvin, c.scale = s.kind, s.value, s.kelvin, s.scale
        if s.kind == SPEC_TABULATED:
            c.count = len(s.table)
            c.offset = len(table)
            for lam, v in s.table:
                table.extend((lam, v))
        return c

    def c(self) -> CScene:
        if self.observer is None:
            raise ValueError("the scene has no observer")
        table: list[float] = []
        media = (CMedium * max(1, len(self.media)))()
        for i, m in enumerate(self.media):
            cm = media[i]
            cm.index_kind, cm.rel_air, cm.opaque, cm.g = m.index_kind, int(m.rel_air), int(m.opaque), m.g
            cm.terms = len(m.a) if m.index_kind == INDEX_SELLMEIER else 0
            for k, v in enumerate(m.a):
                cm.coeff_a[k] = v
            for k, v in enumerate(m.b):
                cm.coeff_b[k] = v
            cm.absorption = self._spectrum(m.absorption, table)
            cm.scattering = self._spectrum(m.scattering, table)
        optics = (COptics * max(1, len(self.optics)))()
        for i, o in enumerate(self.optics):
            co = optics[i]
            co.type, co.roughness, co.back_absorbs = o.type, o.roughness, int(o.back_absorbs)
            co.reflectance = self._spectrum(o.reflectance, table)
            co.conductor_n = self._spectrum(o.conductor_n, table)
            co.conductor_k = self._spectrum(o.conductor_k, table)
        emissions = (CEmission * max(1, len(self.emissions)))()
        for i, e in enumerate(self.emissions):
            ce = emissions[i]
            ce.radiance = self._spectrum(e.radiance, table)
            ce.front, ce.back, ce.nee = int(e.front), int(e.back), int(e.nee)
        bounds = (CBoundary * max(1, len(self.boundaries)))()
        for i, b in enumerate(self.boundaries):
            cb = bounds[i]
            cb.shape, cb.aperture, cb.front, cb.back = b["shape"], b["aperture"], b["front"], b["back"]
            cb.optics, cb.emission, cb.first_triangle, cb.triangle_count = b["optics"], b["emission"], b["first"], b["count"]
            cb.a, cb.b = b["a"], b["b"]
            for k, v in enumerate(b["xf"].reshape(-1)):
                cb.to_world[k] = v
            for k, v in enumerate(b.get("params", ())):
                cb.params[k] = v
            cb.first_wave, cb.wave_count = b.get("waves", (0, 0))
        waves = (CPlaneWave * max(1, len(self.waves)))()
        for i, w in enumerate(self.waves):
            waves[i].amplitude, waves[i].kx, waves[i].ky, waves[i].phase = w
        regions = np.ascontiguousarray(self.region_medium, dtype=np.uint32)
        verts = (np.ascontiguousarray(np.concatenate(self.vertices), dtype=np.float64) if self.vertices
                 else np.zeros((1, 3)))
        tris = (np.ascontiguousarray(np.concatenate(self.triangles), dtype=np.uint32) if self.triangles
                else np.zeros((1, 3), np.uint32))
        env = CEnvironment()
        env.has_sun, env.sky = int(self.env["has_sun"]), self.env["sky"]
        for k in range(3):
            env.sun_dir[k] = self.env["sun_dir"][k]
        env.sun_angular_radius = self.env["sun_angular_radius"]
        env.sun_radiance = self._spectrum(self.env["sun_radiance"], table)
        env.sky_radiance = self._spectrum(self.env["sky_radiance"], table)
        env.sun_nee_share = self.env["sun_nee_share"]
        tab = np.ascontiguousarray(table if table else [0.0], dtype=np.float64)
        ob = CObserver()
        o = self.observer
        ob.kind, ob.width, ob.height, ob.pixel_sigma = o["kind"], o["width"], o["height"], o["pixel_sigma"]
        if o["kind"] == DETECTOR_OBSERVER:
            ob.region = o["region"]
            for k in range(3):
                ob.position[k], ob.look_at[k], ob.up[k] = o["position"][k], o["look_at"][k], o["up"][k]
            ob.fov_y, ob.pupil_radius, ob.focus_distance = o["fov_y"], o["pupil_radius"], o["focus_distance"]
        else:
            ob.sensor_boundary, ob.half_x, ob.half_y = o["sensor_boundary"], o["half_x"], o["half_y"]
            ob.flip_x, ob.flip_y, ob.has_aim = int(o["flip_x"]), int(o["flip_y"]), int(o["has_aim"])
            for k in range(3):
                ob.aim_center[k], ob.aim_normal[k] = o["aim_center"][k], o["aim_normal"][k]
            ob.aim_radius, ob.aim_share = o["aim_radius"], o["aim_share"]
        s = CScene()
        s.n_media, s.n_regions, s.n_optics = len(self.media), len(regions), len(self.optics)
        s.n_emissions, s.n_boundaries = len(self.emissions), len(self.boundaries)
        s.n_vertices, s.n_triangles, s.n_table = self.n_vertices, self.n_triangles, len(table)
        s.media = ctypes.cast(media, ctypes.POINTER(CMedium))
        s.region_medium = regions.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))
        s.optics = ctypes.cast(optics, ctypes.POINTER(COptics))
        s.emissions = ctypes.cast(emissions, ctypes.POINTER(CEmission))
        s.boundaries = ctypes.cast(bounds, ctypes.POINTER(CBoundary))
        s.vertices = verts.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        s.triangles = tris.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))
        s.table = tab.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        s.env, s.observer = env, ob
        s.waves, s.n_waves = ctypes.cast(waves, ctypes.POINTER(CPlaneWave)), len(self.waves)
        self._keep = [media, optics, emissions, bounds, regions, verts, tris, tab, waves]
        return s


def settings(seed=1, max_depth=64, rr_depth=6, spectral=SPECTRAL_HERO4, lambda_min=380.0, lambda_max=780.0,
             fresnel_floor=0.0, integrator=INTEGRATOR_PATH, photon_radius=2.0, photon_alpha=2.0 / 3.0,
             photons_per_pixel=1.0, photon_groups=None) -> CRenderSettings:
    """RenderSettings (scene/scene.hpp:15-40); the photon_* fields are the sppm integrator's, and
    vcm's (merging radius, its shrink, light subpaths per pixel, wavelength groups) and bdpt's.
    photon_groups defaults to the reference's: 4 for sppm, 1 for vcm and bdpt."""
    if photon_groups is None:
        photon_groups = 1 if integrator in (INTEGRATOR_VCM, INTEGRATOR_BDPT) else 4
    s = CRenderSettings()
    s.seed, s.max_depth, s.rr_depth, s.integrator, s.spectral = seed, max_depth, rr_depth, integrator, spectral
    s.lambda_min, s.lambda_max, s.fresnel_floor = lambda_min, lambda_max, fresnel_floor
    s.photon_radius, s.photon_alpha, s.photons_per_pixel = photon_radius, photon_alpha, photons_per_pixel
    s.photon_groups = photon_groups
    return s


# ---------------------------------------------------------------- builders


def prism_mesh(apex: float, side: float, length: float):
    """The reference's buildPrism geometry: 8 triangles, outward winding."""
    hx, hy = side * math.sin(apex / 2), side * math.cos(apex / 2)
    A, B, C = np.array([0, 2 * hy / 3, 0]), np.array([-hx, -hy / 3, 0]), np.array([hx, -hy / 3, 0])
    dz = np.array([0, 0, length / 2])
    pos, tri = [], []

    def add_tri(a, b, c):
        base = len(pos)
        pos.extend([a, b, c])
        tri.append([base, base + 1, base + 2])

    def add_quad(a, b, c, d):
        base = len(pos)
        pos.extend([a, b, c, d])
        tri.append([base, base + 1, base + 2])
        tri.append([base, base + 2, base + 3])

    add_tri(A - dz, B - dz, C - dz)
    add_tri(A + dz, C + dz, B + dz)
    add_quad(A - dz, A + dz, B + dz, B - dz)
    add_quad(B - dz, B + dz, C + dz, C - dz)
    add_quad(C - dz, C + dz, A + dz, A - dz)
    pos = np.array(pos, dtype=np.float64)
    tri = np.array(tri, dtype=np.uint32)
    # Orient outward: each triangle's normal away from the centroid.
    c = pos.mean(axis=0)
    for k, t in enumerate(tri):
        n = np.cross(pos[t[1]] - pos[t[0]], pos[t[2]] - pos[t[0]])
        if np.dot(n, pos[t].mean(axis=0) - c) < 0:
            tri[k] = t[[0, 2, 1]]
    return pos, tri


def direction_from_azimuth_elevation(az: float, el: float):
    """Unit vector at azimuth az (from +y toward +x) and elevation el (core/math.hpp:11-13)."""
    return (math.cos(el) * math.sin(az), math.cos(el) * math.cos(az), math.sin(el))


def dispersion_prism_scene(width=256, height=256, sun_elevation=math.radians(30.0), sun_azimuth=math.radians(250.0),
                           lift=0.0002, sun: bool = True, sky: float = 0.0) -> Scene:
    """Plain spectral path tracing of a dispersion prism in sunlight.

    An equilateral glass prism (40 mm sides, 20 mm tall, Sellmeier B = (1.0396, 0.2318, 1.0105),
    C = (0.0060, 0.0200, 103.56) µm², absolute) stands upright on a 1 m² Lambertian floor of albedo
    0.8, in air of constant n = 1.000277, lit by a 5778 K blackbody sun of 0.53° diameter and
    luminance 1e5. The prism floats `lift` (0.2 mm) above the floor, as the reference's own prism
    scenes do to avoid a shared face (scenes/prism_in_sunlight.owe:27-28). A 256×256 pinhole camera
    looks down at the prism and the spectrum it throws on the floor.
    """
    sc = Scene()
    air = sc.add_medium(Medium.constant(1.000277))
    glass = sc.add_medium(Medium.sellmeier((1.0396, 0.2318, 1.0105), (0.0060, 0.0200, 103.56)))
    amb = sc.add_region(air)
    inside = sc.add_region(glass)
    diel = sc.add_optics(Optics(SURF_DIELECTRIC))
    floor = sc.add_optics(Optics(SURF_DIFFUSE, reflectance=Spectrum.constant(0.8)))
    sc.add_plane(APERTURE_RECT, 0.5, 0.5, transform(), amb, amb, floor)
    pos, tri = prism_mesh(math.radians(60), 0.040, 0.020)
    R = rotation((0, 0, 1), math.radians(-20.0))
    sc.add_mesh(pos, tri, transform(R, (0.0, 0.0, 0.010 + lift)), amb, inside, diel)
    if sun:
        sun_dir = direction_from_azimuth_elevation(sun_azimuth, sun_elevation)
        sc.set_sun(sun_dir, math.radians(0.53) / 2, Spectrum.blackbody(5778.0, 1e5))
    if sky > 0:
        sc.set_uniform_sky(Spectrum.constant(sky))
    sc.set_observer(width, height, position=(0.09, -0.16, 0.16), look_at=(0.035, 0.02, 0.0), fov_y=math.radians(40))
    return sc


def white_furnace_scene(rho: float, size: int = 16, glass: float = 0.0) -> Scene:
    """A closed diffuse sphere of albedo rho emitting unit radiance inward: L = 1 / (1 − ρ)
    (tests/test_transport.cpp:83-114). With glass > 0 a lossless dielectric ball of that radius
    (n = 1.5) sits inside, behind the eye: a uniform radiance field stays uniform through it, so the
    eye still sees 1 / (1 − ρ) everywhere, part of it light that crossed the glass on its way to the
    wall the eye sees."""
    sc = Scene()
    air = sc.add_medium(Medium.air())
    amb = sc.add_region(air)
    inside = sc.add_region(air)
    shell = sc.add_optics(Optics(SURF_DIFFUSE, reflectance=Spectrum.constant(rho)))
    e = sc.add_emission(Emission(Spectrum.constant(1.0), front=False, back=True))
    sc.add_sphere(1.0, transform(), amb, inside, shell, e)
    if glass > 0:
        ball = sc.add_region(sc.add_medium(Medium.constant(1.5)))
        sc.add_sphere(glass, transform(None, (-0.45, 0.0, 0.0)), inside, ball, sc.add_optics(Optics(SURF_DIELECTRIC)))
    sc.set_observer(size, size, position=(0.2, 0.1, -0.3), look_at=(1, 0.3, 0), fov_y=math.radians(70), region=inside)
    return sc


def _as_spectrum(v) -> Spectrum:
    return v if isinstance(v, Spectrum) else Spectrum.constant(float(v))


def scattering_furnace_scene(sigma_s, sigma_a=0.0, g: float = 0.0, rho: float = 0.0, size: int = 16,
                             at_centre: bool = False, bubble: float = 0.0) -> Scene:
    """A unit sphere filled with a participating medium (scattering sigma_s and absorption sigma_a
    per metre, constants or spectra; Henyey–Greenstein asymmetry g), its inner wall diffuse of albedo
    rho and emitting unit radiance inward; the eye is inside, at the centre if `at_centre`.

    With sigma_a = 0 the radiance everywhere inside is 1 / (1 − ρ) whatever sigma_s(λ) and g are: a
    conservative medium leaves a uniform radiance field uniform, and the wall then closes the
    furnace balance L = 1 + ρL. With absorption the eye at the centre sees, in every direction,
    E[ϖ^N] (ρ = 0), N the collisions of a walk from the centre to the wall and ϖ = σs/(σs + σa).

    `bubble` > 0 puts the eye at the centre of a clear sphere of that radius behind a null boundary
    (light tracing then never connects a medium vertex closer than the bubble to the eye: a vertex
    at distance r contributes ∝ 1/r², whose second moment diverges as r → 0)."""
    sc = Scene()
    air = sc.add_medium(Medium.air())
    fog = Medium.constant(1.0)
    fog.scattering, fog.absorption, fog.g = _as_spectrum(sigma_s), _as_spectrum(sigma_a), float(g)
    amb = sc.add_region(air)
    inside = sc.add_region(sc.add_medium(fog))
    shell = sc.add_optics(Optics(SURF_DIFFUSE, reflectance=Spectrum.constant(rho)))
    e = sc.add_emission(Emission(Spectrum.constant(1.0), front=False, back=True))
    sc.add_sphere(1.0, transform(), amb, inside, shell, e)
    eye = inside
    if bubble > 0:
        eye = sc.add_region(sc.add_medium(Medium.constant(1.0)))
        sc.add_sphere(bubble, transform(), inside, eye, sc.add_optics(Optics(SURF_NULL)))
    pos = (0.0, 0.0, 0.0) if at_centre or bubble > 0 else (0.2, 0.1, -0.3)
    sc.set_observer(size, size, position=pos, look_at=(1, 0.3, 0), fov_y=math.radians(70), region=eye)
    return sc


def rough_glass_scene(alpha: float = 0.2, size: int = 32, sun: bool = True, sky: float = 0.2,
                      fog: float = 0.0, dome: bool = False) -> Scene:
    """Frosted glass in sunlight: a rough N-BK7 ball (radius 50 mm, GGX α = `alpha`) and a rough
    plate (100 × 80 × 10 mm, the same α) standing on a diffuse floor of albedo 0.7 under a 0.53° sun
    40° high and a uniform sky; the eye looks at both. Every rough-dielectric vertex takes next-event
    estimation toward the sun (reflection lobe) and samples reflection or transmission through the
    GGX visible normals; light reaches the floor behind the plate only through it. `fog` > 0 fills
    the ball with a scattering medium (σs = fog per metre, g = 0.6). `dome` closes the scene in a
    black sphere of radius 2 m emitting unit radiance inward, in place of sun and sky: every light
    path then starts on a real surface, so light tracing sees everything camera paths see."""
    sc = Scene()
    air = sc.add_medium(Medium.constant(1.000277))
    glass = Medium.sellmeier(**N_BK7)
    amb = sc.add_region(air)
    inside_plate = sc.add_region(sc.add_medium(glass))
    ball_glass = Medium.sellmeier(**N_BK7)
    if fog > 0:
        ball_glass.scattering, ball_glass.g = Spectrum.constant(fog), 0.6
    inside_ball = sc.add_region(sc.add_medium(ball_glass))
    frosted = sc.add_optics(Optics(SURF_DIELECTRIC, roughness=alpha))
    floor = sc.add_optics(Optics(SURF_DIFFUSE, reflectance=Spectrum.constant(0.7)))
    sc.add_plane(APERTURE_RECT, 0.5, 0.5, transform(), amb, amb, floor)
    sc.add_sphere(0.05, transform(None, (-0.06, 0.0, 0.0502)), amb, inside_ball, frosted)
    pos, tri = box_mesh((-0.05, -0.005, 0.0), (0.05, 0.005, 0.08))
    sc.add_mesh(pos, tri, transform(rotation((0, 0, 1), math.radians(15.0)), (0.07, 0.02, 0.0002)), amb,
                inside_plate, frosted)
    if dome:
        black = sc.add_optics(Optics(SURF_DIFFUSE, reflectance=Spectrum.constant(0.0)))
        glow = sc.add_emission(Emission(Spectrum.constant(1.0), front=False, back=True))
        sc.add_sphere(2.0, transform(), sc.add_region(air), amb, black, glow)
    else:
        if sun:
            sc.set_sun(direction_from_azimuth_elevation(math.radians(200.0), math.radians(40.0)),
                       math.radians(0.53) / 2, Spectrum.blackbody(5778.0, 1e3))
        if sky > 0:
            sc.set_uniform_sky(Spectrum.constant(sky))
    sc.set_observer(size, size, position=(0.0, -0.35, 0.2), look_at=(0.0, 0.0, 0.04), fov_y=math.radians(40))
    return sc


def dielectric_sphere_scene(size: int = 24) -> Scene:
    """A lossless N-SF11-like sphere in a uniform sky is invisible (test_transport.cpp:116-142)."""
    sc = Scene()
    air = sc.add_medium(Medium.air())
    glass = sc.add_medium(Medium.sellmeier((1.73759695, 0.313747346, 1.89878101),
                                           (0.013188707, 0.0623068142, 155.23629), rel_air=True))
    amb = sc.add_region(air)
    inside = sc.add_region(glass)
    diel = sc.add_optics(Optics(SURF_DIELECTRIC))
    sc.add_sphere(1.0, transform(), amb, inside, diel)
    sc.set_uniform_sky(Spectrum.constant(1.0))
    sc.set_observer(size, size, position=(0, -4, 0), look_at=(0, 0, 0), fov_y=math.radians(30))
    return sc


def disk_irradiance(L: float, a: float, h: float, r: float) -> float:
    """Irradiance at distance r from the axis on a plane h below a Lambertian disk of radius a and
    radiance L (the exact off-axis formula of tests/analytic.hpp:10-14)."""
    s = h * h + r * r + a * a
    return math.pi * L / 2 * (1 - (h * h + r * r - a * a) / math.sqrt(s * s - 4 * r * r * a * a))


def mean_disk_irradiance(L: float, a: float, h: float, half: float, n: int = 200) -> float:
    """disk_irradiance averaged over a centred square of half-size `half` (tests/analytic.hpp:16-25)."""
    c = -half + (np.arange(n) + 0.5) * 2 * half / n
    x, y = np.meshgrid(c, c)
    r = np.sqrt(x * x + y * y)
    s = h * h + r * r + a * a
    e = math.pi * L / 2 * (1 - (h * h + r * r - a * a) / np.sqrt(s * s - 4 * r * r * a * a))
    return float(e.mean())


def disk_irradiance_scene(a: float = 0.3, h: float = 0.5, half: float = 0.15, aimed: bool = False,
                          screen: bool = False, size: int = 2) -> Scene:
    """A Lambertian disk emitter (radiance 2, radius a) facing down from height h onto an upward
    square sensor of half-size `half` (tests/analytic.hpp:28-60): a detector surface, or a diffuse
    screen of albedo 0.7; optionally aimed at the disk."""
    sc = Scene()
    air = sc.add_medium(Medium.air())
    amb = sc.add_region(air)
    black = sc.add_optics(Optics(SURF_ABSORBER))
    e = sc.add_emission(Emission(Spectrum.constant(2.0)))
    sc.add_plane(APERTURE_DISK, a, 0.0, transform(rotation((1, 0, 0), math.pi), (0.0, 0.0, h)), amb, amb, black, e)
    so = sc.add_optics(Optics(SURF_DIFFUSE, reflectance=Spectrum.constant(0.7)) if screen else Optics(SURF_DETECTOR))
    sb = sc.add_plane(APERTURE_RECT, half, half, transform(), amb, amb, so)
    if aimed:
        sc.set_sensor(sb, size, size, half, half, aim_center=(0.0, 0.0, h), aim_normal=(0.0, 0.0, 1.0), aim_radius=a,
                      aim_share=0.95 if screen else 1.0)
    else:
        sc.set_sensor(sb, size, size, half, half)
    return sc


def ball_caustic_scene(size: int = 16, sun_radius_deg: float = 3.0, dispersive: bool = True) -> Scene:
    """A glass ball (radius 20 mm, centre 40 mm up) focusing a sun straight overhead onto a diffuse
    floor (half-size 0.1 m, albedo 0.8); a pinhole eye 15 mm above the floor, below the ball, looks
    straight down at the caustic. Everything the eye sees is diffuse, so camera paths and particles
    estimate the same image; the light reaching it has crossed two refractions (importance transport
    drops the n² radiance factor that camera paths carry). The sun is widened to `sun_radius_deg` so
    camera paths find it through the ball."""
    sc = Scene()
    air = sc.add_medium(Medium.constant(1.000277))
    glass = (sc.add_medium(Medium.sellmeier((1.03961212, 0.231792344, 1.01046945), (0.00600069867, 0.0200179144,
                                                                                     103.560653)))
             if dispersive else sc.add_medium(Medium.constant(1.5)))
    amb = sc.add_region(air)
    inside = sc.add_region(glass)
    floor = sc.add_optics(Optics(SURF_DIFFUSE, reflectance=Spectrum.constant(0.8)))
    diel = sc.add_optics(Optics(SURF_DIELECTRIC))
    sc.add_plane(APERTURE_RECT, 0.1, 0.1, transform(), amb, amb, floor)
    sc.add_sphere(0.02, transform(None, (0.0, 0.0, 0.04)), amb, inside, diel)
    sc.set_sun((0.0, 0.0, 1.0), math.radians(sun_radius_deg), Spectrum.constant(1.0))
    sc.set_observer(size, size, position=(0.0, 0.0, 0.015), look_at=(0.0, 0.0, 0.0), up=(0.0, 1.0, 0.0),
                    fov_y=math.radians(90))
    return sc


def box_mesh(lo, hi):
    """An axis-aligned box as 12 triangles wound outward (the front side faces out)."""
    lo, hi = np.asarray(lo, float), np.asarray(hi, float)
    pos = np.array([[(hi if (k >> a) & 1 else lo)[a] for a in range(3)] for k in range(8)])
    quads = [(0, 2, 6, 4), (1, 5, 7, 3), (0, 4, 5, 1), (2, 3, 7, 6), (0, 1, 3, 2), (4, 6, 7, 5)]
    tri = []
    c = pos.mean(axis=0)
    for q in quads:
        for t in ((q[0], q[1], q[2]), (q[0], q[2], q[3])):
            n = np.cross(pos[t[1]] - pos[t[0]], pos[t[2]] - pos[t[0]])
            tri.append(t if np.dot(n, pos[list(t)].mean(axis=0) - c) > 0 else (t[0], t[2], t[1]))
    return pos, np.array(tri, dtype=np.uint32)


def submerged_sensor_scene(size: int = 2, n_water: float = 1.333) -> Scene:
    """The disk emitter of disk_irradiance_scene (radius 0.3, radiance 2, 0.5 m up, facing down) over
    a water box (0.4 × 0.4 × 0.1 m, n = 1.333) holding an upward detector sensor (half-size 0.1,
    30 mm under the surface). The light reaching the sensor has refracted into the water: camera
    paths carry radiance (the n² factor at the surface) and particles carry flux (no factor), so
    the two integrators agree only if both transport modes are right."""
    sc = Scene()
    air = sc.add_medium(Medium.air())
    water = sc.add_medium(Medium.constant(n_water))
    amb = sc.add_region(air)
    wet = sc.add_region(water)
    black = sc.add_optics(Optics(SURF_ABSORBER))
    diel = sc.add_optics(Optics(SURF_DIELECTRIC))
    det = sc.add_optics(Optics(SURF_DETECTOR))
    e = sc.add_emission(Emission(Spectrum.constant(2.0)))
    sc.add_plane(APERTURE_DISK, 0.3, 0.0, transform(rotation((1, 0, 0), math.pi), (0.0, 0.0, 0.5)), amb, amb, black, e)
    pos, tri = box_mesh((-0.2, -0.2, -0.1), (0.2, 0.2, 0.0))
    sc.add_mesh(pos, tri, transform(), amb, wet, diel)
    sb = sc.add_plane(APERTURE_RECT, 0.1, 0.1, transform(None, (0.0, 0.0, -0.03)), wet, wet, det)
    sc.set_sensor(sb, size, size, 0.1, 0.1)
    return sc
