"""Bodies built from analytic shapes, assembled as the reference's scene builders assemble them, and
the reference's forward-light test scenes that use them.

* ``add_lens``: buildLens (src/owe/scene/builders.cpp:24-70) — a sag boundary per surface at its
  vertex, a flat annular step where a surface's semi-diameter is below the edge radius, and a
  cylindrical rim per element between the surfaces' edge heights;
* ``design_simple_lens``: designSimpleLens (builders.cpp:72-100), "bi" and "plano" forms;
* ``add_mirror``: buildMirror (builders.cpp:102-126) — the reflecting sag face over an opaque
  substrate, a black back disc and rim;
* ``add_water``: buildWater (builders.cpp:402-424) without a basin — the wave surface over the
  liquid, its floor and four walls facing the liquid;
* ``add_sensor``: a surface sensor (loader_detectors.cpp:224-257) placed as the loader places bodies;
* the scenes of tests/test_references.cpp: ``sun_lens_scene`` (77-120), ``sun_water_scene``
  (122-147), ``sun_mirror_scene`` (149-175), ``ripple_scene`` (322-346), with the power budgets
  those tests compute in ``lens_budget``, ``water_budget`` and ``mirror_budget``.

Units are metres.
"""
from __future__ import annotations

import functools
import math

import numpy as np

from . import scene as S

# Daimon & Masumura (2007) water at 20 °C, absorption after Pope & Fry (1997) (medium.cpp:136-142).
WATER = dict(B=(5.684027565e-1, 1.726177391e-1, 2.086189578e-2, 1.130748688e-1),
             C=(5.101829712e-3, 1.821153936e-2, 2.620722293e-2, 1.069792721e1))
WATER_ABSORPTION = ((380, 0.0114), (400, 0.00663), (450, 0.0092), (500, 0.0204), (550, 0.0565), (600, 0.2224),
                    (650, 0.34), (700, 0.65), (750, 2.47), (800, 2.07))
SUN_HALF_ANGLE = math.radians(0.533) / 2   # the loader's default angular diameter (scene_loader.cpp:98)


def water_medium() -> S.Medium:
    m = S.Medium.sellmeier(WATER["B"], WATER["C"])
    m.absorption = S.Spectrum.tabulated(*zip(*WATER_ABSORPTION))
    return m


def glass_medium(entry=S.N_BK7) -> S.Medium:
    return S.Medium.sellmeier(entry["B"], entry["C"], entry["rel_air"])


def compose(A, B) -> np.ndarray:
    """The rigid transform x ↦ A(B x) of two 3 × 4 [R | t] transforms."""
    A, B = np.asarray(A, float), np.asarray(B, float)
    M = np.zeros((3, 4))
    M[:, :3] = A[:, :3] @ B[:, :3]
    M[:, 3] = A[:, :3] @ B[:, 3] + A[:, 3]
    return M


def align_z(axis, up=(0.0, 1.0, 0.0)) -> np.ndarray:
    """Transform::alignZ (src/owe/core/math.hpp:123-132): local z along `axis`, y toward `up`."""
    z = np.asarray(axis, float)
    z = z / np.linalg.norm(z)
    u = np.asarray(up, float)
    if abs(np.dot(u / np.linalg.norm(u), z)) > 0.999:
        u = np.array([1.0, 0.0, 0.0]) if abs(z[0]) < 0.9 else np.array([0.0, 0.0, 1.0])
    x = np.cross(u, z)
    x /= np.linalg.norm(x)
    y = np.cross(z, x)
    return S.transform(np.stack([x, y, z], axis=1))


def placement(position=(0.0, 0.0, 0.0), axis=(0.0, 0.0, 1.0), up=(0.0, 1.0, 0.0)) -> np.ndarray:
    return compose(S.transform(None, position), align_z(axis, up))


def sag_of(curvature: float, conic: float, r: float, asphere=()) -> float:
    """z = c r² / (1 + √(1 − (1 + k) c² r²)) + Σ A₂ₘ r²ᵐ (m = 2, 3, …), the sag of SagSurface."""
    z = curvature * r * r / (1 + math.sqrt(max(0.0, 1 - (1 + conic) * curvature * curvature * r * r)))
    r2 = r * r
    rp = r2
    for a in asphere:
        rp *= r2
        z += a * rp
    return z


def design_simple_lens(f: float, n: float, t: float, form: str = "bi"):
    """Radii (R1, R2) of designSimpleLens for focal length f, relative index n, thickness t."""
    if form == "plano":
        return (n - 1) * f, math.inf
    disc = 1 - t / (n * f)
    if disc < 0:
        raise ValueError("lens too thick for the requested focal length")
    R = (n - 1) * f * (1 + math.sqrt(disc))
    return R, -R


def add_lens(sc: S.Scene, radii, semi_diameters, thicknesses, media, xf, ambient: int, rim_optics: int,
             dielectric: int | None = None, conics=None, aspheres=None, edge_radius: float = 0.0) -> list[int]:
    """buildLens: len(media) elements between len(media) + 1 surfaces (radii in metres, inf for
    flat; semi-diameters; centre thicknesses; optional conic constants and asphere terms A4, A6, … in
    SI units per surface); the body placed by xf, its vertex 0 at the origin and axis +z; the rim at
    the largest of the semi-diameters and edge_radius. Returns the boundaries added."""
    n = len(media)
    if len(radii) != n + 1 or len(semi_diameters) != n + 1 or len(thicknesses) != n:
        raise ValueError("a lens needs n media, n thicknesses and n + 1 surfaces")
    conics = [0.0] * (n + 1) if conics is None else list(conics)
    aspheres = [()] * (n + 1) if aspheres is None else [tuple(a) for a in aspheres]
    if dielectric is None:
        dielectric = sc.add_optics(S.Optics(S.SURF_DIELECTRIC))
    edge = max(max(semi_diameters), edge_radius)
    regions = [sc.add_region(sc.add_medium(m)) for m in media]
    zv = [0.0]
    for t in thicknesses:
        zv.append(zv[-1] + t)
    out, z_edge = [], []
    for i in range(n + 1):
        c = 0.0 if math.isinf(radii[i]) else 1.0 / radii[i]
        sd = semi_diameters[i]
        z_edge.append(zv[i] + sag_of(c, conics[i], sd, aspheres[i]))
        front = regions[i] if i < n else ambient
        back = regions[i - 1] if i > 0 else ambient
        out.append(sc.add_sag(c, conics[i], aspheres[i], sd, 0.0, compose(xf, S.transform(None, (0, 0, zv[i]))),
                              front, back, dielectric))
        if sd < edge:
            out.append(sc.add_plane(S.APERTURE_DISK, edge, sd, compose(xf, S.transform(None, (0, 0, z_edge[i]))),
                                    front, back, rim_optics))
    for i in range(n):
        if not z_edge[i + 1] > z_edge[i]:
            raise ValueError(f"non-positive edge thickness in element {i + 1}")
        out.append(sc.add_cylinder(edge, z_edge[i], z_edge[i + 1], xf, ambient, regions[i], rim_optics))
    return out


def add_mirror(sc: S.Scene, radius: float, conic: float, semi_diameter: float, thickness: float, optics: int, xf,
               ambient: int, hole: float = 0.0, asphere=()) -> list[int]:
    """buildMirror: the reflecting sag face (normal +z, the substrate behind it on −z), a black
    back disc and rim (and bore, with a central hole) around an opaque substrate."""
    sub = sc.add_region(sc.add_medium(S.Medium(S.INDEX_CONSTANT, (1.0,), opaque=True)))
    black = sc.add_optics(S.Optics(S.SURF_ABSORBER))
    c = 0.0 if math.isinf(radius) else 1.0 / radius
    out = [sc.add_sag(c, conic, tuple(asphere), semi_diameter, hole, xf, ambient, sub, optics)]
    z_edge = sag_of(c, conic, semi_diameter, asphere)
    z_back = min(0.0, z_edge) - thickness
    flip = S.transform(S.rotation((1, 0, 0), math.pi), (0, 0, z_back))
    out.append(sc.add_plane(S.APERTURE_DISK, semi_diameter, hole, compose(xf, flip), ambient, sub, black))
    out.append(sc.add_cylinder(semi_diameter, z_back, z_edge, xf, ambient, sub, black))
    if hole > 0:
        out.append(sc.add_cylinder(hole, z_back, sag_of(c, conic, hole, asphere), xf, sub, ambient, black))
    return out


def add_water(sc: S.Scene, waves, half_x: float, half_y: float, margin: float, depth: float, medium: S.Medium, xf,
              ambient: int, spacing: float = -1.0) -> tuple[int, list[int]]:
    """buildWater without a basin: the liquid under the wave surface, its floor and walls
    dielectric and facing the liquid. Returns (liquid region, boundaries)."""
    liquid = sc.add_region(sc.add_medium(medium))
    diel = sc.add_optics(S.Optics(S.SURF_DIELECTRIC))
    out = [sc.add_wave(waves, half_x, half_y, margin, xf, ambient, liquid, diel, spacing)]
    out.append(sc.add_plane(S.APERTURE_RECT, half_x, half_y, compose(xf, S.transform(None, (0, 0, -depth))),
                            liquid, ambient, diel))
    hz = depth / 2
    for at, inward, half in (((half_x, 0, -hz), (-1, 0, 0), half_y), ((-half_x, 0, -hz), (1, 0, 0), half_y),
                             ((0, half_y, -hz), (0, -1, 0), half_x), ((0, -half_y, -hz), (0, 1, 0), half_x)):
        wall = compose(S.transform(None, at), align_z(inward, (0, 0, 1)))
        out.append(sc.add_plane(S.APERTURE_RECT, half, hz, compose(xf, wall), liquid, ambient, diel))
    return liquid, out


def add_sensor(sc: S.Scene, size, resolution, position, axis=(0.0, 0.0, 1.0), up=(0.0, 1.0, 0.0),
               region: int = 0) -> int:
    """A detector plane of `size` (metres) and `resolution` pixels, measuring irradiance on its
    front (+axis) face; it becomes the scene's detector."""
    det = sc.add_optics(S.Optics(S.SURF_DETECTOR))
    b = sc.add_plane(S.APERTURE_RECT, size[0] / 2, size[1] / 2, placement(position, axis, up), region, region, det)
    sc.set_sensor(b, resolution[0], resolution[1], size[0] / 2, size[1] / 2)
    return b


def sun_direction(elevation: float, azimuth: float = 0.0) -> tuple:
    """directionFromAzimuthElevation (src/owe/core/math.hpp:9): azimuth from +y toward +x."""
    return (math.cos(elevation) * math.sin(azimuth), math.cos(elevation) * math.cos(azimuth), math.sin(elevation))


def _sun(sc: S.Scene, elevation: float, azimuth: float = 0.0):
    sc.set_sun(sun_direction(elevation, azimuth), SUN_HALF_ANGLE, S.Spectrum.constant(1.0))


def sun_irradiance() -> float:
    """A sun of unit spectral radiance on a plane facing it: π sin²θ_s."""
    return math.pi * math.sin(SUN_HALF_ANGLE) ** 2


# ---------------------------------------------------------------- the reference's scenes


def sun_lens_scene(resolution: int = 48) -> S.Scene:
    """An overhead sun, an f = 100 mm biconvex N-BK7 lens (40 mm, 8 mm thick, black rim) 100 mm up,
    a 24 × 24 mm sensor 1.5 mm below the origin facing up (test_references.cpp:77-88)."""
    sc = S.Scene()
    air = sc.add_region(sc.add_medium(S.Medium.air()))
    _sun(sc, math.radians(90))
    n = relative_index(S.N_BK7, 587.56)
    R1, R2 = design_simple_lens(0.100, n, 0.008)
    black = sc.add_optics(S.Optics(S.SURF_ABSORBER))
    add_lens(sc, (R1, R2), (0.020, 0.020), (0.008,), [glass_medium()], placement((0, 0, 0.100)), air, black)
    add_sensor(sc, (0.024, 0.024), (resolution, resolution), (0, 0, -0.0015), region=air)
    return sc


def sun_water_scene(resolution: int = 16) -> S.Scene:
    """A sun at 40° elevation over a 1.2 × 1.2 × 0.1 m water box (its centre 0.3 m up), a 0.2 m
    sensor on the ground facing up (test_references.cpp:122-131)."""
    sc = S.Scene()
    air = sc.add_region(sc.add_medium(S.Medium.air()))
    _sun(sc, math.radians(40))
    water = sc.add_region(sc.add_medium(water_medium()))
    diel = sc.add_optics(S.Optics(S.SURF_DIELECTRIC))
    pos, tri = S.box_mesh((-0.6, -0.6, 0.25), (0.6, 0.6, 0.35))
    sc.add_mesh(pos, tri, S.transform(), air, water, diel)
    add_sensor(sc, (0.2, 0.2), (resolution, resolution), (0, 0, 0), region=air)
    return sc


def sun_mirror_scene(resolution: int = 20) -> S.Scene:
    """An overhead sun, a spherical mirror (R = 400 mm, 100 mm aperture, reflectance 0.9) at the
    origin facing up, a 10 mm sensor at 199 mm facing down (test_references.cpp:149-160)."""
    sc = S.Scene()
    air = sc.add_region(sc.add_medium(S.Medium.air()))
    _sun(sc, math.radians(90))
    coating = sc.add_optics(S.Optics(S.SURF_MIRROR, reflectance=S.Spectrum.constant(0.9)))
    add_mirror(sc, 0.400, 0.0, 0.050, 0.0, coating, placement(), air)
    add_sensor(sc, (0.010, 0.010), (resolution, resolution), (0, 0, 0.199), axis=(0, 0, -1), region=air)
    return sc


def ripple_scene(a: float = 0.0005, wavelength: float = 0.02, depth: float = 0.08, gap: float = 0.001,
                 columns: int = 160, rows: int = 4, spacing: float = -1.0) -> S.Scene:
    """Sunlight through a rippled water surface h = a sin(kx) (0.12 × 0.12 m pool, 3 cm margin) onto
    a sensor `gap` below its floor (test_references.cpp:322-346)."""
    sc = S.Scene()
    air = sc.add_region(sc.add_medium(S.Medium.air()))
    _sun(sc, math.radians(90))
    k = 2 * math.pi / wavelength
    add_water(sc, [(a, k, 0.0, 0.0)], 0.06, 0.06, 0.03, depth, water_medium(), placement(), air, spacing)
    add_sensor(sc, (0.04, 0.02), (columns, rows), (0, 0, -depth - gap), region=air)
    return sc


# ---------------------------------------------------------------- budgets


def cie_y(lam: float) -> float:
    return S._cie_y(lam)


def y_weighted(f) -> float:
    """∫ f ȳ dλ / ∫ ȳ dλ over 360–830 nm in 0.5 nm steps (test_references.cpp:38-46)."""
    num = den = 0.0
    lam = 360.0
    while lam <= 830.0:
        w = cie_y(lam)
        num += f(lam) * w
        den += w
        lam += 0.5
    return num / den


def fresnel(n1: float, n2: float, ci: float) -> float:
    si2 = 1 - ci * ci
    st2 = (n1 / n2) ** 2 * si2
    if st2 >= 1:
        return 1.0
    ct = math.sqrt(1 - st2)
    rs = (n1 * ci - n2 * ct) / (n1 * ci + n2 * ct)
    rp = (n2 * ci - n1 * ct) / (n2 * ci + n1 * ct)
    return 0.5 * (rs * rs + rp * rp)


def ciddor(lam: float) -> float:
    s2 = (1e3 / lam) ** 2
    return 1 + 0.05792105 / (238.0185 - s2) + 0.00167917 / (57.362 - s2)


def sellmeier_index(entry, lam: float) -> float:
    l2 = (lam * 1e-3) ** 2
    n = math.sqrt(1 + sum(b * l2 / (l2 - c) for b, c in zip(entry["B"], entry["C"])))
    return n * ciddor(lam) if entry.get("rel_air") else n


def relative_index(entry, lam: float) -> float:
    return sellmeier_index(entry, lam) / ciddor(lam)


def absorption(lam: float) -> float:
    xs, ys = zip(*WATER_ABSORPTION)
    return float(np.interp(lam, xs, ys))


def lens_budget() -> float:
    """P = E_sun π (20 mm)² ⟨(1 − R₁)(1 − R₂)⟩ over the aperture (test_references.cpp:90-117)."""
    t, f = 8.0, 100.0

    def transmitted(lam):
        n = relative_index(S.N_BK7, lam)
        nd = relative_index(S.N_BK7, 587.56)
        a, b = 2 * (nd - 1), -((nd - 1) ** 2) * t / nd
        R = (a * f + math.sqrt((a * f) ** 2 + 4 * b * f)) / 2
        s = w = 0.0
        for i in range(400):
            h = 20.0 * (i + 0.5) / 400
            c1 = math.sqrt(1 - (h / R) ** 2)
            s2 = math.sqrt(1 - c1 * c1) / n
            th1, th2 = math.asin(math.sqrt(1 - c1 * c1)), math.asin(s2)
            tilt = th1 - th2
            h2 = h - t * math.tan(tilt)
            c3 = math.cos(math.asin(h2 / R) + tilt)
            s += h * (1 - fresnel(1, n, c1)) * (1 - fresnel(n, 1, c3))
            w += h
        return s / w

    return sun_irradiance() * math.pi * 0.020 ** 2 * y_weighted(transmitted)


def water_budget() -> float:
    """E = E_sun cos θᵢ T(θᵢ) T(θₜ) exp(−σₐ d / cos θₜ) at 50° incidence (test_references.cpp:132-141)."""
    ci = math.cos(math.radians(50))

    def f(lam):
        n = sellmeier_index(WATER, lam) / ciddor(lam)
        ct = math.sqrt(1 - (1 - ci * ci) / (n * n))
        return ci * (1 - fresnel(1, n, ci)) * (1 - fresnel(n, 1, ct)) * math.exp(-absorption(lam) * 0.1 / ct)

    return sun_irradiance() * y_weighted(f)


def mirror_budget() -> float:
    """P = E_sun (π (50 mm)² − (10 mm)²) · 0.9 (test_references.cpp:162)."""
    return sun_irradiance() * (math.pi * 0.050 ** 2 - 0.010 ** 2) * 0.9


def ripple_analytic(a: float = 0.0005, wavelength: float = 0.02, depth: float = 0.08, gap: float = 0.001,
                    n: int = 4_000_000, seed: int = 99, bins: int = 160) -> np.ndarray:
    """The ripple test's independent estimate (test_references.cpp:348-385): sunlight refracted
    analytically at h = a sin(kx) and the flat floor, Fresnel-weighted and absorbed along the way,
    histogrammed over the sensor's columns; irradiance per column in units of the sun's radiance."""
    rng = np.random.default_rng(seed)
    k = 2 * math.pi / wavelength
    X, Y = 0.03, 0.012
    lams = np.arange(360.0, 830.0 + 1e-9, 0.5)
    cdf = np.cumsum([cie_y(l) for l in lams])
    cos_s = math.cos(SUN_HALF_ANGLE)
    cone = 2 * math.pi * (1 - cos_s)
    x = (2 * rng.random(n) - 1) * X
    y = (2 * rng.random(n) - 1) * Y
    lam = lams[np.minimum(np.searchsorted(cdf, rng.random(n) * cdf[-1], side="right"), len(lams) - 1)]
    u1, u2 = rng.random(n), rng.random(n)
    cz = 1 - u1 * (1 - cos_s)
    sz = np.sqrt(np.maximum(0.0, 1 - cz * cz))
    d0 = np.stack([sz * np.cos(2 * math.pi * u2), sz * np.sin(2 * math.pi * u2), -cz], axis=1)
    nw = np.array([sellmeier_index(WATER, l) for l in lams]) / np.array([ciddor(l) for l in lams])
    eta = np.interp(lam, lams, nw)
    h = a * np.sin(k * x)
    nrm = np.stack([-a * k * np.cos(k * x), np.zeros(n), np.ones(n)], axis=1)
    nrm /= np.linalg.norm(nrm, axis=1)[:, None]

    def refract(d, nn, n1, n2):
        flip = np.sum(d * nn, axis=1) > 0
        nn = np.where(flip[:, None], -nn, nn)
        ci = -np.sum(d * nn, axis=1)
        e = n1 / n2
        kk = 1 - e * e * (1 - ci * ci)
        ok = kk >= 0
        out = d * e[:, None] + nn * (e * ci - np.sqrt(np.maximum(kk, 0)))[:, None]
        out /= np.linalg.norm(out, axis=1)[:, None]
        return out, ok

    one = np.ones(n)
    d1, ok1 = refract(d0, nrm, one, eta)
    d2, ok2 = refract(d1, np.tile([0.0, 0.0, 1.0], (n, 1)), eta, one)
    c1 = np.abs(np.sum(d0 * nrm, axis=1))
    c2 = np.abs(d1[:, 2])
    t1 = (-depth - h) / d1[:, 2]
    qx, qy = x + d1[:, 0] * t1, y + d1[:, 1] * t1
    t2 = -gap / d2[:, 2]
    u, v = qx + d2[:, 0] * t2, qy + d2[:, 1] * t2
    keep = ok1 & ok2 & (np.abs(v) < 0.01) & (np.abs(u) < 0.02)
    fres = np.vectorize(fresnel)
    R1 = fres(np.ones(keep.sum()), eta[keep], c1[keep])
    R2 = fres(eta[keep], np.ones(keep.sum()), c2[keep])
    sig = np.interp(lam[keep], *zip(*WATER_ABSORPTION))
    w = np.abs(d0[keep, 2]) * cone * (2 * X) * (2 * Y) / n * (1 - R1) * (1 - R2) * np.exp(-sig * t1[keep])
    col = ((u[keep] + 0.02) / 0.04 * bins).astype(int)
    hist = np.bincount(np.clip(col, 0, bins - 1), weights=w, minlength=bins)
    return hist / ((0.04 / bins) * 0.02)


def uniform_sky_optics_scene(kind: str = "lens", size: int = 32) -> S.Scene:
    """Lossless dielectric optics under a uniform sky of unit radiance, seen by a camera: a thick
    biconvex lens with polished (dielectric) rims and an annular step, a rippled water pool
    (n = 1.333, no absorption) with a sampled or analytic wave surface, or a concave dielectric
    meniscus. Fresnel reflection and refraction only redistribute a uniform radiance field, so every
    pixel sees radiance 1 (the invisible-sphere argument of tests/test_transport.cpp:116-142)."""
    sc = S.Scene()
    air = sc.add_region(sc.add_medium(S.Medium.air()))
    sc.set_uniform_sky(S.Spectrum.constant(1.0))
    diel = sc.add_optics(S.Optics(S.SURF_DIELECTRIC))
    if kind == "lens":
        n = relative_index(S.N_BK7, 587.56)
        R1, R2 = design_simple_lens(0.05, n, 0.012)
        add_lens(sc, (R1, R2), (0.015, 0.012), (0.012,), [glass_medium()], placement((0, 0, -0.006)), air, diel,
                 dielectric=diel)
        sc.set_observer(size, size, (0.02, 0.01, 0.12), (0, 0, 0), up=(0, 1, 0), fov_y=math.radians(20), region=air)
    elif kind in ("waves", "waves-sampled"):
        waves = [(0.002, 2 * math.pi / 0.03, 0.0, 0.0), (0.001, 60.0, 120.0, 0.7)]
        add_water(sc, waves, 0.05, 0.05, 0.01, 0.03, S.Medium.constant(1.333), placement(), air,
                  spacing=0.001 if kind == "waves-sampled" else 0.0)
        sc.set_observer(size, size, (0.03, -0.05, 0.12), (0, 0, -0.01), up=(0, 0, 1), fov_y=math.radians(45),
                        region=air)
    else:
        # An aspheric meniscus: a concave conic face over a convex one, with a cylindrical rim.
        glass = sc.add_region(sc.add_medium(S.Medium.constant(1.6)))
        sc.add_sag(1 / 0.05, 0.0, (), 0.02, 0.0, S.transform(), glass, air, diel)
        sc.add_sag(1 / 0.08, -0.5, (0.0, 2.0e3), 0.02, 0.0, S.transform(None, (0, 0, 0.004)), air, glass, diel)
        z0, z1 = sag_of(1 / 0.05, 0.0, 0.02), 0.004 + sag_of(1 / 0.08, -0.5, 0.02) + 2.0e3 * 0.02 ** 6
        sc.add_cylinder(0.02, z0, z1, S.transform(), air, glass, diel)
        sc.set_observer(size, size, (0.0, 0.03, 0.1), (0, 0, 0), up=(0, 1, 0), fov_y=math.radians(30), region=air)
    return sc


# ---------------------------------------------------------------- instruments as matter


def add_prescription(sc: S.Scene, p, xf, ambient: int, mirror_optics: int, rim_optics: int, tube_radius: float = 0.0,
                     media_of=None) -> dict:
    """buildPrescription (src/owe/scene/prescription.cpp:354-432): a lens prescription (owe_hip.lens,
    millimetres) as matter placed by xf — each mirror row a mirror body (its frame turned over when
    met travelling +z, so the face meets the light; substrate, rim, bore, spider struts behind it),
    each group of glass rows a lens with rims (and, inside a tube, a black mount from its rim to the
    tube), each stop row in air a black annulus. ``mirror_optics`` is an optics index, or a function
    of a mirror row's material name. Returns the vertex positions (metres, along the body's z), the
    index of the last vertex, and the rear group's nearest z and edge radius."""
    from . import lens as Lm

    if media_of is None:
        def media_of(name):
            if Lm.CATALOG[name] is None:
                return S.Medium.air()
            B, C = Lm.CATALOG[name]
            if B == "cauchy":
                return S.Medium.cauchy(C[0], C[1])
            return S.Medium.sellmeier(B, C, True)

    rows = p.surfaces
    rear_nearest, rear_edge = 0.0, 0.0
    z = [0.0]
    for r in rows[:-1]:
        z.append(z[-1] + r.t * 1e-3)
    black = sc.add_optics(S.Optics(S.SURF_ABSORBER))
    direction, i = 1.0, 0
    while i < len(rows):
        r = rows[i]
        before = p.ambient if i == 0 else rows[i - 1].medium
        if r.reflect:
            R = (-r.R if direction > 0 else r.R) * 1e-3
            at = S.transform(None, (0, 0, z[i]))
            if direction > 0:
                at = compose(at, S.transform(S.rotation((1, 0, 0), math.pi)))
            substrate = getattr(r, "substrate", 0.0) or 4.0
            optics = mirror_optics(getattr(r, "material", "")) if callable(mirror_optics) else mirror_optics
            A = tuple((-a if direction > 0 else a) for a in r.A)
            add_mirror(sc, R, r.k, r.sd * 1e-3, substrate * 1e-3, optics, compose(xf, at), ambient,
                       hole=r.hole * 1e-3, asphere=A)
            for k in range(getattr(r, "spider_count", 0)):
                a = getattr(r, "spider_angle", 0.0) + 2 * math.pi * k / r.spider_count
                outer = r.spider_outer * 1e-3
                z_back = min(0.0, sag_of(1 / R, r.k, r.sd * 1e-3)) - substrate * 1e-3
                strut = compose(at, compose(S.transform(S.rotation((0, 0, 1), a)),
                                            S.transform(None, (0.5 * outer, 0, z_back - 0.5 * substrate * 1e-3 - 1e-6))))
                pos, tri = S.box_mesh((-0.5 * outer, -0.5 * r.spider_width * 1e-3, -0.5 * substrate * 1e-3),
                                      (0.5 * outer, 0.5 * r.spider_width * 1e-3, 0.5 * substrate * 1e-3))
                inside = sc.add_region(sc.add_medium(S.Medium(S.INDEX_CONSTANT, (1.0,), opaque=True)))
                sc.add_mesh(pos, tri, compose(xf, strut), ambient, inside, black)
            direction = -direction
            i += 1
            continue
        if before == p.ambient and r.medium == p.ambient:
            if r.stop and r.sd > 0:
                outer = tube_radius if tube_radius > r.sd * 1e-3 else 3 * r.sd * 1e-3
                sc.add_plane(S.APERTURE_DISK, outer, r.sd * 1e-3, compose(xf, S.transform(None, (0, 0, z[i]))),
                             ambient, ambient, black)
            i += 1
            continue
        j = i
        while j < len(rows) and rows[j].medium != p.ambient:
            j += 1
        if j >= len(rows):
            raise ValueError("prescription: last surface must exit into the ambient medium")
        group = rows[i:j + 1]
        radii = [(math.inf if math.isinf(q.R) else q.R * 1e-3) for q in group]
        sds = [q.sd * 1e-3 for q in group]
        add_lens(sc, radii, sds, [q.t * 1e-3 for q in rows[i:j]], [media_of(q.medium) for q in rows[i:j]],
                 compose(xf, S.transform(None, (0, 0, z[i]))), ambient, rim_optics,
                 conics=[q.k for q in group], aspheres=[q.A for q in group])
        edge = max(sds)
        c0 = 0.0 if math.isinf(radii[0]) else 1 / radii[0]
        c1 = 0.0 if math.isinf(radii[-1]) else 1 / radii[-1]
        z_first = z[i] + sag_of(c0, group[0].k, sds[0], group[0].A)
        z_last = z[j] + sag_of(c1, group[-1].k, sds[-1], group[-1].A)
        rear_nearest, rear_edge = max(z[j], z_last, z[i], z_first), edge
        if tube_radius > edge:
            # The mount: an annulus from just outside the rim (1e-9 relative, so the two bodies share
            # no edge) to the tube, halfway between the group's edges.
            sc.add_plane(S.APERTURE_DISK, tube_radius, edge * (1 + 1e-9),
                         compose(xf, S.transform(None, (0, 0, 0.5 * (z_first + z_last)))), ambient, ambient, black)
        i = j + 1
    return dict(vertex_z=z, last=len(rows) - 1, rear_nearest_z=rear_nearest, rear_edge_radius=rear_edge)


@functools.lru_cache(maxsize=2)
def _statue(frequency: int, seed: int):
    """The room's statue mesh (a displaced icosphere), generated once per (frequency, seed)."""
    from .workloads import icosphere

    return icosphere(frequency, seed=seed, amplitude=0.05, frequency=3.0)


def telescope_room_scene(p, size: int = 512, statue_frequency: int = 317, fov_deg: float = 50.0,
                         seed: int = 1) -> S.Scene:
    """The differentiable-telescope scene: telescope p (owe_hip.lens, solved afocal) at the origin
    looking along −z into a closed room — a diffuse floor and walls, a rough aluminium wall, an
    emitting ceiling panel, and a displaced icosphere statue of 20 · statue_frequency² triangles
    (2,009,780 at 317) in rough gold 5 m in front — seen by an eye at the ocular's eye point (the
    last row's gap) through a pinhole, apparent field fov_deg."""
    sc = S.Scene()
    air = sc.add_region(sc.add_medium(S.Medium.air()))
    mirror = sc.add_optics(S.Optics(S.SURF_CONDUCTOR, conductor_n=S.Spectrum.constant(1.2),
                                    conductor_k=S.Spectrum.constant(4.0)))
    black = sc.add_optics(S.Optics(S.SURF_ABSORBER))
    info = add_prescription(sc, p, S.transform(), air, mirror, black)
    # The room: x in [−2, 2], y in [−1, 2], z in [−8, 1].
    floor = sc.add_optics(S.Optics(S.SURF_DIFFUSE, reflectance=S.Spectrum.constant(0.5)))
    wall = sc.add_optics(S.Optics(S.SURF_DIFFUSE, reflectance=S.Spectrum.constant(0.7)))
    alu = sc.add_optics(S.Optics(S.SURF_CONDUCTOR, roughness=0.15, conductor_n=S.Spectrum.constant(1.0),
                                 conductor_k=S.Spectrum.constant(6.5)))
    gold = sc.add_optics(S.Optics(S.SURF_CONDUCTOR, roughness=0.2, conductor_n=S.Spectrum.constant(0.3),
                                  conductor_k=S.Spectrum.constant(3.0)))
    light = sc.add_emission(S.Emission(S.Spectrum.constant(8.0)))
    sc.add_plane(S.APERTURE_RECT, 2.0, 4.5, placement((0, -1, -3.5), (0, 1, 0), (0, 0, 1)), air, air, floor)
    sc.add_plane(S.APERTURE_RECT, 2.0, 4.5, placement((0, 2, -3.5), (0, -1, 0), (0, 0, 1)), air, air, wall)
    sc.add_plane(S.APERTURE_RECT, 1.0, 1.0, placement((0, 1.99, -4.0), (0, -1, 0), (0, 0, 1)), air, air, black, light)
    sc.add_plane(S.APERTURE_RECT, 2.0, 1.5, placement((0, 0.5, -8.0), (0, 0, 1), (0, 1, 0)), air, air, alu)
    sc.add_plane(S.APERTURE_RECT, 2.0, 1.5, placement((0, 0.5, 1.0), (0, 0, -1), (0, 1, 0)), air, air, wall)
    sc.add_plane(S.APERTURE_RECT, 4.5, 1.5, placement((-2, 0.5, -3.5), (1, 0, 0), (0, 1, 0)), air, air, wall)
    sc.add_plane(S.APERTURE_RECT, 4.5, 1.5, placement((2, 0.5, -3.5), (-1, 0, 0), (0, 1, 0)), air, air, wall)
    pos, tri = _statue(statue_frequency, seed)
    inside = sc.add_region(sc.add_medium(S.Medium(S.INDEX_CONSTANT, (1.0,), opaque=True)))
    sc.add_mesh(np.asarray(pos, np.float64) * 0.4, tri, S.transform(None, (0.0, -0.2, -5.0)), air, inside, gold)
    eye = info["vertex_z"][info["last"]] + p.surfaces[-1].t * 1e-3
    sc.set_observer(size, size, (0.0, 0.0, eye), (0.0, 0.0, eye - 1.0), up=(0, 1, 0), fov_y=math.radians(fov_deg),
                    region=air)
    return sc
