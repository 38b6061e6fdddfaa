"""Sequential lenses for the batched lens analysis (include/owe_hip_lens.h).

* ``CATALOG``: the reference catalog's glasses (Schott Sellmeier, relative to air) and Ciddor air
  (src/owe/core/medium.cpp:109-127);
* ``Surface``/``Prescription``: rows of a lens table in millimetres, as the reference's ``.lens``
  text (src/owe/scene/prescription.hpp:16-44), converted to metres for the C ABI;
* ``SINGLET``/``DOUBLET``: the reference's own test prescriptions (tests/test_instruments.cpp:209-221);
* ``seeded_achromats``: cemented achromat doublets from a fixed seed — focal length U(50, 150) mm,
  front radius ±U(30, 120) mm, element thicknesses U(2, 8) mm, a crown and a flint from the catalog,
  the inner and rear radii from the thin-lens achromat condition at the F, d and C lines;
* ``LensBatch``: the persistent GPU analysis over the C ABI;
* ``spot``: focusedSpot (src/owe/analysis/analysis.cpp:85-113) from the nine sums a run returns.
"""
from __future__ import annotations

import ctypes
import dataclasses
import math
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from .scene import CMedium, INDEX_CAUCHY, INDEX_CIDDOR, INDEX_SELLMEIER

LAMBDA_F, LAMBDA_D, LAMBDA_C = 486.13, 587.56, 656.27
# Fraunhofer h, g, F, e, d, C, r lines (nm): the seven wavelengths of the batched analysis.
FRAUNHOFER7 = (404.66, 435.83, 486.13, 546.07, 587.56, 656.27, 706.52)

CATALOG = {
    "air": None,
    "N-BK7": ((1.03961212, 0.231792344, 1.01046945), (0.00600069867, 0.0200179144, 103.560653)),
    "N-SK16": ((1.34317774, 0.241144399, 0.994317969), (0.00704687339, 0.0229005, 92.7508526)),
    "N-BAF10": ((1.5851495, 0.143559385, 1.08521269), (0.00926681282, 0.0424489805, 105.613573)),
    "F2": ((1.34533359, 0.209073176, 0.937357162), (0.00997743871, 0.0470450767, 111.886764)),
    "N-F2": ((1.39757037, 0.159201403, 1.2686543), (0.00995906143, 0.0546931752, 119.248346)),
    "N-SF6": ((1.77931763, 0.338149866, 2.08734474), (0.0133714182, 0.0617533621, 174.01759)),
    "N-SF11": ((1.73759695, 0.313747346, 1.89878101), (0.013188707, 0.0623068142, 155.23629)),
    "crown": ("cauchy", (1.5046, 0.00420)),   # IndexModel::cauchy, absolute (medium.cpp:145-147)
}
CROWNS = ("N-BK7", "N-SK16", "N-BAF10")
FLINTS = ("F2", "N-F2", "N-SF6", "N-SF11")
MEDIA_ORDER = tuple(CATALOG)


def media_array():
    """The catalog as owe_medium records, in MEDIA_ORDER (index 0 is air)."""
    arr = (CMedium * len(MEDIA_ORDER))()
    for i, name in enumerate(MEDIA_ORDER):
        m = arr[i]
        if CATALOG[name] is None:
            m.index_kind = INDEX_CIDDOR
            continue
        B, C = CATALOG[name]
        if B == "cauchy":
            m.index_kind, m.rel_air = INDEX_CAUCHY, 0
            for k, v in enumerate(C):
                m.coeff_a[k] = v
            continue
        m.index_kind, m.terms, m.rel_air = INDEX_SELLMEIER, len(B), 1
        for k in range(len(B)):
            m.coeff_a[k], m.coeff_b[k] = B[k], C[k]
    return arr


def ciddor_air(lnm: float) -> float:
    s2 = (1e3 / lnm) ** 2
    return 1 + 0.05792105 / (238.0185 - s2) + 0.00167917 / (57.362 - s2)


def index(name: str, lnm: float) -> float:
    """Absolute refractive index (catalog glasses are relative to air)."""
    if CATALOG[name] is None:
        return ciddor_air(lnm)
    B, C = CATALOG[name]
    if B == "cauchy":
        l2 = (lnm * 1e-3) ** 2
        return C[0] + C[1] / l2
    l2 = (lnm * 1e-3) ** 2
    n2 = 1 + sum(b * l2 / (l2 - c) for b, c in zip(B, C))
    return math.sqrt(n2) * ciddor_air(lnm)


@dataclass
class Surface:
    R: float                  # mm, math.inf for flat
    t: float                  # mm to the next vertex, signed along the fixed chart
    medium: str               # after the surface
    sd: float                 # semi-diameter, mm
    k: float = 0.0
    A: tuple = ()             # A4, A6, ... in SI units
    stop: bool = False
    reflect: bool = False     # a mirror row
    hole: float = 0.0         # mirror: central hole radius, mm
    solve: bool = False       # the gap solve_afocal adjusts
    material: str = ""        # mirror: the scene material of its face
    substrate: float = 0.0    # mirror: axial depth of its backing, mm
    spider_count: int = 0     # mirror: opaque struts behind it
    spider_width: float = 0.0
    spider_angle: float = 0.0  # radians
    spider_outer: float = 0.0
    label: str = ""


@dataclass
class Prescription:
    surfaces: list
    ambient: str = "air"
    rays: int = 4096
    name: str = ""
    rings: int = 0            # > 0: spotAt's hexapolar pupil rings (then rays = 1 + 3 r (r + 1))
    wavelength: float = 587.5618   # design wavelength for paraxial data, nm
    object_distance: float = math.inf  # mm

    def with_rings(self, rings: int) -> "Prescription":
        return dataclasses.replace(self, rays=1 + 3 * rings * (rings + 1), rings=rings)


SINGLET = Prescription([Surface(50.0, 8.0, "N-BK7", 12.5), Surface(-50.0, 40.0, "air", 12.5)], name="biconvex singlet")
DOUBLET = Prescription([Surface(44.78, 6.0, "N-BK7", 12.5), Surface(-44.78, 2.5, "F2", 12.5),
                        Surface(-810.8, 95.0, "air", 12.5)], name="cemented achromat f=100")
# The reference's Gregorian telescope (lenses/gregorian_54x.lens): a paraboloidal primary with a
# 26 mm hole, an ellipsoidal secondary, a Huygenian ocular of two plano-convex crown lenses ("crown"
# is the catalog's Cauchy glass, medium.cpp:145-150). The secondary's spider lies behind it.
GREGORIAN = Prescription([
    Surface(-900.0, -600.0, "air", 54.0, k=-1.0, stop=True, reflect=True, hole=13.0),
    Surface(247.368421053, 667.255310082, "air", 21.0, k=-0.42136041859, reflect=True, solve=True),
    Surface(41.3076651664, 3.0, "crown", 13.0), Surface(math.inf, 50.3333333333, "air", 13.0),
    Surface(13.7692217221, 2.5, "crown", 6.0), Surface(math.inf, 14.7400784369, "air", 6.0)],
    name="Gregorian 54x")


class CLensSurface(ctypes.Structure):
    _fields_ = [("curvature", ctypes.c_double), ("conic", ctypes.c_double), ("asphere", ctypes.c_double * 4),
                ("thickness", ctypes.c_double), ("semi_diameter", ctypes.c_double), ("medium", ctypes.c_int32),
                ("stop", ctypes.c_int32), ("reflect", ctypes.c_int32), ("reserved", ctypes.c_int32),
                ("hole_radius", ctypes.c_double)]


class CLens(ctypes.Structure):
    _fields_ = [("first_surface", ctypes.c_uint32), ("surface_count", ctypes.c_uint32), ("rays", ctypes.c_uint64),
                ("ambient", ctypes.c_int32), ("pupil_rings", ctypes.c_int32)]


class CLensSettings(ctypes.Structure):
    _fields_ = [("field_deg", ctypes.c_double), ("wavelengths", ctypes.c_double * 16), ("n_wavelengths", ctypes.c_int32),
                ("detector_pixels", ctypes.c_int32), ("detector_half_size", ctypes.c_double),
                ("reference_wavelength", ctypes.c_double), ("focus_planes", ctypes.c_int32),
                ("reserved", ctypes.c_int32), ("focus_step", ctypes.c_double)]


class CLensResult(ctypes.Structure):
    _fields_ = [("launched", ctypes.c_uint64), ("arrived", ctypes.c_uint64), ("efl", ctypes.c_double),
                ("bfl", ctypes.c_double), ("entrance_pupil_z", ctypes.c_double),
                ("entrance_pupil_radius", ctypes.c_double), ("last_vertex_z", ctypes.c_double),
                ("best_focus_z", ctypes.c_double), ("rms_best", ctypes.c_double), ("rms_paraxial", ctypes.c_double),
                ("centroid_y", ctypes.c_double), ("transmitted", ctypes.c_double), ("afocal", ctypes.c_int32),
                ("reserved", ctypes.c_int32), ("angular_magnification", ctypes.c_double),
                ("apparent_angle", ctypes.c_double), ("angular_rms", ctypes.c_double)]


class CLensParam(ctypes.Structure):
    _fields_ = [("surface", ctypes.c_uint32), ("field", ctypes.c_int32)]


PARAM_CURVATURE, PARAM_CONIC, PARAM_THICKNESS, PARAM_A4 = 0, 1, 2, 3


MOMENTS = 9


_LENS_UNITS = {"m": 1e3, "cm": 10.0, "mm": 1.0, "um": 1e-3, "nm": 1e-6, "km": 1e6, "in": 25.4}


def _lens_number(tok: str, ctx: str) -> float:
    if tok in ("inf", "infinity", "INF", "Infinity"):
        return math.inf
    if tok == "-inf":
        return -math.inf
    try:
        return float(tok)
    except ValueError:
        raise ValueError(f"{ctx}: expected a number, got '{tok}'") from None


def _lens_option(x: str, ctx: str, unit: float, s: Surface) -> None:
    if x in ("stop", "reflect", "solve"):
        setattr(s, x, True)
        return
    if "=" not in x:
        raise ValueError(f"{ctx}: unexpected token '{x}'")
    k, v = x.split("=", 1)
    if k == "material":
        s.material = v
    elif k == "out_axis":
        raise ValueError(f"{ctx}: out_axis (folded, non-coaxial mirrors) is not supported")
    elif k == "support":
        if not (v.startswith("spider(") and v.endswith(")")):
            raise ValueError(f"{ctx}: support must be spider(...)")
        for kv in v[7:-1].split(","):
            if "=" not in kv:
                raise ValueError(f"{ctx}: spider takes name=value pairs")
            n, val = kv.split("=", 1)
            val = _lens_number(val, ctx)
            if n == "count":
                s.spider_count = int(val)
            elif n == "width":
                s.spider_width = val * unit
            elif n == "angle":
                s.spider_angle = math.radians(val)
            elif n == "outer_radius":
                s.spider_outer = val * unit
            else:
                raise ValueError(f"{ctx}: unknown spider parameter '{n}'")
        if s.spider_count < 1 or not s.spider_width > 0 or not s.spider_outer > 0:
            raise ValueError(f"{ctx}: spider needs count >= 1, width > 0 and outer_radius > 0")
    else:
        val = _lens_number(v, ctx)
        if k == "substrate":
            s.substrate = val * unit
        elif k == "hole":
            s.hole = val * unit / 2       # the file gives the bore's diameter
        elif k == "k":
            s.k = val
        elif len(k) > 1 and k[0] == "A" and k[1:].isdigit():
            order = int(k[1:])
            if order < 4 or order % 2:
                raise ValueError(f"{ctx}: aspheric terms are A4, A6, A8, ...")
            slot = (order - 4) // 2
            A = list(s.A) + [0.0] * max(0, slot + 1 - len(s.A))
            A[slot] = val * (unit * 1e-3) ** (1 - order)   # coefficient of r^order in SI units
            s.A = tuple(A)
        else:
            raise ValueError(f"{ctx}: unknown key '{k}'")


def parse_prescription(text: str, origin: str = "<string>") -> Prescription:
    """A ``.lens`` table (src/owe/scene/prescription.cpp:97-140): ``name``, ``units``, ``ambient``,
    ``object_distance``, ``wavelength`` lines and one row per surface — label, R, t, the medium after
    it, semi-diameter, then options (``k=``, ``A4=``…, ``stop``, ``reflect``, ``material=``,
    ``substrate=``, ``hole=`` (a diameter), ``support=spider(…)``, ``solve``). Lengths are converted to
    millimetres, asphere terms to SI units; a mirror system's legs are checked (t signed with the
    travel, glass only on +z legs)."""
    p = Prescription([])
    unit = 1.0
    for n, raw in enumerate(text.splitlines(), 1):
        line = raw.split("#", 1)[0]
        tok = line.split()
        if not tok:
            continue
        ctx = f"{origin}:{n}"
        key = tok[0]
        if key == "name":
            p.name = line[line.index("name") + 4:].strip().strip('"').strip()
        elif key == "units":
            if len(tok) < 2 or tok[1] not in _LENS_UNITS:
                raise ValueError(f"{ctx}: unknown unit")
            unit = _LENS_UNITS[tok[1]]
        elif key == "ambient":
            p.ambient = tok[1]
        elif key == "object_distance":
            p.object_distance = _lens_number(tok[1], ctx) * unit
        elif key == "wavelength":
            p.wavelength = _lens_number(tok[1], ctx)
        else:
            if len(tok) < 5:
                raise ValueError(f"{ctx}: surface lines need: label R t medium semi-diameter")
            s = Surface(_lens_number(tok[1], ctx) * unit, _lens_number(tok[2], ctx) * unit, tok[3],
                        _lens_number(tok[4], ctx) * unit, label=tok[0])
            for x in tok[5:]:
                _lens_option(x, ctx, unit, s)
            if math.isinf(s.t):
                raise ValueError(f"{ctx}: thickness must be finite")
            if s.reflect and (not s.material or not s.substrate > 0):
                raise ValueError(f"{ctx}: a reflecting row needs material= and substrate= > 0")
            if s.reflect and s.hole >= s.sd:
                raise ValueError(f"{ctx}: the hole must be smaller than the mirror")
            p.surfaces.append(s)
    if not p.surfaces:
        raise ValueError(f"{origin}: prescription has no surfaces")
    if any(s.reflect for s in p.surfaces):
        direction = 1.0
        for i, s in enumerate(p.surfaces):
            if s.reflect:
                if s.medium != p.ambient:
                    raise ValueError(f"{origin}: {s.label}: a mirror reflects into the ambient medium")
                direction = -direction
            elif direction < 0:
                raise ValueError(f"{origin}: {s.label}: refraction on a leg travelling -z is not supported")
            if i + 1 < len(p.surfaces) and s.t * direction <= 0:
                raise ValueError(f"{origin}: {s.label}: t must have the sign of the travel after this row")
    return p


def load_prescription(path: str) -> Prescription:
    with open(path) as f:
        return parse_prescription(f.read(), path)


def _unfolded(p: Prescription, lam: float, index_fn):
    """Per surface: curvature (1/m), index after, power, gap to the next vertex along the travel (m),
    and the travel direction after the last surface (makeSystem, prescription.cpp:197-215)."""
    n_prev, direction = index_fn(p.ambient, lam), 1.0
    rows = []
    for sf in p.surfaces:
        cur = 0.0 if math.isinf(sf.R) else 1.0 / (sf.R * 1e-3)
        n = index_fn(sf.medium, lam)
        phi = -2 * direction * n * cur if sf.reflect else (n - n_prev) * cur
        if sf.reflect:
            direction = -direction
        rows.append((n_prev, n, phi, abs(sf.t * 1e-3)))
        n_prev = n
    return rows, direction


def system_matrix(p: Prescription, lam: float, index_fn=None):
    """The paraxial (y, n u) matrix (A, B, C, D) of paraxialAnalysis (prescription.cpp:193-230): a
    refracting row's power (n' − n) c, a mirror's −2 dir n c (dir the travel, which it reverses),
    the gaps unfolded."""
    rows, _ = _unfolded(p, lam, index_fn or index)
    a, b, c, d = 1.0, 0.0, 0.0, 1.0
    for i, (_, n, phi, t) in enumerate(rows):
        c, d = c - phi * a, d - phi * b
        if i + 1 < len(rows):
            k = t / n
            a, b = a + k * c, b + k * d
    return a, b, c, d


def stop_index(p: Prescription) -> int:
    """The last row marked stop, or −1."""
    stop = -1
    for i, sf in enumerate(p.surfaces):
        if sf.stop:
            stop = i
    return stop


def paraxial(p: Prescription, lam: float | None = None, index_fn=None) -> dict:
    """paraxialAnalysis (prescription.cpp:232-322), lengths in metres along the fixed chart: the system
    matrix; efl, bfl, ffl, or the angular magnification when afocal; the image distance of the object;
    the aperture stop (marked, else the surface that limits a marginal ray from the object); the exit
    pupil (the stop imaged by the surfaces after it: its z from the first vertex and radius) and the
    entrance pupil (imaged backwards by the surfaces before it); the f-number."""
    index_fn = index_fn or index
    lam = p.wavelength if lam is None else lam
    rows, dir_last = _unfolded(p, lam, index_fn)
    N = len(rows)
    A, B, C, D = system_matrix(p, lam, index_fn)
    n0, nk = rows[0][0], rows[-1][1]
    length = sum(r[3] for r in rows[:-1])
    z_chart = [0.0]
    for sf in p.surfaces[:-1]:
        z_chart.append(z_chart[-1] + sf.t * 1e-3)
    scale = 1e-12 * (1 + abs(A) + abs(D))
    out = dict(A=A, B=B, C=C, D=D, length=length, efl=0.0, bfl=0.0, ffl=0.0, angular_magnification=0.0,
               exit_pupil_z=0.0, exit_pupil_radius=0.0, entrance_pupil_z=0.0, entrance_pupil_radius=0.0, f_number=0.0)
    out["afocal"] = abs(C) * max(1.0, length) < scale * 1e3 or abs(C) < 1e-9
    if not out["afocal"]:
        out["efl"], out["bfl"], out["ffl"] = -nk / C, -nk * A / C, n0 * D / C
    else:
        out["angular_magnification"] = D * n0 / nk
    obj = p.object_distance * 1e-3
    if math.isinf(obj):
        out["image_distance"] = math.inf if out["afocal"] else out["bfl"]
    else:
        u = 1e-3
        y0 = obj * u
        yk, nuk = A * y0 + B * n0 * u, C * y0 + D * n0 * u
        out["image_distance"] = math.inf if nuk == 0 else -yk * nk / nuk
    stop = stop_index(p)
    if stop < 0:
        best, y = -1.0, 1.0
        nu = 0.0 if math.isinf(obj) else n0 / obj
        for i in range(N):
            ratio = abs(y) / max(p.surfaces[i].sd * 1e-3, 1e-12)
            if ratio > best:
                best, stop = ratio, i
            nu -= rows[i][2] * y
            if i + 1 < N:
                y += rows[i][3] * nu / rows[i][1]
    out["stop_index"] = stop
    stop_r = p.surfaces[stop].sd * 1e-3
    y, nu, ym, num = 0.0, 1e-3 * rows[stop][1], stop_r, 0.0
    for i in range(stop + 1, N):
        y += rows[i - 1][3] * nu / rows[i - 1][1]
        ym += rows[i - 1][3] * num / rows[i - 1][1]
        nu -= rows[i][2] * y
        num -= rows[i][2] * ym
    u = nu / nk
    if stop == N - 1:
        out["exit_pupil_z"], out["exit_pupil_radius"] = z_chart[-1], stop_r
    elif u != 0:
        dz = -y / u
        out["exit_pupil_z"] = z_chart[-1] + dir_last * dz
        out["exit_pupil_radius"] = abs(ym + dz * num / nk)
    y, nu, ym, num = 0.0, 1e-3 * rows[stop][0], stop_r, 0.0
    for i in range(stop - 1, -1, -1):
        y -= rows[i][3] * nu / rows[i][1]
        ym -= rows[i][3] * num / rows[i][1]
        nu += rows[i][2] * y
        num += rows[i][2] * ym
    if stop == 0:
        out["entrance_pupil_z"], out["entrance_pupil_radius"] = 0.0, stop_r
    else:
        u = nu / n0
        if u != 0:
            dz = -y / u
            out["entrance_pupil_z"], out["entrance_pupil_radius"] = dz, abs(ym + dz * num / n0)
    if not out["afocal"] and out["entrance_pupil_radius"] > 0:
        out["f_number"] = abs(out["efl"]) / (2 * out["entrance_pupil_radius"])
    return out


def solve_afocal(p: Prescription, lam: float = LAMBDA_D, index_fn=None) -> Prescription:
    """solveAfocal (prescription.cpp:324-352): the gap marked solve (a mirror system needs exactly
    one), else the largest air gap, set so the system's power C vanishes (C is affine in the gap),
    its sign (the leg's travel) kept."""
    rows = p.surfaces
    marked = [i for i, sf in enumerate(rows[:-1]) if sf.solve]
    if len(marked) > 1:
        raise ValueError("solveAfocal: more than one gap is marked solve")
    if not marked and any(sf.reflect for sf in rows):
        raise ValueError("solveAfocal: mark the gap to adjust with solve")
    if marked:
        g = marked[0]
    else:
        air = [i for i, sf in enumerate(rows[:-1]) if sf.medium == p.ambient]
        if not air:
            raise ValueError("solveAfocal: no air gap to adjust")
        g = max(air, key=lambda i: (rows[i].t, -i))
    old = rows[g].t
    sign = -1.0 if old < 0 else 1.0

    def power(t):
        q = [dataclasses.replace(sf) for sf in rows]
        q[g].t = t
        return system_matrix(dataclasses.replace(p, surfaces=q), lam, index_fn)[2]

    c0, c1 = power(0.0), power(sign)
    if c1 == c0:
        raise ValueError("solveAfocal: gap does not affect power")
    t = -c0 / (c1 - c0)
    if not t > 0:
        raise ValueError("solveAfocal: no positive afocal separation")
    q = [dataclasses.replace(sf) for sf in rows]
    q[g].t = sign * t
    return dataclasses.replace(p, surfaces=q)


def afocal_spot(m) -> dict:
    """afocalSpot (analysis.cpp:70-81) from the nine sums of angles: the apparent angle of the mean
    ray and the angular RMS about it (radians)."""
    m = np.asarray(m, dtype=np.float64)
    n = m[0]
    if n < 3:
        return dict(arrived=int(n), apparent_angle=math.nan, angular_rms=math.nan)
    c = np.arctan(m[3:5] / n)
    v = m[5] / n - 2 * (c @ m[1:3]) / n + c @ c
    return dict(arrived=int(n), apparent_angle=float(c[1]), angular_rms=math.sqrt(max(0.0, v)))


def gregorian_variants(n: int = 64, rings: int = 200, seed: int = 5) -> list:
    """n perturbed copies of the reference's Gregorian (lenses/gregorian_54x.lens), the tolerancing
    family of the differentiable-telescope workload: both mirrors' radii ±0.2 %, the secondary's
    conic ±0.002, the ocular's radii ±0.5 %, each re-solved afocal; `rings` pupil rings each."""
    base = solve_afocal(GREGORIAN, 587.5618).with_rings(rings)
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        sf = [Surface(**{**q.__dict__}) for q in base.surfaces]
        sf[0].R *= 1 + rng.uniform(-2e-3, 2e-3)
        sf[1].R *= 1 + rng.uniform(-2e-3, 2e-3)
        sf[1].k += rng.uniform(-2e-3, 2e-3)
        sf[2].R *= 1 + rng.uniform(-5e-3, 5e-3)
        sf[4].R *= 1 + rng.uniform(-5e-3, 5e-3)
        out.append(solve_afocal(Prescription(sf, base.ambient, base.rays, base.name, base.rings), 587.5618))
    return out


def c_surface(s: Surface) -> CLensSurface:
    c = CLensSurface()
    c.curvature = 0.0 if math.isinf(s.R) else 1.0 / (s.R * 1e-3)
    c.conic = s.k
    for i, a in enumerate(s.A):
        c.asphere[i] = a
    c.thickness = s.t * 1e-3
    c.semi_diameter = s.sd * 1e-3
    c.medium = MEDIA_ORDER.index(s.medium)
    c.stop = int(s.stop)
    c.reflect = int(s.reflect)
    c.hole_radius = s.hole * 1e-3
    return c


def c_batch(lenses):
    """(surfaces, lenses) C arrays for a list of Prescriptions."""
    rows = [c_surface(s) for p in lenses for s in p.surfaces]
    surf = (CLensSurface * max(1, len(rows)))(*rows)
    arr = (CLens * len(lenses))()
    first = 0
    for i, p in enumerate(lenses):
        arr[i].first_surface, arr[i].surface_count = first, len(p.surfaces)
        arr[i].rays, arr[i].ambient, arr[i].pupil_rings = int(p.rays), MEDIA_ORDER.index(p.ambient), int(p.rings)
        first += len(p.surfaces)
    return surf, arr


def c_settings(field_deg=0.5, wavelengths=FRAUNHOFER7, pixels=128, half_size=0.0, reference=LAMBDA_D, planes=1,
               focus_step=0.0):
    """Batch settings; `planes` grids spaced `focus_step` m (0: 1e-3 × |EFL|) about the paraxial image."""
    s = CLensSettings()
    s.field_deg = field_deg
    for i, l in enumerate(wavelengths):
        s.wavelengths[i] = l
    s.n_wavelengths = len(wavelengths)
    s.detector_pixels, s.detector_half_size, s.reference_wavelength = pixels, half_size, reference
    s.focus_planes, s.focus_step = planes, focus_step
    return s


def sag_mm(R: float, r: float) -> float:
    return 0.0 if math.isinf(R) else R - math.copysign(math.sqrt(R * R - r * r), R)


def seeded_achromats(n: int = 256, seed: int = 2024, rays_min: int = 4096, rays_max: int = 1 << 20):
    """n cemented achromats: f U(50, 150) mm, front radius ±U(30, 120) mm, element thicknesses
    U(2, 8) mm, crown/flint pairs from the catalog, f/8 clear aperture, rays log-uniform in
    [rays_min, rays_max]. Thin-lens achromat powers φ₁ = φ V₁/(V₁ − V₂), φ₂ = −φ V₂/(V₁ − V₂);
    a design whose radii leave no room for the aperture is drawn again."""
    rng = np.random.default_rng(seed)
    out = []
    while len(out) < n:
        f = rng.uniform(50.0, 150.0)
        R1 = rng.uniform(30.0, 120.0) * rng.choice((-1.0, 1.0))
        t1, t2 = rng.uniform(2.0, 8.0), rng.uniform(2.0, 8.0)
        crown, flint = CROWNS[rng.integers(len(CROWNS))], FLINTS[rng.integers(len(FLINTS))]
        rays = int(round(math.exp(rng.uniform(math.log(rays_min), math.log(rays_max)))))
        n1 = {l: index(crown, l) / ciddor_air(l) for l in (LAMBDA_F, LAMBDA_D, LAMBDA_C)}
        n2 = {l: index(flint, l) / ciddor_air(l) for l in (LAMBDA_F, LAMBDA_D, LAMBDA_C)}
        V1 = (n1[LAMBDA_D] - 1) / (n1[LAMBDA_F] - n1[LAMBDA_C])
        V2 = (n2[LAMBDA_D] - 1) / (n2[LAMBDA_F] - n2[LAMBDA_C])
        phi = 1.0 / f
        phi1, phi2 = phi * V1 / (V1 - V2), -phi * V2 / (V1 - V2)
        c1 = 1.0 / R1
        c2 = c1 - phi1 / (n1[LAMBDA_D] - 1)
        c3 = c2 - phi2 / (n2[LAMBDA_D] - 1)
        sd = f / 16.0
        radii = [1.0 / c if abs(c) > 1e-12 else math.inf for c in (c1, c2, c3)]
        if any(abs(R) < 1.2 * sd for R in radii):
            continue
        # Edge thickness of each element at least 1 mm.
        e1 = t1 + sag_mm(radii[1], sd) - sag_mm(radii[0], sd)
        e2 = t2 + sag_mm(radii[2], sd) - sag_mm(radii[1], sd)
        t1 += max(0.0, 1.0 - e1)
        t2 += max(0.0, 1.0 - e2)
        out.append(Prescription([Surface(radii[0], t1, crown, sd, stop=True), Surface(radii[1], t2, flint, sd),
                                 Surface(radii[2], f, "air", sd)], rays=rays,
                                name=f"achromat f={f:.1f} {crown}/{flint}"))
    return out


def spot(m, z_img: float, last_vertex_z: float) -> dict:
    """focusedSpot from the nine sums (count, ΣA, ΣB, Σ|A|², Σ|B|², ΣA·B, ΣT): the best-focus plane
    z = −cov/var, its RMS radius, and the RMS radius and centroid height at z_img."""
    m = np.asarray(m, dtype=np.float64)
    n = m[0]
    if n < 3:
        return dict(arrived=int(n), best_focus_z=math.nan, rms_best=math.nan, rms_paraxial=math.nan,
                    centroid_y=math.nan)
    ma, mb = m[1:3] / n, m[3:5] / n
    var_a = m[5] / n - ma @ ma
    var_b = m[6] / n - mb @ mb
    cov = m[7] / n - ma @ mb
    z_best = -cov / var_b if var_b > 0 else z_img

    def rms(z):
        return math.sqrt(max(0.0, var_a + 2 * z * cov + z * z * var_b))

    return dict(arrived=int(n), best_focus_z=z_best - last_vertex_z, rms_best=rms(z_best), rms_paraxial=rms(z_img),
                centroid_y=ma[1] + mb[1] * z_img)


class LensBatch:
    """Many lenses traced in one persistent launch on the GPU (owe_hip_lens_batch_*)."""

    def __init__(self, lenses, settings: CLensSettings | None = None, device: int = 0):
        self.lenses = list(lenses)
        self.settings = settings if settings is not None else c_settings()
        self._media = media_array()
        self._surf, self._arr = c_batch(self.lenses)
        h = ctypes.c_void_p()
        lib = _lib.load()
        _lib.check(lib.owe_hip_lens_batch_create(device, self._media, len(MEDIA_ORDER), self._surf,
                                                 sum(len(p.surfaces) for p in self.lenses), self._arr, len(self.lenses),
                                                 ctypes.byref(self.settings), ctypes.byref(h)))
        self._h = h
        self.n_params = 0

    def close(self):
        if self._h:
            _lib.check(_lib.load().owe_hip_lens_batch_destroy(self._h))
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run(self):
        _lib.check(_lib.load().owe_hip_lens_batch_run(self._h))

    def results(self, irradiance: bool = False):
        n, nw, px = len(self.lenses), self.settings.n_wavelengths, self.settings.detector_pixels
        planes = max(1, self.settings.focus_planes)
        res = (CLensResult * n)()
        mom = np.zeros((n, nw, MOMENTS))
        irr = np.zeros((n, planes, px, px), dtype=np.float32) if irradiance and px > 0 else None
        _lib.check(_lib.load().owe_hip_lens_batch_results(self._h, res, mom.ctypes.data,
                                                          irr.ctypes.data if irr is not None else None))
        per = [{f: getattr(r, f) for f, _ in CLensResult._fields_} for r in res]
        return per, mom, irr

    def set_shape(self, blocks_per_cu: int = 0, rays_per_item: int = 0):
        _lib.check(_lib.load().owe_hip_lens_batch_set_shape(self._h, int(blocks_per_cu), int(rays_per_item)))

    def autotune(self) -> float:
        ms = ctypes.c_double()
        _lib.check(_lib.load().owe_hip_lens_batch_autotune(self._h, ctypes.byref(ms)))
        return ms.value

    def last_ms(self) -> float:
        ms = ctypes.c_double()
        _lib.check(_lib.load().owe_hip_lens_batch_last_ms(self._h, ctypes.byref(ms)))
        return ms.value

    def set_params(self, params):
        """Differentiable parameters: (surface within each lens, PARAM_*) pairs, at most 8."""
        arr = (CLensParam * max(1, len(params)))()
        for k, (surface, field) in enumerate(params):
            arr[k].surface, arr[k].field = int(surface), int(field)
        _lib.check(_lib.load().owe_hip_lens_batch_set_params(self._h, arr, len(params)))
        self.n_params = len(params)

    def gradients(self) -> np.ndarray:
        """d(spot metric)/dθ of the last run, [lens, parameter] (SI units): rms_best when focused,
        angular_rms when afocal."""
        out = np.zeros((len(self.lenses), self.n_params))
        _lib.check(_lib.load().owe_hip_lens_batch_gradients(self._h, out.ctypes.data))
        return out

    def trace(self, lens: int, wavelength: int, index) -> np.ndarray:
        index = np.ascontiguousarray(index, dtype=np.uint64)
        out = np.zeros((len(index), 7))
        _lib.check(_lib.load().owe_hip_lens_batch_trace(self._h, int(lens), int(wavelength), index.ctypes.data,
                                                        len(index), out.ctypes.data))
        return out
