"""TEST INFRASTRUCTURE ONLY — ctypes loaders of the double-precision CPU oracle.

``Oracle`` wraps build/liboracle.so (the repository's restatement of the reference transport, see
oracle_core.hpp for the citations); ``Reference`` wraps _ref/libowe_ref.so, the harness over the
reference engine's own sources, which exists only where the reference was mounted at build time.
Only tests, __graft_entry__.smoke() and bench.py's CPU-baseline leg import this module.
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ORACLE_LIB = HERE / "build" / "liboracle.so"
REF_LIB = HERE / "_ref" / "libowe_ref.so"

_d, _i, _u32, _u64, _vp = ctypes.c_double, ctypes.c_int, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_void_p
_pd = ctypes.POINTER(ctypes.c_double)


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


class _Lib:
    prefix = ""

    def __init__(self, path: Path):
        if not path.exists():
            raise FileNotFoundError(f"{path} is not built (make -C oracle{' ref' if 'ref' in path.name else ''})")
        self.lib = ctypes.CDLL(os.fspath(path))
        p = self.prefix
        sigs = {
            "pcg_u32": (None, [_u64, _u64, _i, _vp]),
            "pcg_uniform": (None, [_u64, _u64, _i, _vp]),
            "mix64": (_u64, [_u64]),
            "hash_pair": (_u64, [_u64, _u64]),
            "fresnel_dielectric": (_d, [_d, _d]),
            "fresnel_conductor": (_d, [_d, _d, _d]),
            "sample_visible": (_d, [_d]),
            "visible_pdf": (_d, [_d]),
            "cie": (None, [_d, _vp]),
            "cie_y_integral": (_d, []),
            "planck": (_d, [_d, _d]),
            "hg_phase": (_d, [_d, _d]),
            "sample_hg": (_d, [_vp, _d, _d, _d, _vp]),
            "ggx_d": (_d, [_d, _vp]),
            "ggx_pdf": (_d, [_d, _vp, _vp]),
            "ggx_sample_wm": (None, [_d, _vp, _d, _d, _vp]),
            "sample_rough_dielectric": (_i, [_d, _vp, _d, _d, _d, _d, _i, _vp]),
            "eval_rough_dielectric": (_d, [_d, _vp, _vp, _d, _i, _vp]),
            "sample_rough_conductor": (_i, [_d, _vp, _d, _d, _d, _d, _d, _vp]),
            "eval_rough_conductor": (_d, [_d, _vp, _vp, _d, _d, _d, _vp]),
            "line_crosses_triangle": (_i, [_vp, _vp, _vp]),
            "mesh_create": (_vp, [_vp, _u32, _vp, _u32]),
            "mesh_destroy": (None, [_vp]),
            "mesh_node_count": (_u64, [_vp]),
            "mesh_order": (None, [_vp, _vp]),
            "mesh_closest": (None, [_vp, _vp, _u64, _vp, _vp, _i]),
        }
        sigs.update(self.extra_sigs())
        for name, (res, args) in sigs.items():
            fn = getattr(self.lib, p + name, None)
            if fn is None:
                continue
            fn.restype = res
            fn.argtypes = args
            setattr(self, name, fn)

    def extra_sigs(self):
        return {}

    # ---- convenience wrappers shared by the oracle and the reference harness

    def pcg(self, seed: int, stream: int, n: int) -> np.ndarray:
        out = np.empty(n, dtype=np.uint32)
        self.pcg_u32(seed, stream, n, _ptr(out))
        return out

    def uniforms(self, seed: int, stream: int, n: int) -> np.ndarray:
        out = np.empty(n, dtype=np.float64)
        self.pcg_uniform(seed, stream, n, _ptr(out))
        return out

    def cie_xyz(self, lam: float) -> np.ndarray:
        out = np.empty(3)
        self.cie(lam, _ptr(out))
        return out

    def mesh(self, positions: np.ndarray, triangles: np.ndarray):
        pos = np.ascontiguousarray(positions, dtype=np.float64).reshape(-1, 3)
        tri = np.ascontiguousarray(triangles, dtype=np.uint32).reshape(-1, 3)
        h = self.mesh_create(_ptr(pos), len(pos), _ptr(tri), len(tri))
        return _Mesh(self, h, len(tri))


class _Mesh:
    def __init__(self, owner: _Lib, handle, n_tris: int):
        self.o, self.h, self.n_tris = owner, handle, n_tris

    def __del__(self):
        try:
            self.o.mesh_destroy(self.h)
        except Exception:
            pass

    def _rays(self, rays):
        r = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 8)
        return r, len(r)

    def closest(self, rays, threads: int = 0):
        r, n = self._rays(rays)
        prim = np.empty(n, dtype=np.uint32)
        t = np.empty(n, dtype=np.float64)
        self.o.mesh_closest(self.h, _ptr(r), n, _ptr(prim), _ptr(t), threads)
        return prim, t

    def any(self, rays, threads: int = 0):
        r, n = self._rays(rays)
        hit = np.empty(n, dtype=np.uint8)
        self.o.mesh_any(self.h, _ptr(r), n, _ptr(hit), threads)
        return hit.astype(bool)

    def brute(self, rays, threads: int = 0):
        r, n = self._rays(rays)
        prim = np.empty(n, dtype=np.uint32)
        t = np.empty(n, dtype=np.float64)
        self.o.mesh_brute(self.h, _ptr(r), n, _ptr(prim), _ptr(t), threads)
        return prim, t

    def traversal_stats(self, rays, threads: int = 0) -> dict:
        r, n = self._rays(rays)
        out = np.zeros(4)
        self.o.mesh_traversal_stats(self.h, _ptr(r), n, _ptr(out), threads)
        return {"slab_tests": out[0], "triangle_tests": out[1], "inner_nodes": out[3]}

    def order(self) -> np.ndarray:
        out = np.empty(self.n_tris, dtype=np.uint32)
        self.o.mesh_order(self.h, _ptr(out))
        return out

    def node_count(self) -> int:
        return int(self.o.mesh_node_count(self.h))


class Oracle(_Lib):
    """The repository's fp64 restatement (liboracle.so)."""

    prefix = "oracle_"

    def __init__(self, path: Path = ORACLE_LIB):
        super().__init__(path)

    def extra_sigs(self):
        return {
            "world_create": (_vp, [_vp]),
            "world_destroy": (None, [_vp]),
            "world_contacts": (_i, [_vp]),
            "light_table": (_i, [_vp, _vp, _vp, _i]),
            "render": (None, [_vp, _vp, _i, _i, _i, _i, _vp, _vp, _vp]),
            "camera_paths": (None, [_vp, _vp, _i, _i, _u64, _vp, _vp, _vp, _vp, _i]),
            "render_particles": (None, [_vp, _vp, _i, _i, _i, _vp, _vp]),
            "sppm": (None, [_vp, _vp, _i, _i, _i, _vp, _vp, _vp, _vp]),
            "sppm_radius": (None, [_vp, _vp, _vp]),
            "vcm": (None, [_vp, _vp, _i, _i, _i, _i, _vp, _vp, _vp, _vp]),
            "particles": (None, [_vp, _vp, _i, _u64, _vp, _vp, _vp, _i]),
            "emission_info": (None, [_vp, _vp]),
            "world_intersect": (None, [_vp, _u64, _vp, _vp, _vp, _vp]),
            "lens_trace": (None, [_vp, _vp, _vp, _d, _d, _d, _vp, _u64, _vp]),
            "lens_moments": (None, [_vp, _vp, _vp, _d, _d, _d, _i, _vp, _vp]),
            "sag_intersect": (None, [_vp, _vp, _vp, _d, _d, _vp]),
            "sag_value": (_d, [_vp, _d]),
            "lens_trace_ray": (None, [_vp, _vp, _vp, _d, _vp, _vp, _vp]),
            "lens_trace_perturbed": (None, [_vp, _vp, _vp, _d, _d, _d, _u32, _vp, _vp, _vp, _u64, _vp]),
            "caustic_forward": (None, [_vp, _vp, _u64, _vp, _i]),
            "caustic_fd": (None, [_vp, _vp, _vp, _u64, _vp, _i, _d, _vp]),
            "spectrum_luminance_blackbody": (_d, [_d]),
            "medium_index": (_d, [_vp, _d]),
            "ciddor_air": (_d, [_d]),
            "mesh_any": (None, [_vp, _vp, _u64, _vp, _i]),
            "mesh_brute": (None, [_vp, _vp, _u64, _vp, _vp, _i]),
            "mesh_traversal_stats": (None, [_vp, _vp, _u64, _vp, _i]),
        }


class Reference(_Lib):
    """The reference engine's own code (libowe_ref.so), built from /root/reference in this container."""

    prefix = "ref_"

    def __init__(self, path: Path = REF_LIB):
        super().__init__(path)

    def extra_sigs(self):
        return {
            "value_noise3": (_d, [_vp]),
            "catalog_index": (_d, [ctypes.c_char_p, _d]),
            "sellmeier_index": (_d, [_vp, _vp, _i, _i, _d]),
            "blackbody": (_d, [_d, _d, _d]),
            "wavelengths": (None, [_d, _vp]),
        }


def reference_available() -> bool:
    return REF_LIB.exists()


class World:
    """The oracle's fp64 world built from a flattened scene (owe_hip.scene.Scene)."""

    def __init__(self, oracle: Oracle, scene):
        self.o = oracle
        self.scene = scene
        self.c = scene.c()
        self.h = oracle.world_create(ctypes.byref(self.c))
        ob = scene.observer
        self.width, self.height = ob["width"], ob["height"]

    def __del__(self):
        try:
            self.o.world_destroy(self.h)
        except Exception:
            pass

    @property
    def contacts(self) -> bool:
        return bool(self.o.world_contacts(self.h))

    def light_table(self):
        pdf = np.zeros(64)
        b = np.zeros(64, dtype=np.uint32)
        n = self.o.light_table(self.h, _ptr(pdf), _ptr(b), 64)
        return pdf[:n], b[:n]

    def render(self, settings, spp: int, passes: int = 1, layout: int = 1, threads: int = 0, first_pass: int = 0):
        """Resolved XYZ image [H, W, 3] (film ÷ samples) and stats (paths, segments, inconsistencies, leaks)."""
        xyz = np.zeros((self.height, self.width, 3))
        samples = np.zeros((self.height, self.width))
        stats = np.zeros(4, dtype=np.uint64)
        for p in range(first_pass, first_pass + passes):
            self.o.render(self.h, ctypes.byref(settings), spp, p, layout, threads, _ptr(xyz), _ptr(samples), _ptr(stats))
        return xyz / np.maximum(samples, 1)[..., None], stats

    def camera_paths(self, settings, spp: int, pixel, sample, pass_index: int = 0, threads: int = 0):
        pixel = np.ascontiguousarray(pixel, dtype=np.uint32)
        sample = np.ascontiguousarray(sample, dtype=np.uint32)
        n = len(pixel)
        xyz = np.zeros((n, 3))
        segs = np.zeros(n, dtype=np.uint32)
        self.o.camera_paths(self.h, ctypes.byref(settings), spp, pass_index, n, _ptr(pixel), _ptr(sample), _ptr(xyz),
                            _ptr(segs), threads)
        return xyz, segs

    def intersect(self, origins, directions, departed=None) -> np.ndarray:
        """World::intersect per ray: [n, 8] = t, point, normal, boundary (inf, NaN, 2^32 − 2 on a miss)."""
        o = np.ascontiguousarray(origins, dtype=np.float64).reshape(-1, 3)
        d = np.ascontiguousarray(directions, dtype=np.float64).reshape(-1, 3)
        dep = None if departed is None else np.ascontiguousarray(departed, dtype=np.uint32)
        out = np.zeros((len(o), 8))
        self.o.world_intersect(self.h, len(o), _ptr(o), _ptr(d), None if dep is None else _ptr(dep), _ptr(out))
        return out

    def render_particles(self, settings, ppp: int, passes: int = 1, threads: int = 0, first_pass: int = 0):
        """Light tracing: resolved XYZ image [H, W, 3] (splats ÷ particles emitted) and stats."""
        xyz = np.zeros((self.height, self.width, 3))
        stats = np.zeros(4, dtype=np.uint64)
        for p in range(first_pass, first_pass + passes):
            self.o.render_particles(self.h, ctypes.byref(settings), ppp, p, threads, _ptr(xyz), _ptr(stats))
        return xyz / float(self.width * self.height * ppp * passes), stats

    def sppm(self, settings, iterations: int, state=None, threads: int = 0):
        """SPPM iterations continuing `state` (a dict from an earlier call, or None to start): returns the
        new state with "direct" (W·H·3 camera-pass XYZ sums), "tau" (W·H·3), "N", "r", "iterations",
        "photons" (per iteration) and "stats"; sppm_resolve turns it into the image."""
        W, H = self.width, self.height
        if state is None:
            state = dict(direct=np.zeros((H, W, 3)), pix=np.zeros((H, W, 5)), iterations=0,
                         stats=np.zeros(4, dtype=np.uint64), guide=np.zeros(128 * 128, dtype=np.uint64))
        direct = np.ascontiguousarray(state["direct"]).copy()
        pix = np.ascontiguousarray(state["pix"]).copy()
        stats = state["stats"].copy()
        guide = state["guide"].copy()
        self.o.sppm(self.h, ctypes.byref(settings), iterations, state["iterations"], threads, _ptr(direct), _ptr(pix),
                    _ptr(guide), _ptr(stats))
        ppp = settings.photons_per_pixel if settings.photons_per_pixel > 0 else 1.0
        return dict(direct=direct, pix=pix, iterations=state["iterations"] + iterations, stats=stats, guide=guide,
                    photons=max(1, int(round(W * H * ppp))), tau=pix[..., :3], N=pix[..., 3], r=pix[..., 4])

    def vcm(self, settings, iterations: int, merging: bool = True, start: int = 0, threads: int = 0):
        """Vertex connection and merging (bdpt when merging is off), iterations [start, start + n):
        the camera paths' XYZ and the light subpaths' splats summed over the iterations, each [H, W, 3],
        the light subpaths per iteration N, the stats and each iteration's cache size. The image is
        camera / iterations + light / (N · iterations)."""
        W, H = self.width, self.height
        camera, light = np.zeros((H, W, 3)), np.zeros((H, W, 3))
        stats = np.zeros(4, dtype=np.uint64)
        cached = np.zeros(max(iterations, 1), dtype=np.uint64)
        self.o.vcm(self.h, ctypes.byref(settings), int(bool(merging)), iterations, start, threads, _ptr(camera),
                   _ptr(light), _ptr(stats), _ptr(cached))
        G = settings.photon_groups if settings.photon_groups > 0 else 1
        ppp = settings.photons_per_pixel if settings.photons_per_pixel > 0 else 1.0
        want = max(1, int(round(W * H * ppp)))
        N = (want + G - 1) // G * G
        return dict(camera=camera, light=light, paths=N, stats=stats, cached=cached[:iterations],
                    image=camera / max(iterations, 1) + light / (N * max(iterations, 1)))

    def sppm_radius(self, settings):
        """(initial radius per metre of path, radius floor)."""
        out = np.zeros(2)
        self.o.sppm_radius(self.h, ctypes.byref(settings), _ptr(out))
        return float(out[0]), float(out[1])

    @staticmethod
    def sppm_resolve(state) -> np.ndarray:
        """L = L_direct / iterations + τ / (N_emitted π r²) per pixel (sppm.slang:19-20)."""
        it = state["iterations"]
        r = state["pix"][..., 4]
        emitted = float(state["photons"]) * it
        with np.errstate(divide="ignore", invalid="ignore"):
            photon = np.where(r[..., None] > 0, state["pix"][..., :3] / (emitted * np.pi * r[..., None] ** 2), 0.0)
        return state["direct"] / max(it, 1) + photon

    def particles(self, settings, index, pass_index: int = 0, threads: int = 0):
        """Per-particle XYZ (summed over its splats, not normalised) and segments."""
        index = np.ascontiguousarray(index, dtype=np.uint64)
        n = len(index)
        xyz = np.zeros((n, 3))
        segs = np.zeros(n, dtype=np.uint32)
        self.o.particles(self.h, ctypes.byref(settings), pass_index, n, _ptr(index), _ptr(xyz), _ptr(segs), threads)
        return xyz, segs

    def emission_info(self) -> dict:
        out = np.zeros(7)
        self.o.emission_info(self.h, _ptr(out))
        return dict(sky_share=out[0], scene_radius=out[1], center=out[2:5].copy(), environment_power=out[5],
                    light_power=out[6])


class LensOracle:
    """The oracle's fp64 sequential lens trace over the same C records as the GPU batch
    (owe_hip.lens): per-ray traces, spot sums and the paraxial quantities."""

    def __init__(self, oracle: Oracle, lenses, field_deg: float = 0.5, reference: float = 587.56):
        from owe_hip import lens as L

        self.o = oracle
        self.lenses = list(lenses)
        self.field_deg, self.reference = field_deg, reference
        self.media = L.media_array()
        self.surf, self.arr = L.c_batch(self.lenses)

    def trace(self, lens: int, lam: float, index) -> np.ndarray:
        index = np.ascontiguousarray(index, dtype=np.uint64)
        out = np.zeros((len(index), 7))
        self.o.lens_trace(self.media, self.surf, ctypes.byref(self.arr[lens]), self.field_deg, lam, self.reference,
                          _ptr(index), len(index), _ptr(out))
        return out

    def trace_ray(self, lens: int, lam: float, o, d) -> np.ndarray:
        o = np.ascontiguousarray(o, dtype=np.float64)
        d = np.ascontiguousarray(d, dtype=np.float64)
        out = np.zeros(7)
        self.o.lens_trace_ray(self.media, self.surf, ctypes.byref(self.arr[lens]), lam, _ptr(o), _ptr(d), _ptr(out))
        return out

    def sag_intersect(self, row, o, d, tmin: float = 0.0, tmax: float = np.inf) -> np.ndarray:
        o = np.ascontiguousarray(o, dtype=np.float64)
        d = np.ascontiguousarray(d, dtype=np.float64)
        out = np.zeros(7)
        self.o.sag_intersect(ctypes.byref(row), _ptr(o), _ptr(d), tmin, tmax, _ptr(out))
        return out

    def moments(self, lens: int, lam: float, threads: int = 0):
        m = np.zeros(9)
        px = np.zeros(8)
        self.o.lens_moments(self.media, self.surf, ctypes.byref(self.arr[lens]), self.field_deg, lam, self.reference,
                            threads, _ptr(m), _ptr(px))
        return m, dict(efl=px[0], bfl=px[1], entrance_pupil_z=px[2], entrance_pupil_radius=px[3], length=px[4],
                       afocal=bool(px[5]), angular_magnification=px[6], last_vertex_z=px[7])

    def trace_perturbed(self, lens: int, lam: float, params, delta, index) -> np.ndarray:
        """trace() with parameter k (surface, field) of the lens moved by delta[k], from the
        unperturbed lens's launch rays."""
        index = np.ascontiguousarray(index, dtype=np.uint64)
        pr = np.ascontiguousarray(np.asarray(params, dtype=np.int32).reshape(-1))
        dl = np.ascontiguousarray(delta, dtype=np.float64)
        out = np.zeros((len(index), 7))
        self.o.lens_trace_perturbed(self.media, self.surf, ctypes.byref(self.arr[lens]), self.field_deg, lam,
                                    self.reference, len(dl), _ptr(pr), _ptr(dl), _ptr(index), len(index), _ptr(out))
        return out


class CausticOracle:
    """The oracle's fp64 caustic slab over the same settings record as the GPU (owe_hip.caustic)."""

    def __init__(self, oracle: Oracle, settings):
        self.o, self.s = oracle, settings

    def forward(self, heights, pass_index: int = 0, threads: int = 0) -> np.ndarray:
        h = np.ascontiguousarray(heights, dtype=np.float64)
        E = np.zeros((self.s.pixels, self.s.pixels))
        self.o.caustic_forward(ctypes.byref(self.s), _ptr(h), pass_index, _ptr(E), threads)
        return E

    def fd(self, heights, target, nodes, h: float = 1e-7, pass_index: int = 0) -> np.ndarray:
        """Central differences of mean((E − target)²) for the heights at flat indices `nodes`."""
        hh = np.ascontiguousarray(heights, dtype=np.float64)
        tt = np.ascontiguousarray(target, dtype=np.float64)
        nn = np.ascontiguousarray(nodes, dtype=np.uint32)
        g = np.zeros(len(nn))
        self.o.caustic_fd(ctypes.byref(self.s), _ptr(hh), _ptr(tt), pass_index, _ptr(nn), len(nn), h, _ptr(g))
        return g

