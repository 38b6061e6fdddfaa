"""Synthetic, seeded inputs of the measured workloads.

Batch ray queries on a BVH: a subdivided icosphere of radius 1 with radial value-noise
displacement of amplitude 0.02, stored as float32 vertices and uint32 indices, and rays from a
radius-3 sphere aimed at uniform points of the unit ball.

Triangle count. An icosphere subdivided at frequency n (each of the 20 faces cut into n² triangles)
has 20·n² triangles, which never equals 2^20 = 1,048,576: ``icosphere(n=229)`` gives 1,048,820
(+0.023 %), the nearest count at or above 2^20. Edges are cut canonically (from the lower-indexed
corner) so the mesh is watertight: shared edge vertices are one vertex.

Displacement. radius(p) = 1 + 0.02·(2·valueNoise3(4·p̂ + offset(seed)) − 1) with p̂ the unit
direction and valueNoise3 the reference's lattice value noise (src/owe/core/shared/textures.hpp:13-56),
restated here in numpy.
"""
from __future__ import annotations

import numpy as np

_M1 = np.uint64(0x7FB5D329728EA185)
_M2 = np.uint64(0x81DADEF4BC2DD44D)
_GOLD = np.uint64(0x9E3779B97F4A7C15)


def _mix64(v: np.ndarray) -> np.ndarray:
    v = np.asarray(v, dtype=np.uint64)
    with np.errstate(over="ignore"):
        v = (v ^ (v >> np.uint64(31))) * _M1
        v = (v ^ (v >> np.uint64(27))) * _M2
    return v ^ (v >> np.uint64(33))


def value_noise3(p: np.ndarray) -> np.ndarray:
    """Lattice value noise in [0, 1) with a smoothstep fade, for points p [..., 3] (float64).

    Corner value = latticeValue(hash(hash(x, y), z)) with hash(a, b) = mix64(a ^ (mix64(b) + φ)),
    as textures.hpp:21-56; the 53-bit lattice value keeps the high bits of the hash.
    """
    p = np.asarray(p, dtype=np.float64)
    f = np.floor(p)
    i = f.astype(np.int64)
    u = p - f
    t = u * u * (3 - 2 * u)
    with np.errstate(over="ignore"):
        def key(k):  # a coordinate as the uint32 the reference converts it to
            return (k.astype(np.int64) & 0xFFFFFFFF).astype(np.uint64)
        my0 = _mix64(key(i[..., 1])) + _GOLD
        my1 = _mix64(key(i[..., 1] + 1)) + _GOLD
        mz0 = _mix64(key(i[..., 2])) + _GOLD
        mz1 = _mix64(key(i[..., 2] + 1)) + _GOLD
        r = np.zeros(p.shape[:-1])
        for c in range(8):
            dx, dy, dz = c & 1, (c >> 1) & 1, (c >> 2) & 1
            kxy = _mix64(key(i[..., 0] + dx) ^ (my1 if dy else my0))
            h = _mix64(kxy ^ (mz1 if dz else mz0))
            v = (h >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
            wx = t[..., 0] if dx else 1 - t[..., 0]
            wy = t[..., 1] if dy else 1 - t[..., 1]
            wz = t[..., 2] if dz else 1 - t[..., 2]
            r += wx * wy * wz * v
    return r


def _icosahedron():
    g = (1 + 5 ** 0.5) / 2
    v = np.array([[-1, g, 0], [1, g, 0], [-1, -g, 0], [1, -g, 0], [0, -1, g], [0, 1, g], [0, -1, -g], [0, 1, -g],
                  [g, 0, -1], [g, 0, 1], [-g, 0, -1], [-g, 0, 1]], dtype=np.float64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4], [11, 10, 2],
                  [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5],
                  [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], dtype=np.int64)
    return v, f


def icosphere(n: int = 229, seed: int = 1, amplitude: float = 0.02, frequency: float = 4.0):
    """Watertight geodesic icosphere at subdivision frequency n, displaced radially.

    Returns (positions float32 [V, 3], triangles uint32 [20·n², 3]) with outward winding.
    """
    if n < 1:
        raise ValueError("frequency must be >= 1")
    cv, cf = _icosahedron()
    edges = {}
    for a, b, c in cf:
        for x, y in ((a, b), (b, c), (c, a)):
            edges.setdefault((min(x, y), max(x, y)), len(edges))
    n_edge_pts = n - 1
    n_face_pts = (n - 1) * (n - 2) // 2
    n_vert = 12 + len(edges) * n_edge_pts + 20 * n_face_pts
    pos = np.empty((n_vert, 3), dtype=np.float64)
    pos[:12] = cv
    k = np.arange(1, n, dtype=np.float64) / n
    for (a, b), e in edges.items():
        base = 12 + e * n_edge_pts
        pos[base:base + n_edge_pts] = cv[a][None, :] * (1 - k)[:, None] + cv[b][None, :] * k[:, None]
    face_base = 12 + len(edges) * n_edge_pts
    tris = []
    ii, jj = np.meshgrid(np.arange(n + 1), np.arange(n + 1), indexing="ij")
    valid = ii + jj <= n
    gi, gj = ii[valid], jj[valid]
    for fi, (A, B, C) in enumerate(cf):
        vid = np.full((n + 1, n + 1), -1, dtype=np.int64)
        vid[0, 0], vid[n, 0], vid[0, n] = A, B, C

        def edge_ids(x, y, count_from_x):
            e = edges[(min(x, y), max(x, y))]
            kk = np.asarray(count_from_x)
            kk = kk if x < y else n - kk
            return 12 + e * n_edge_pts + (kk - 1)

        r = np.arange(1, n)
        vid[r, 0] = edge_ids(A, B, r)            # A + (B-A) i/n
        vid[0, r] = edge_ids(A, C, r)            # A + (C-A) j/n
        vid[n - r, r] = edge_ids(B, C, r)        # B + (C-B) j/n
        inner = (gi >= 1) & (gj >= 1) & (gi + gj <= n - 1)
        fi_i, fi_j = gi[inner], gj[inner]
        ids = face_base + fi * n_face_pts + np.arange(len(fi_i))
        vid[fi_i, fi_j] = ids
        pos[ids] = (cv[A][None, :] + (cv[B] - cv[A])[None, :] * (fi_i / n)[:, None]
                    + (cv[C] - cv[A])[None, :] * (fi_j / n)[:, None])
        up_i, up_j = gi[gi + gj <= n - 1], gj[gi + gj <= n - 1]
        tris.append(np.stack([vid[up_i, up_j], vid[up_i + 1, up_j], vid[up_i, up_j + 1]], axis=1))
        dn = gi + gj <= n - 2
        dn_i, dn_j = gi[dn], gj[dn]
        tris.append(np.stack([vid[dn_i + 1, dn_j], vid[dn_i + 1, dn_j + 1], vid[dn_i, dn_j + 1]], axis=1))
    tri = np.concatenate(tris).astype(np.uint32)
    unit = pos / np.linalg.norm(pos, axis=1, keepdims=True)
    rng = np.random.Generator(np.random.PCG64(seed))
    offset = rng.uniform(0, 1024, size=3)
    radius = 1.0 + amplitude * (2.0 * value_noise3(unit * frequency + offset) - 1.0)
    return (unit * radius[:, None]).astype(np.float32), tri


def query_rays(count: int = 8 * 1024 * 1024, seed: int = 2, origin_radius: float = 3.0):
    """Seeded query rays [count, 8] float32 and their targets.

    Origins are uniform on the sphere of radius 3; each direction points at a uniform point of the
    unit ball. Closest-hit rays use tmin = 0, tmax = +inf. The any-hit batch asks whether the segment
    from the origin to its target is blocked: tmax = |target − origin| in units of |d| (d is the unit
    direction rounded to float32).
    """
    rng = np.random.Generator(np.random.PCG64(seed))
    o = rng.standard_normal((count, 3))
    o *= origin_radius / np.linalg.norm(o, axis=1, keepdims=True)
    g = rng.standard_normal((count, 3))
    g /= np.linalg.norm(g, axis=1, keepdims=True)
    target = g * np.cbrt(rng.uniform(0, 1, size=(count, 1)))
    d = target - o
    dist = np.linalg.norm(d, axis=1)
    d /= dist[:, None]
    closest = np.empty((count, 8), dtype=np.float32)
    closest[:, 0:3] = o
    closest[:, 3] = 0.0
    closest[:, 4:7] = d
    closest[:, 7] = np.inf
    any_hit = closest.copy()
    d32 = any_hit[:, 4:7].astype(np.float64)
    any_hit[:, 7] = (dist / np.linalg.norm(d32, axis=1)).astype(np.float32)
    return closest, any_hit
