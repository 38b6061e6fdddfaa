"""Comparison of float32 ray-query answers with the fp64 oracle, with ties and grazing cases named.

The bar: primitive IDs bit-exact, except
  * ties — the two answers' triangles are crossed at plane distances within 1e-6 relative (two
    triangles at one distance: the ray passes a shared edge or vertex); and
  * grazing flips — hit against miss, or two triangles at different distances, where the ray
    passes the disputed triangle's boundary closer than float32 can resolve;
and t within 1e-5 relative wherever the IDs agree.

"Closer than float32 can resolve" is an absolute distance: 16 float32 ulps of the coordinates
involved, 16 · 2⁻²⁴ · (|o| + t·|d|) — about 3e-6 m for rays starting at radius 3. The GPU's own
triangle must be crossed within that distance of its edges for any exception to apply.
"""
from __future__ import annotations

import numpy as np

MISS = 0xFFFFFFFF
TIE_REL = 1e-6
T_REL = 1e-5
ULPS = 16


def _seg_dist(p, a, b):
    ab = b - a
    s = np.clip(((p - a) * ab).sum(-1) / np.maximum((ab * ab).sum(-1), 1e-300), 0, 1)
    return np.linalg.norm(p - (a + ab * s[..., None]), axis=-1)


def plane_t_and_outside(pos: np.ndarray, tri: np.ndarray, rays: np.ndarray, prims: np.ndarray):
    """fp64 plane distance t of each ray on triangle prims, how far outside the triangle the plane
    point lies (0 inside), and the float32 resolution at that point."""
    v = pos.astype(np.float64)[tri[prims].astype(np.int64)]  # [k, 3, 3]
    o = rays[:, 0:3].astype(np.float64)
    d = rays[:, 4:7].astype(np.float64)
    e1, e2 = v[:, 1] - v[:, 0], v[:, 2] - v[:, 0]
    n = np.cross(e1, e2)
    t = -np.einsum("ij,ij->i", o - v[:, 0], n) / np.einsum("ij,ij->i", d, n)
    p = o + d * t[:, None]
    q = p - v[:, 0]
    d00, d01, d11 = (e1 * e1).sum(1), (e1 * e2).sum(1), (e2 * e2).sum(1)
    d20, d21 = (q * e1).sum(1), (q * e2).sum(1)
    den = d00 * d11 - d01 * d01
    b1 = (d11 * d20 - d01 * d21) / den
    b2 = (d00 * d21 - d01 * d20) / den
    b0 = 1 - b1 - b2
    inside = (b0 >= 0) & (b1 >= 0) & (b2 >= 0)
    dist = np.minimum(np.minimum(_seg_dist(p, v[:, 0], v[:, 1]), _seg_dist(p, v[:, 1], v[:, 2])),
                      _seg_dist(p, v[:, 2], v[:, 0]))
    outside = np.where(inside, 0.0, dist)
    edge = np.where(inside, dist, 0.0)  # how far inside from the nearest edge
    resolution = ULPS * 2.0 ** -24 * (np.abs(o).max(1) + np.abs(t) * np.linalg.norm(d, axis=1))
    return t, outside, edge, resolution


def classify(pos, tri, rays, gpu_prim, gpu_t, ora_prim, ora_t):
    """Returns a dict of counts and the indices of unexplained disagreements."""
    gpu_prim = gpu_prim.astype(np.uint32)
    same = gpu_prim == ora_prim
    both_hit = same & (ora_prim != MISS)
    rel = np.abs(gpu_t[both_hit].astype(np.float64) - ora_t[both_hit]) / np.abs(ora_t[both_hit])
    diff = np.nonzero(~same)[0]
    ties = grazing = 0
    bad = []
    for i in diff:
        g, o = int(gpu_prim[i]), int(ora_prim[i])
        r = rays[i:i + 1]
        if g != MISS:
            tg, out_g, _, res_g = plane_t_and_outside(pos, tri, r, np.array([g]))
            if out_g[0] > res_g[0]:
                bad.append(int(i))  # the GPU reports a triangle the line clearly misses
                continue
        if g != MISS and o != MISS and abs(tg[0] - ora_t[i]) <= TIE_REL * abs(ora_t[i]):
            ties += 1
            continue
        if o != MISS:
            _, _, edge_o, res_o = plane_t_and_outside(pos, tri, r, np.array([o]))
            if edge_o[0] <= res_o[0]:
                grazing += 1  # the oracle's crossing is within float32 resolution of its edge
                continue
        elif g != MISS:
            grazing += 1  # the GPU crossed, within resolution of the edge, what fp64 misses
            continue
        bad.append(int(i))
    return {
        "n": len(rays),
        "agree": int(same.sum()),
        "ties": ties,
        "grazing": grazing,
        "bad": bad,
        "max_rel_t": float(rel.max()) if rel.size else 0.0,
        "hits": int((ora_prim != MISS).sum()),
    }
