"""Comparison of float32 ray-query answers with the fp64 oracle, with ties and grazing cases named.

The bar: primitive IDs bit-exact, except
  * ties — the other answer's triangle is crossed at a t within 1e-6 relative of the oracle's t
    (two triangles at one distance: a shared edge or vertex); and
  * grazing flips — the ray passes within 1e-5 (relative barycentric margin) of the edge of the
    triangle the two answers disagree about, where float32 and float64 rounding decide differently;
and t within 1e-5 relative wherever the IDs agree.
"""
from __future__ import annotations

import numpy as np

MISS = 0xFFFFFFFF
TIE_REL = 1e-6
T_REL = 1e-5
EDGE_MARGIN = 1e-5


def plane_t_and_margin(pos: np.ndarray, tri: np.ndarray, rays: np.ndarray, prims: np.ndarray):
    """fp64 plane distance and smallest barycentric coordinate of each ray's line on triangle prims."""
    v = pos.astype(np.float64)[tri[prims].astype(np.int64)]  # [k, 3, 3]
    o = rays[:, 0:3].astype(np.float64)
    d = rays[:, 4:7].astype(np.float64)
    e1, e2 = v[:, 1] - v[:, 0], v[:, 2] - v[:, 0]
    n = np.cross(e1, e2)
    t = -np.einsum("ij,ij->i", o - v[:, 0], n) / np.einsum("ij,ij->i", d, n)
    p = o + d * t[:, None]
    # Barycentrics of p.
    q = p - v[:, 0]
    d00, d01, d11 = (e1 * e1).sum(1), (e1 * e2).sum(1), (e2 * e2).sum(1)
    d20, d21 = (q * e1).sum(1), (q * e2).sum(1)
    den = d00 * d11 - d01 * d01
    b1 = (d11 * d20 - d01 * d21) / den
    b2 = (d00 * d21 - d01 * d20) / den
    b0 = 1 - b1 - b2
    return t, np.minimum(np.minimum(b0, b1), b2)


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
        if g != MISS and o != MISS:
            tg, mg = plane_t_and_margin(pos, tri, r, np.array([g]))
            if abs(tg[0] - ora_t[i]) <= TIE_REL * abs(ora_t[i]) and mg[0] > -EDGE_MARGIN:
                ties += 1
                continue
            to, mo = plane_t_and_margin(pos, tri, r, np.array([o]))
            if mo[0] < EDGE_MARGIN and mg[0] > -EDGE_MARGIN:
                grazing += 1
                continue
        elif g != MISS:  # the GPU hit what the oracle missed
            _, mg = plane_t_and_margin(pos, tri, r, np.array([g]))
            if mg[0] > -EDGE_MARGIN:
                grazing += 1
                continue
        else:  # the oracle hit what the GPU missed
            _, mo = plane_t_and_margin(pos, tri, r, np.array([o]))
            if mo[0] < EDGE_MARGIN:
                grazing += 1
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
