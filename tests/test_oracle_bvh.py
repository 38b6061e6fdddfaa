"""The oracle's BVH and mesh queries against the reference and against brute force.

The reference's own check is "mesh BVH = brute force" on a seeded triangle soup
(tests/test_geometry.cpp:170-194); the fixture tests/golden/mesh_small.npz holds the reference's
MeshShape::intersect answers (primitive, t) and its BVH primitive order for an icosphere.
"""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

from conftest import GOLDEN
from owe_hip import workloads


def test_mesh_query_matches_reference_fixture(oracle):
    g = np.load(GOLDEN / "mesh_small.npz")
    m = oracle.mesh(g["positions"].astype(np.float64), g["triangles"])
    prim, t = m.closest(g["rays"])
    assert np.array_equal(prim, g["prim"])
    hit = np.isfinite(g["t"])
    assert np.array_equal(t[hit], g["t"][hit])  # bit-exact: same tree, same order, same arithmetic
    assert np.array_equal(m.order(), g["order"])


def test_mesh_bvh_matches_brute_force_on_triangle_soup(oracle):
    rng = np.random.default_rng(7)
    c = rng.uniform(0, 10, size=(3000, 3))
    tri_pos = np.stack([c, c + rng.uniform(0, 0.5, (3000, 3)), c + rng.uniform(0, 0.5, (3000, 3))], axis=1)
    pos = tri_pos.reshape(-1, 3)
    tri = np.arange(len(pos), dtype=np.uint32).reshape(-1, 3)
    m = oracle.mesh(pos, tri)
    n = 400
    rays = np.zeros((n, 8), dtype=np.float32)
    rays[:, 0:2] = rng.uniform(0, 10, (n, 2))
    rays[:, 2] = -1
    d = np.column_stack([rng.uniform(-0.5, 0.5, (n, 2)), np.ones(n)])
    rays[:, 4:7] = d / np.linalg.norm(d, axis=1, keepdims=True)
    rays[:, 7] = np.inf
    p, t = m.closest(rays)
    pb, tb = m.brute(rays)
    assert np.array_equal(np.isfinite(t), np.isfinite(tb))
    f = np.isfinite(t)
    assert np.allclose(t[f], tb[f], rtol=0, atol=1e-12)
    assert f.sum() > 100


def test_mesh_query_matches_reference_build(oracle, reference):
    pos, tri = workloads.icosphere(n=48, seed=4)
    rays, _ = workloads.query_rays(30000, seed=8)
    mo, mr = oracle.mesh(pos, tri), reference.mesh(pos, tri)
    assert mo.node_count() == mr.node_count()
    assert np.array_equal(mo.order(), mr.order())
    po, to = mo.closest(rays)
    pr, tr = mr.closest(rays)
    assert np.array_equal(po, pr)
    f = np.isfinite(tr)
    assert np.array_equal(to[f], tr[f])


def test_any_hit_agrees_with_closest_hit(oracle):
    pos, tri = workloads.icosphere(n=20, seed=2)
    _, rays = workloads.query_rays(20000, seed=9)
    m = oracle.mesh(pos, tri)
    hit = m.any(rays)
    _, t = m.closest(rays)
    assert np.array_equal(hit, t < rays[:, 7])
    assert 0.5 < hit.mean() < 1.0  # most targets lie inside the surface, a few beyond it


def test_icosphere_is_closed_outward_and_counts_20n2():
    for n in (1, 3, 17):
        pos, tri = workloads.icosphere(n=n, seed=1)
        assert len(tri) == 20 * n * n
        assert len(pos) == 10 * n * n + 2
        e = np.sort(np.concatenate([tri[:, [0, 1]], tri[:, [1, 2]], tri[:, [2, 0]]]), axis=1)
        _, cnt = np.unique(e, axis=0, return_counts=True)
        assert (cnt == 2).all()  # every edge shared by exactly two triangles: watertight
        c = pos[tri].mean(axis=1)
        nrm = np.cross(pos[tri[:, 1]] - pos[tri[:, 0]], pos[tri[:, 2]] - pos[tri[:, 0]])
        assert ((nrm * c).sum(axis=1) > 0).all()
        r = np.linalg.norm(pos, axis=1)
        assert r.min() >= 0.98 - 1e-6 and r.max() <= 1.02 + 1e-6


def test_value_noise_matches_reference_fixture():
    import json

    g = json.loads((GOLDEN / "formulas.json").read_text())["value_noise3"]
    pts = np.array([row[:3] for row in g])
    expect = np.array([row[3] for row in g])
    assert np.array_equal(workloads.value_noise3(pts), expect)


def test_query_rays_geometry():
    closest, any_hit = workloads.query_rays(5000, seed=1)
    o = closest[:, 0:3].astype(np.float64)
    assert np.allclose(np.linalg.norm(o, axis=1), 3.0, rtol=1e-6)
    d = closest[:, 4:7].astype(np.float64)
    target = o + d * any_hit[:, 7:8].astype(np.float64)
    assert np.linalg.norm(target, axis=1).max() <= 1.0 + 1e-5
    assert (closest[:, 3] == 0).all() and np.isinf(closest[:, 7]).all()
