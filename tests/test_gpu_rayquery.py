"""Batch ray queries on the GPU (owe_hip_trace_batch) against the fp64 oracle.

Parity bar (rq_check.py): primitive IDs bit-exact except documented ties (t within 1e-6 relative)
and grazing flips at triangle edges (barycentric margin below 1e-5); t within 1e-5 relative.
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

from owe_hip import _lib, rayquery, workloads
from rq_check import MISS, classify

pytestmark = pytest.mark.gpu


def run(mesh, rays_np, mode=_lib.QUERY_CLOSEST):
    rays = torch.from_numpy(np.ascontiguousarray(rays_np)).cuda()
    prim, t = mesh.trace(rays, mode)
    torch.cuda.synchronize()
    return prim.cpu().numpy().astype(np.uint32), t.cpu().numpy()


def assert_parity(res, max_bad=0, max_rel_t=1e-5):
    print(f"parity: {res['agree']}/{res['n']} agree, {res['ties']} ties, {res['grazing']} grazing, "
          f"{len(res['bad'])} bad, max rel t {res['max_rel_t']:.2e}", flush=True)
    assert len(res["bad"]) <= max_bad, f"unexplained mismatches at rays {res['bad'][:10]}"
    assert res["max_rel_t"] <= max_rel_t
    assert res["ties"] + res["grazing"] <= 1e-4 * res["n"] + 2


def test_closest_hit_matches_oracle_on_icosphere(cuda, oracle):
    pos, tri = workloads.icosphere(n=64, seed=1)
    rays, _ = workloads.query_rays(200_000, seed=21)
    mesh = rayquery.Mesh(pos, tri)
    st = mesh.stats()
    print(f"mesh: {st.triangles} tris, {st.nodes} nodes, depth {st.depth}, stack bound {st.stack_bound}", flush=True)
    gp, gt = run(mesh, rays)
    op, ot = oracle.mesh(pos, tri).closest(rays)
    res = classify(pos, tri, rays, gp, gt, op, ot)
    assert_parity(res)
    assert res["hits"] > 0.99 * len(rays)


def test_any_hit_agrees_with_oracle_segments(cuda, oracle):
    pos, tri = workloads.icosphere(n=48, seed=3)
    _, rays = workloads.query_rays(100_000, seed=22)
    mesh = rayquery.Mesh(pos, tri)
    gp, gt = run(mesh, rays, _lib.QUERY_ANY)
    _, ot = oracle.mesh(pos, tri).closest(rays)
    tmax = rays[:, 7].astype(np.float64)
    expect = ot < tmax
    got = gp != MISS
    disagree = np.nonzero(got != expect)[0]
    # Disagreement only where the oracle's nearest crossing sits at the segment's end (a tie).
    near_end = np.abs(ot[disagree] - tmax[disagree]) <= 1e-5 * tmax[disagree]
    assert near_end.all(), f"any-hit disagrees away from the segment end at {disagree[~near_end][:10]}"
    # Every reported any-hit crossing lies inside (tmin, tmax).
    assert (gt[got] > rays[got, 3]).all() and (gt[got] < rays[got, 7]).all()
    assert 0.5 < got.mean() < 1.0


def test_triangle_soup_axis_aligned_and_offset_rays(cuda, oracle):
    rng = np.random.default_rng(7)
    c = rng.uniform(0, 10, size=(3000, 3))
    pos = np.stack([c, c + rng.uniform(0, 0.5, (3000, 3)), c + rng.uniform(0, 0.5, (3000, 3))], axis=1).reshape(-1, 3)
    tri = np.arange(len(pos), dtype=np.uint32).reshape(-1, 3)
    pos = pos.astype(np.float32)
    n = 20000
    rays = np.zeros((n, 8), dtype=np.float32)
    rays[:, 0:3] = rng.uniform(-1, 11, (n, 3))
    d = rng.standard_normal((n, 3))
    # A third of the rays run along coordinate axes (zero direction components, infinite inverses).
    axis = rng.integers(0, 3, n)
    aligned = rng.uniform(size=n) < 0.33
    d[aligned] = 0
    d[aligned, axis[aligned]] = rng.choice([-1.0, 1.0], aligned.sum())
    rays[:, 4:7] = d / np.linalg.norm(d, axis=1, keepdims=True)
    rays[:, 3] = rng.uniform(0, 2, n) * (rng.uniform(size=n) < 0.5)  # some rays start their interval late
    rays[:, 7] = np.where(rng.uniform(size=n) < 0.5, np.inf, rng.uniform(1, 8, n))
    mesh = rayquery.Mesh(pos, tri)
    gp, gt = run(mesh, rays)
    op, ot = oracle.mesh(pos, tri).closest(rays)
    assert_parity(classify(pos, tri, rays, gp, gt, op, ot))


def test_degenerate_meshes_and_batches(cuda):
    rays, _ = workloads.query_rays(1000, seed=3)
    # No triangles: every ray misses.
    empty = rayquery.Mesh(np.zeros((0, 3), np.float32), np.zeros((0, 3), np.uint32))
    gp, gt = run(empty, rays)
    assert (gp == MISS).all() and np.isinf(gt).all()
    # A zero-area triangle is seen edge-on from every direction: never crossed.
    flat = rayquery.Mesh(np.array([[0, 0, 0], [1, 1, 1], [2, 2, 2]], np.float32), np.array([[0, 1, 2]], np.uint32))
    gp, _ = run(flat, rays)
    assert (gp == MISS).all()
    # An empty batch launches nothing and succeeds.
    mesh = rayquery.Mesh(*workloads.icosphere(n=4))
    z = torch.zeros((0, 8), dtype=torch.float32, device="cuda")
    p, t = mesh.trace(z)
    assert p.numel() == 0
    # Host memory is refused with a clear error, not read by the kernel.
    lib = _lib.load()
    host = np.zeros((4, 8), np.float32)
    st = lib.owe_hip_trace_batch(mesh._h, 0, host.ctypes.data, 4, host.ctypes.data, host.ctypes.data, None)
    assert st == _lib.ERR_INVALID and b"device memory" in lib.owe_hip_last_error()


def test_results_are_deterministic(cuda):
    pos, tri = workloads.icosphere(n=32, seed=2)
    rays, _ = workloads.query_rays(50_000, seed=4)
    mesh = rayquery.Mesh(pos, tri)
    a = run(mesh, rays)
    b = run(mesh, rays)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_full_size_batch(cuda, oracle):
    """The full workload: 1,048,820 triangles (icosphere frequency 229), 8,388,608 rays.

    Size-independent checks over every ray: any-hit equals (closest t < tmax) up to segment-end ties,
    and answers are deterministic; a seeded subset of 65,536 rays is checked against the oracle.
    """
    pos, tri = workloads.icosphere(n=229, seed=1)
    closest, any_rays = workloads.query_rays(8 * 1024 * 1024, seed=2)
    mesh = rayquery.Mesh(pos, tri)
    st = mesh.stats()
    print(f"full mesh: {st.triangles} tris, {st.nodes} nodes, depth {st.depth}, bound {st.stack_bound}, "
          f"build {st.build_seconds:.2f}s", flush=True)
    gp, gt = run(mesh, closest)
    ap, at = run(mesh, any_rays, _lib.QUERY_ANY)
    hit_any = ap != MISS
    tmax = any_rays[:, 7]
    expect = gt < tmax
    disagree = np.nonzero(hit_any != expect)[0]
    assert (np.abs(gt[disagree].astype(np.float64) - tmax[disagree]) <= 2e-5 * tmax[disagree]).all()
    assert (gp != MISS).mean() > 0.999
    sub = np.random.default_rng(0).choice(len(closest), 65536, replace=False)
    op, ot = oracle.mesh(pos, tri).closest(closest[sub])
    assert_parity(classify(pos, tri, closest[sub], gp[sub], gt[sub], op, ot))
    gp2, gt2 = run(mesh, closest)
    assert np.array_equal(gp, gp2) and np.array_equal(gt, gt2)


@pytest.mark.parametrize("n", [1, 777, 200_000])
def test_refill_form_returns_the_per_lane_forms_hits_bit_for_bit(cuda, n):
    """Persistent lanes with refill walk each ray as one lane per ray does: the same hits, any shape."""
    pos, tri = workloads.icosphere(n=64, seed=5)
    closest, any_rays = workloads.query_rays(n, seed=31)
    mesh = rayquery.Mesh(pos, tri)
    mesh.set_launch(rayquery.Mesh.PER_LANE)
    ref = [run(mesh, closest), run(mesh, any_rays, _lib.QUERY_ANY)]
    for bpc, refill in ((1, 1), (6, 16), (10, 64), (8, 32)):
        mesh.set_launch(rayquery.Mesh.REFILL, bpc, refill)
        assert mesh.launch() == mesh.launch(_lib.QUERY_ANY) == dict(form=1, blocks_per_cu=bpc, refill=refill)
        got = [run(mesh, closest), run(mesh, any_rays, _lib.QUERY_ANY)]
        for (rp, rt), (gp, gt) in zip(ref, got):
            assert np.array_equal(rp, gp) and np.array_equal(rt, gt), (bpc, refill)
    # Each mode keeps its own form.
    mesh.set_launch(rayquery.Mesh.PER_LANE, mode=_lib.QUERY_ANY)
    assert mesh.launch(_lib.QUERY_ANY)["form"] == 0 and mesh.launch(_lib.QUERY_CLOSEST)["form"] == 1
    assert [run(mesh, closest), run(mesh, any_rays, _lib.QUERY_ANY)][1][0].tolist() == ref[1][0].tolist()
    print(f"n {n}: refill form identical over 4 shapes", flush=True)
