"""Benchmark of the hip transport backend on one MI355X; prints ONE JSON line.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload rayquery]

Workload ``rayquery`` (batch ray queries on a BVH): 1,048,820-triangle displaced icosphere
(frequency 229, the nearest 20·n² count at or above 2^20), 8,388,608 seeded rays resident in HBM.
One step = one closest-hit launch plus one any-hit launch over the whole batch; ``value`` is
queries per second over both kinds.

roofline: the closest-hit traversal kernel, priced at the algorithmic bytes per query of the
reference algorithm (binary binned-SAH BVH, leaf 4, ordered traversal, src/owe/scene/bvh.hpp:31-78)
measured on this very batch by the oracle: 40 B of ray and hit + 32 B per node visited (50.685) +
36 B per triangle tested (7.060) = 1,916.2 B per query; its average duration comes from HIP events
around each launch on the launch stream. cpu_baseline: the reference's own MeshShape::intersect
(oracle/_ref, built from the reference sources) on the host threads over a bounded sample, or the
oracle restatement when that harness is absent.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "optical-pathtracer_amd"))

HBM_PEAK_GBS = 8000.0
RQ_NODES_PER_QUERY = 50.684757232666016  # slab tests per closest-hit query, reference algorithm, full batch
RQ_TRIS_PER_QUERY = 7.060239195823669    # triangle tests per closest-hit query, same
RQ_BYTES_PER_QUERY = 40 + 32 * RQ_NODES_PER_QUERY + 36 * RQ_TRIS_PER_QUERY
PROFILE_TRAFFIC = ROOT / "profiles" / "rayquery_closest_traffic.md"


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def measured_traffic(path: Path):
    """HBM bytes per launch recorded by the committed rocprofv3 --pmc summary, if present."""
    if not path.exists():
        return None
    m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", path.read_text())
    return float(m.group(1)) if m else None


def cpu_baseline_rayquery(pos, tri, rays, seconds_budget: float = 15.0):
    sys.path.insert(0, str(ROOT / "oracle"))
    import numpy as np
    from pyoracle import Oracle, Reference, reference_available

    threads = min(16, os.cpu_count() or 1)
    kind = "reference" if reference_available() else "port"
    lib = Reference() if kind == "reference" else Oracle()
    mesh = lib.mesh(pos.astype(np.float64), tri)
    n = 65536
    done, elapsed = 0, 0.0
    while elapsed < seconds_budget and done < len(rays):
        chunk = rays[done:done + n]
        t0 = time.perf_counter()
        mesh.closest(chunk, threads=threads)
        elapsed += time.perf_counter() - t0
        done += len(chunk)
        n = min(4 * n, 1 << 21)
    return {"value": done / elapsed, "unit": "rays/s", "cores": threads, "kind": kind,
            "sample": f"closest hit of the first {done} rays of the batch on the same mesh ({elapsed:.1f} s, "
                      f"fp64 MeshShape::intersect, {threads} threads)"}


def bench_rayquery(args):
    import numpy as np
    import torch

    from owe_hip import _lib, rayquery, workloads

    dev = torch.device("cuda:0")
    log("rayquery: generating mesh and rays")
    pos, tri = workloads.icosphere(n=229, seed=1)
    closest_np, any_np = workloads.query_rays(8 * 1024 * 1024, seed=2)
    n = len(closest_np)
    mesh = rayquery.Mesh(pos, tri, device=0)
    st = mesh.stats()
    log(f"rayquery: mesh {st.triangles} tris, {st.nodes} wide nodes, depth {st.depth}, stack {st.stack_bound}, "
        f"build {st.build_seconds:.2f}s")
    closest = torch.from_numpy(closest_np).to(dev)
    anyq = torch.from_numpy(any_np).to(dev)
    prim = torch.empty(n, dtype=torch.int32, device=dev)
    tt = torch.empty(n, dtype=torch.float32, device=dev)
    prim2 = torch.empty(n, dtype=torch.int32, device=dev)
    tt2 = torch.empty(n, dtype=torch.float32, device=dev)
    stream = torch.cuda.current_stream(dev)

    def step(ev=None):
        if ev is not None:
            ev[0].record(stream)
        mesh.trace(closest, _lib.QUERY_CLOSEST, prim, tt, stream)
        if ev is not None:
            ev[1].record(stream)
        mesh.trace(anyq, _lib.QUERY_ANY, prim2, tt2, stream)
        if ev is not None:
            ev[2].record(stream)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    events = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        step(events[k])
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    closest_ms = sum(e[0].elapsed_time(e[1]) for e in events) / args.steps
    any_ms = sum(e[1].elapsed_time(e[2]) for e in events) / args.steps
    hits = int((prim != -1).sum().item())
    log(f"rayquery: closest {closest_ms:.3f} ms, any {any_ms:.3f} ms per launch, wall {wall:.3f}s, hit {hits / n:.4f}")
    algo_bytes = RQ_BYTES_PER_QUERY * n
    achieved = algo_bytes / (closest_ms * 1e-3) / 1e9
    result = {
        "metric": "spectral paths/s at fixed rel. RMSE vs fp64 ref; adjoint grads/s; HBM GB/s",
        "value": 2 * n * args.steps / wall,
        "unit": "ray queries/s (closest + any hit)",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded icosphere mesh and rays)",
        "config": {"workload": "batch ray queries on a BVH: closest + any hit", "triangles": int(st.triangles),
                   "rays": n, "mesh": "icosphere frequency 229, radial value noise 0.02", "bvh": "4-wide SAH",
                   "closest_ms": closest_ms, "any_ms": any_ms,
                   "closest_rays_per_s": n / (closest_ms * 1e-3), "any_rays_per_s": n / (any_ms * 1e-3)},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(PROFILE_TRAFFIC),
                     "kernel": "traceBatchKernel<closest>", "algorithmic_bytes_per_launch": algo_bytes},
    }
    if not args.no_cpu_baseline:
        log("rayquery: timing the CPU baseline")
        result["cpu_baseline"] = cpu_baseline_rayquery(pos, tri, closest_np)
    else:
        result["cpu_baseline"] = None
    mesh.close()
    return result


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="rayquery", choices=["rayquery"])
    ap.add_argument("--no-cpu-baseline", action="store_true", help="skip the CPU baseline leg")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print(f"bench.py measures one GPU only (asked for {args.gpus}); every workload here runs on a single "
              "MI355X and no multi-GPU figure is produced.", file=sys.stderr)
        return 2
    result = bench_rayquery(args)
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
