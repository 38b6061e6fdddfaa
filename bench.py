"""Benchmark of the hip transport backend on one MI355X; prints ONE JSON line.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload prism|rayquery|light|lens|caustic|telescope]

Workload ``prism`` (the default; plain spectral path tracing of a dispersion prism): an equilateral
Sellmeier-glass prism (40 mm sides, 20 mm tall) standing on a 1 m² Lambertian floor of albedo 0.8
(its base face in the floor's plane) in air n = 1.000277, a 5778 K blackbody sun of 0.53°; a 256×256
pinhole camera, 64 spp, one wavelength per path uniform over 380–780 nm, max depth 16, Russian
roulette from depth 4. One step = one pass of 64 spp = 4,194,304 camera paths. Both forms run — the
wavefront kernels and the persistent batched kernel replayed from a captured hipGraph, each
autotuned first — and ``value`` is the faster form's spectral paths per second (the films of the two
forms are compared bit for bit). Beside it: the same scene with the prism lifted 0.2 mm (no
contact), and the quality of the images against the fp64 oracle's 16,384-spp image
(tests/golden/prism_reference.npz): rel. RMSE as the reference's benchmark runner scores it, for the
timed image and for fresh images of 64 to 4,096 spp. roofline: the persistent kernel is bound by
VALU issue and lane divergence, so it is priced in useful lane-instructions per second (profiled
VALU instructions per launch × 64 × active-lane fraction, over the live launch time) against the
chip's VALU issue peak; the SoA byte model and the measured HBM bytes sit beside it. cpu_baseline:
the oracle's fp64 restatement of Tracer::radiance (the reference's Tracer needs stb_image and cannot
be built here) on the host threads, full frames of the same scene and spp.

Beside it: ``rayquery`` (below), ``light`` (light tracing of the same prism scene, 256×256×64
particles per pass), ``lens`` (256 seeded achromats × 7 wavelengths in one persistent launch) and
``caustic`` (a gradient step of the caustic slab: 1,048,576 particles forward, then the path-replay
adjoint into the 256×256 heights) and ``telescope`` (64 perturbed Gregorian telescopes, the angular
RMS of each afocal spot and its derivatives with respect to 8 prescription parameters).

Workload ``rayquery`` (batch ray queries on a BVH): 1,048,820-triangle displaced icosphere
(frequency 229, the nearest 20·n² count at or above 2^20), 8,388,608 seeded rays resident in HBM.
One step = one closest-hit launch plus one any-hit launch over the whole batch; ``value`` is
queries per second over both kinds.

roofline: the closest-hit traversal kernel, priced at the algorithmic bytes per query of the
reference algorithm (binary binned-SAH BVH, leaf 4, ordered traversal, src/owe/scene/bvh.hpp:31-78)
measured on this very batch by the oracle: 40 B of ray and hit + 32 B per node visited (50.685) +
36 B per triangle tested (7.060) = 1,916.2 B per query; its average duration comes from HIP events
around each launch on the launch stream. cpu_baseline: the oracle's fp64 restatement of
MeshShape::intersect (its BVH equals the reference's node for node) on the host threads over a
bounded sample of the batch.
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
PROFILE_PRISM_TRAFFIC = ROOT / "profiles" / "prism_camera_paths_traffic.md"
PROFILE_LIGHT_TRAFFIC = ROOT / "profiles" / "light_particles_traffic.md"
PROFILE_LENS_TRAFFIC = ROOT / "profiles" / "lens_batch_traffic.md"
PROFILE_VCM = ROOT / "profiles" / "vcm_kernels.md"
PROFILE_CAUSTIC_TRAFFIC = ROOT / "profiles" / "caustic_adjoint_traffic.md"
PROFILE_COUNTERS = ROOT / "profiles" / "occupancy_valu_l2_counters.md"


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def measured_traffic(path: Path, kernel: str | None = None):
    """HBM bytes per launch of `kernel` recorded by the committed rocprofv3 --pmc summary
    (a line `traffic_bytes_per_launch[kernel]: B`), if present."""
    if not path.exists():
        return None
    text = path.read_text()
    if kernel is not None:
        m = re.search(r"traffic_bytes_per_launch\[" + re.escape(kernel) + r"\]:\s*([0-9.eE+]+)", text)
        return float(m.group(1)) if m else None
    m = re.search(r"traffic_bytes_per_launch(?:\[[^\]]*\])?:\s*([0-9.eE+]+)", text)
    return float(m.group(1)) if m else None


def measured_counters(kernel: str):
    """Occupancy, VALU and L2 figures of `kernel` recorded by a committed rocprofv3 counter summary
    under profiles/ (a line `counters[kernel]: name=value ...`), if present."""
    for path in sorted((ROOT / "profiles").glob("*.md")):
        m = re.search(r"counters\[" + re.escape(kernel) + r"\]:([^\n]*)", path.read_text())
        if m:
            return {k: float(v) for k, v in re.findall(r"(\w+)=([0-9.eE+-]+)", m.group(1))}
    return None


def cpu_baseline_rayquery(pos, tri, rays, seconds_budget: float = 15.0):
    sys.path.insert(0, str(ROOT / "oracle"))
    import numpy as np
    from pyoracle import Oracle

    threads = min(16, os.cpu_count() or 1)
    kind = "port"
    mesh = Oracle().mesh(pos.astype(np.float64), tri)
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
                      f"the oracle's fp64 MeshShape::intersect, {threads} threads)"}


PRISM_SPP = 64
PRISM_SIZE = 256
# Algorithmic bytes per live path of each wavefront kernel (path state 64 B, hit 16 B, shadow 32 B).
KERNEL_BYTES = {"generate": 64, "intersect": 32 + 16, "shade": 2 * 64 + 16 + 32, "shadow": 32, "splat": 24 + 64}


def cpu_baseline_prism(sc, st, seconds_budget: float = 12.0):
    sys.path.insert(0, str(ROOT / "oracle"))
    from pyoracle import Oracle, World

    threads = min(16, os.cpu_count() or 1)
    w = World(Oracle(), sc)
    frames, elapsed = 0, 0.0
    while elapsed < seconds_budget and frames < 64:
        t0 = time.perf_counter()
        w.render(st, PRISM_SPP, passes=1, layout=0, threads=threads, first_pass=frames)
        elapsed += time.perf_counter() - t0
        frames += 1
    paths = frames * PRISM_SIZE * PRISM_SIZE * PRISM_SPP
    return {"value": paths / elapsed, "unit": "spectral paths/s", "cores": threads, "kind": "port",
            "sample": f"{frames} full frames ({PRISM_SIZE}x{PRISM_SIZE}, {PRISM_SPP} spp, {paths} paths, {elapsed:.1f} s) "
                      f"of the same scene through the oracle's fp64 Tracer::radiance, {threads} threads"}


def _time_passes(r, steps, warmup, events_in_timed: bool):
    """Wall time of `steps` passes after `warmup`, the renderer's statistics and image over them,
    and per-kernel HIP-event times. With events_in_timed the events run inside the timed passes
    (one pair per pass for the persistent form's single launch); otherwise — the wavefront form,
    whose events would sit between its many small kernels — the same passes are repeated with
    events after the timed region and the breakdown comes from that repeat."""
    import torch

    for _ in range(warmup):
        r.run_pass(PRISM_SPP)
    r.reset()
    r.set_timing(events_in_timed)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        r.run_pass(PRISM_SPP)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    stats, img = r.stats(), r.resolve()
    if not events_in_timed:
        r.reset()
        r.set_timing(True)
        for _ in range(steps):
            r.run_pass(PRISM_SPP)
        torch.cuda.synchronize()
    times = r.kernel_times()
    r.set_timing(False)
    return wall, stats, times, img


PRISM_REFERENCE = ROOT / "tests" / "golden" / "prism_reference.npz"
# gfx950 VALU issue peak: 256 CUs x 4 SIMDs x 16 lanes per clock at 2.4 GHz, one lane-instruction
# per lane per clock (packed FP32 is disabled in every build here).
VALU_PEAK_TLANE = 256 * 4 * 16 * 2.4e9 / 1e12


def prism_quality(sc, st, timed_img, timed_spp):
    """Relative error of this backend's prism images against the fp64 oracle's 16,384-spp image
    (tests/golden/prism_reference.npz), as the reference's benchmark runner scores a render:
    relMSE over pixels and linear-sRGB channels (docs/BENCHMARKS.md:94-101), with and without the
    sun-glint mask; rel. RMSE = sqrt(relMSE). Images of 64·4^k spp come from fresh renderers on
    seeds no other measurement uses."""
    import numpy as np

    from owe_hip import output as O
    from owe_hip import render

    if not PRISM_REFERENCE.exists():
        return None
    ref = np.load(PRISM_REFERENCE)
    rgb = lambda xyz: np.asarray(xyz, np.float64) @ O.XYZ_TO_LINEAR_SRGB.T  # noqa: E731
    ref_rgb = rgb(ref["xyz"])
    mask = O.glint_mask(ref_rgb)

    def score(img):
        e = O.benchmark_errors(rgb(img), ref_rgb, mask)
        full, masked = e["rel_mse"], e["rel_mse_no_glints"]
        return {"rel_rmse": None if full is None else full ** 0.5,
                "rel_rmse_no_glints": None if masked is None else masked ** 0.5, "mean_ratio": e["mean_ratio"]}

    ladder = {}
    for k, spp in enumerate((64, 256, 1024, 4096)):
        st.seed = 500 + k
        r = render.Renderer(sc, st)
        for _ in range(spp // PRISM_SPP):
            r.run_pass(PRISM_SPP)
        ladder[str(spp)] = score(r.resolve())
        r.close()
    reach = [int(k) for k, v in ladder.items() if v["rel_rmse_no_glints"] is not None and v["rel_rmse_no_glints"] <= 0.02]
    return {"reference": f"fp64 oracle, {int(ref['spp'])} spp, seed {int(ref['seed'])}",
            "glint_pixels": int(mask.sum()), "timed_image": dict(spp=timed_spp, **score(timed_img)),
            "by_spp": ladder, "spp_for_2pct_no_glints": min(reach) if reach else None}


def _prism_forms(sc, st, args, forms_wanted=("wavefront", "persistent")):
    from owe_hip import render

    forms = {}
    # The plain wavefront form, then the persistent batched form with its autotuned launch shape
    # and graph replay; the two images are equal bit for bit (checked on the timed passes).
    for form in forms_wanted:
        r = render.Renderer(sc, st)
        if form == "persistent":
            r.set_mode(render.Renderer.PERSISTENT, use_graph=True)
        tuned_ms = r.autotune(PRISM_SPP)
        wall, stats, times, img = _time_passes(r, args.steps, args.warmup, events_in_timed=form == "persistent")
        forms[form] = dict(wall=wall, stats=stats, times=times, mode=r.mode(), tuned_ms=tuned_ms, img=img)
        paths = stats["paths"]
        log(f"prism {form}: {args.steps} passes in {wall:.3f}s; {paths / wall / 1e6:.1f} M paths/s; "
            f"E_seg {stats['segments'] / max(1, paths):.3f}; shape {r.mode()}; inconsistencies "
            f"{stats['inconsistencies']}, leaks {stats['leaks']}; "
            + ", ".join(f"{k} {ms / args.steps:.3f} ms/pass ({n})" for k, (ms, n) in times.items() if n))
        r.close()
    return forms


def bench_prism(args):
    from owe_hip import scene as S

    log("prism: building scene")
    # The prism stands on the floor (its base face lies in the floor's plane: contact resolution).
    sc = S.dispersion_prism_scene(PRISM_SIZE, PRISM_SIZE, lift=0.0)
    st = S.settings(seed=1, max_depth=16, rr_depth=4, spectral=S.SPECTRAL_UNIFORM1, lambda_min=380.0, lambda_max=780.0)
    forms = _prism_forms(sc, st, args)
    same = bool((forms["wavefront"]["img"] == forms["persistent"]["img"]).all())
    best = max(forms, key=lambda f: forms[f]["stats"]["paths"] / forms[f]["wall"])
    F = forms[best]
    wall, stats, times = F["wall"], F["stats"], F["times"]
    paths = stats["paths"]
    eseg = stats["segments"] / max(1, paths)
    pass_bytes = paths * (eseg * 224 + 24)
    if best == "persistent":
        dominant = "persistent"
        dom_ms, dom_launches = times["persistent"]
        dom_bytes = pass_bytes
    else:
        dominant = max(("generate", "intersect", "shade", "shadow", "splat"), key=lambda k: times[k][0])
        dom_ms, dom_launches = times[dominant]
        live = stats["segments"] if dominant in ("intersect", "shade", "shadow") else paths
        dom_bytes = KERNEL_BYTES[dominant] * live
    total_ms = sum(ms for ms, _ in times.values())
    launch_ms = dom_ms / max(1, dom_launches)
    model_GBs = dom_bytes / (dom_ms * 1e-3) / 1e9
    kernel = f"{dominant}Kernel"
    cnt = measured_counters(kernel)
    traffic = measured_traffic(PROFILE_PRISM_TRAFFIC, kernel)
    # The kernel's bound is its VALU issue (counters: busy VALU, half the lanes active), priced in
    # useful lane-instructions per second: the profiled VALU instructions per launch x 64 lanes x
    # the active-lane fraction, over the live launch time.
    roofline = {"bound": "valu", "kernel": kernel, "unit": "T lane-instructions/s", "peak": VALU_PEAK_TLANE,
                "achieved": None, "frac": None, "traffic": traffic, "counters": cnt}
    if cnt and "valu_insts" in cnt and "lane_utilisation" in cnt:
        lanes = cnt["valu_insts"] * 64 * cnt["lane_utilisation"]
        roofline["achieved"] = lanes / (launch_ms * 1e-3) / 1e12
        roofline["frac"] = roofline["achieved"] / VALU_PEAK_TLANE
        roofline["issue_frac"] = cnt["valu_insts"] * 64 / (launch_ms * 1e-3) / 1e12 / VALU_PEAK_TLANE
    roofline["hbm_model"] = {"note": "SoA path-record byte model (224 B per segment + 24 B per path); the "
                                     "persistent kernel keeps path state in registers, so this prices traffic "
                                     "it does not make", "achieved_GBs": model_GBs, "peak_GBs": HBM_PEAK_GBS,
                             "frac": model_GBs / HBM_PEAK_GBS,
                             "algorithmic_bytes_per_launch": dom_bytes / max(1, dom_launches)}
    if traffic:
        roofline["hbm_measured"] = {"bytes_per_launch": traffic,
                                    "achieved_GBs": traffic / (launch_ms * 1e-3) / 1e9,
                                    "frac": traffic / (launch_ms * 1e-3) / 1e9 / HBM_PEAK_GBS}
    roofline["pass"] = {"kernel_ms": total_ms / args.steps, "launch_ms": launch_ms}
    img = F["img"]
    quality = None
    if not getattr(args, "no_quality", False):
        log("prism: image quality against the fp64 reference")
        quality = prism_quality(sc, S.settings(seed=1, max_depth=16, rr_depth=4, spectral=S.SPECTRAL_UNIFORM1),
                                img, args.steps * PRISM_SPP)
        log(f"prism quality: {quality}")
    lifted = None
    if not getattr(args, "no_lifted", False):
        # The same scene with the prism lifted 0.2 mm off the floor (no contact), persistent form.
        log("prism: the lifted variant")
        lifted_sc = S.dispersion_prism_scene(PRISM_SIZE, PRISM_SIZE, lift=0.0002)
        lifted = _prism_forms(lifted_sc, st, args, ("persistent",))["persistent"]
    result = {
        "metric": "spectral paths/s at fixed rel. RMSE vs fp64 ref; adjoint grads/s; HBM GB/s",
        "value": paths / wall,
        "unit": "spectral paths/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (procedural dispersion-prism scene, seeded streams)",
        "config": {"workload": f"plain spectral path tracing, dispersion prism standing on the floor ({best} form)",
                   "resolution": [PRISM_SIZE, PRISM_SIZE], "spp": PRISM_SPP, "lift_m": 0.0,
                   "paths_per_step": PRISM_SIZE * PRISM_SIZE * PRISM_SPP, "max_depth": 16, "rr_depth": 4,
                   "spectral": "1 wavelength per path, uniform 380-780 nm", "segments_per_path": eseg,
                   "kernel_ms_per_step": {k: ms / args.steps for k, (ms, n) in times.items() if n},
                   "inconsistencies": stats["inconsistencies"], "leaks": stats["leaks"],
                   "mean_Y": float(img[..., 1].mean()),
                   "forms": {f: {"paths_per_s": forms[f]["stats"]["paths"] / forms[f]["wall"],
                                 "ms_per_step": forms[f]["wall"] * 1e3 / args.steps, "launch": forms[f]["mode"],
                                 "autotuned_pass_ms": forms[f]["tuned_ms"]} for f in forms},
                   "persistent_over_wavefront": (forms["persistent"]["stats"]["paths"] / forms["persistent"]["wall"]) /
                                                (forms["wavefront"]["stats"]["paths"] / forms["wavefront"]["wall"]),
                   "forms_bit_identical": same,
                   "quality": quality,
                   "lifted_prism": None if lifted is None else {
                       "lift_m": 0.0002, "paths_per_s": lifted["stats"]["paths"] / lifted["wall"],
                       "ms_per_step": lifted["wall"] * 1e3 / args.steps, "launch": lifted["mode"],
                       "inconsistencies": lifted["stats"]["inconsistencies"], "leaks": lifted["stats"]["leaks"]}},
        "roofline": roofline,
    }
    result["cpu_baseline"] = None if args.no_cpu_baseline else cpu_baseline_prism(sc, st)
    return result


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

    # Launch forms: one lane per ray, and persistent lanes with refill over a sweep of shapes; the
    # fastest closest + any pair is kept. Both forms must return the same hits bit for bit.
    def pair_ms(reps=3):
        for _ in range(2):
            step()
        ev = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(reps)]
        for k in range(reps):
            step(ev[k])
        torch.cuda.synchronize()
        return (sum(e[0].elapsed_time(e[1]) for e in ev) / reps, sum(e[1].elapsed_time(e[2]) for e in ev) / reps)

    mesh.set_launch(rayquery.Mesh.PER_LANE)
    per_lane = pair_ms()
    ref_prim, ref_t = prim.clone(), tt.clone()
    ref_prim2, ref_t2 = prim2.clone(), tt2.clone()
    shapes = []
    for bpc in (6, 8, 10):
        for refill in (1, 8, 16, 32):
            mesh.set_launch(rayquery.Mesh.REFILL, bpc, refill)
            shapes.append((pair_ms(), bpc, refill))
    # Each query mode keeps its own fastest form and shape.
    launch = {}
    refill_ms = [0.0, 0.0]
    for q, name in ((_lib.QUERY_CLOSEST, "closest"), (_lib.QUERY_ANY, "any")):
        ms, bpc, refill = min(shapes, key=lambda x: x[0][q])
        refill_ms[q] = ms[q]
        if ms[q] < per_lane[q]:
            mesh.set_launch(rayquery.Mesh.REFILL, bpc, refill, mode=q)
        else:
            mesh.set_launch(rayquery.Mesh.PER_LANE, mode=q)
        launch[name] = mesh.launch(q) | {"refill_ms": ms[q], "refill_blocks_per_cu": bpc, "refill_refill": refill}
    step()
    torch.cuda.synchronize()
    identical = bool(torch.equal(prim, ref_prim) and torch.equal(tt, ref_t) and torch.equal(prim2, ref_prim2)
                     and torch.equal(tt2, ref_t2))
    log(f"rayquery: per lane closest {per_lane[0]:.3f} + any {per_lane[1]:.3f} ms; best refill closest "
        f"{refill_ms[0]:.3f} + any {refill_ms[1]:.3f} ms; identical {identical}; using {launch}")
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
    kernel = "traceRefillKernel" if launch["closest"]["form"] == rayquery.Mesh.REFILL else "traceBatchKernel"
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
                   "closest_rays_per_s": n / (closest_ms * 1e-3), "any_rays_per_s": n / (any_ms * 1e-3),
                   "launch": launch,
                   "forms": {"per_lane": {"closest_ms": per_lane[0], "any_ms": per_lane[1]},
                             "refill_best": {"closest_ms": refill_ms[0], "any_ms": refill_ms[1]},
                             "bit_identical": identical}},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(PROFILE_TRAFFIC, f"{kernel}<closest>"),
                     "counters": measured_counters(f"{kernel}<closest>"),
                     "kernel": f"{kernel}<closest>", "algorithmic_bytes_per_launch": algo_bytes},
    }
    if not args.no_cpu_baseline:
        log("rayquery: timing the CPU baseline")
        result["cpu_baseline"] = cpu_baseline_rayquery(pos, tri, closest_np)
    else:
        result["cpu_baseline"] = None
    mesh.close()
    return result


def cpu_baseline_light(sc, st, seconds_budget: float = 10.0):
    sys.path.insert(0, str(ROOT / "oracle"))
    from pyoracle import Oracle, World

    threads = min(16, os.cpu_count() or 1)
    w = World(Oracle(), sc)
    frames, elapsed = 0, 0.0
    while elapsed < seconds_budget and frames < 64:
        t0 = time.perf_counter()
        w.render_particles(st, PRISM_SPP, passes=1, threads=threads, first_pass=frames)
        elapsed += time.perf_counter() - t0
        frames += 1
    n = frames * PRISM_SIZE * PRISM_SIZE * PRISM_SPP
    return {"value": n / elapsed, "unit": "particles/s", "cores": threads, "kind": "port",
            "sample": f"{frames} full passes ({PRISM_SIZE}x{PRISM_SIZE}x{PRISM_SPP} particles, {n} particles, {elapsed:.1f} s) "
                      f"of the same scene through the oracle's fp64 particle restatement, {threads} threads"}


def bench_light(args):
    """Light tracing of the prism scene: 256×256×64 particles per pass from the sun, connected to
    the eye from every diffuse vertex, splatted exactly — in both forms of the particle pass: one
    thread per particle (the reference's form) and the lane-refill form (autotuned), which must give
    the same image bit for bit; the faster is the leg's value."""
    import numpy as np
    import torch

    from owe_hip import render
    from owe_hip import scene as S

    sc = S.dispersion_prism_scene(PRISM_SIZE, PRISM_SIZE)
    st = S.settings(seed=1, max_depth=16, rr_depth=4, spectral=S.SPECTRAL_UNIFORM1, lambda_min=380.0, lambda_max=780.0,
                    integrator=S.INTEGRATOR_LIGHT)
    forms = {}
    images = {}
    for name, mode in (("per_particle", render.Renderer.WAVEFRONT), ("refill", render.Renderer.PERSISTENT)):
        r = render.Renderer(sc, st)
        r.set_mode(mode)
        if mode == render.Renderer.PERSISTENT:
            r.autotune(PRISM_SPP)
        for _ in range(args.warmup):
            r.run_pass(PRISM_SPP)
        r.reset()
        r.set_timing(True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            r.run_pass(PRISM_SPP)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        stats = r.stats()
        ms, _ = r.kernel_times()["shade"]
        images[name] = r.resolve()
        forms[name] = {"particles_per_s": stats["paths"] / wall, "ms_per_step": wall * 1e3 / args.steps,
                       "kernel_ms_per_step": ms / args.steps, "stats": stats, "launch": r.mode()}
        log(f"light ({name}): {args.steps} passes in {wall:.3f}s; {stats['paths'] / wall / 1e6:.1f} M particles/s; "
            f"segments/particle {stats['segments'] / max(1, stats['paths']):.3f}; kernel {ms / args.steps:.3f} ms/pass; "
            f"shape {r.mode()}")
        r.close()
    best = max(forms, key=lambda k: forms[k]["particles_per_s"])
    f = forms[best]
    stats = f["stats"]
    n = stats["paths"]
    kernel = "particleRefillKernel" if best == "refill" else "particleKernel"
    # Algorithmic HBM bytes: each particle vertex connects to the eye at most once and deposits three
    # 64-bit bin adds (24 B); the particle's own state never leaves registers.
    splat_bytes = 24.0 * stats["segments"] / args.steps
    k_ms = f["kernel_ms_per_step"]
    hbm = {"bound": "hbm", "achieved": splat_bytes / (k_ms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": splat_bytes / (k_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
           "traffic": measured_traffic(PROFILE_LIGHT_TRAFFIC, kernel), "kernel": kernel,
           "algorithmic_bytes_per_launch": splat_bytes}
    roofline = valu_roofline(kernel, k_ms)
    if roofline:
        roofline["traffic"] = hbm["traffic"]
        roofline["hbm_model"] = hbm
    else:
        roofline = hbm
    out = {"value": f["particles_per_s"], "unit": "particles/s", "ms_per_step": f["ms_per_step"],
           "config": {"workload": "light tracing (particles), dispersion prism", "resolution": [PRISM_SIZE, PRISM_SIZE],
                      "particles_per_step": PRISM_SIZE * PRISM_SIZE * PRISM_SPP, "max_depth": 16, "rr_depth": 4,
                      "segments_per_particle": stats["segments"] / max(1, n), "form": best,
                      "kernel_ms_per_step": k_ms, "mean_Y": float(images[best][..., 1].mean()),
                      "forms": {k: {kk: v[kk] for kk in ("particles_per_s", "ms_per_step", "kernel_ms_per_step", "launch")}
                                for k, v in forms.items()},
                      "refill_over_per_particle": forms["refill"]["particles_per_s"] / forms["per_particle"]["particles_per_s"],
                      "forms_bit_identical": bool(np.array_equal(images["per_particle"], images["refill"]))},
           "roofline": roofline}
    out["cpu_baseline"] = None if args.no_cpu_baseline else cpu_baseline_light(sc, st)
    return out


def valu_roofline(kernel: str, launch_ms: float):
    """The VALU roofline of `kernel` from its profiled counters (VALU instructions per launch and
    active-lane fraction) over the live launch time, or None without a profile line."""
    cnt = measured_counters(kernel)
    if not cnt or "valu_insts" not in cnt or "lane_utilisation" not in cnt or not launch_ms > 0:
        return None
    achieved = cnt["valu_insts"] * 64 * cnt["lane_utilisation"] / (launch_ms * 1e-3) / 1e12
    return {"bound": "valu", "kernel": kernel, "unit": "T lane-instructions/s", "peak": VALU_PEAK_TLANE,
            "achieved": achieved, "frac": achieved / VALU_PEAK_TLANE,
            "issue_frac": cnt["valu_insts"] * 64 / (launch_ms * 1e-3) / 1e12 / VALU_PEAK_TLANE, "counters": cnt}


LENS_PLANES = 5  # through-focus grids per lens, 1e-3 x EFL apart about the paraxial image


def bench_lens(args):
    """Batched lens analysis: 256 seeded cemented achromats, each tracing its log-uniform number of
    rays (4,096 to 1,048,576) at all seven Fraunhofer wavelengths in one persistent launch, into a
    through-focus stack of five 128 x 128 grids per lens."""
    from owe_hip import lens as L

    lenses = L.seeded_achromats(256)
    traces = 7 * sum(p.rays for p in lenses)
    b = L.LensBatch(lenses, L.c_settings(planes=LENS_PLANES))
    tuned = b.autotune()
    for _ in range(args.warmup):
        b.run()
    t0 = time.perf_counter()
    kernel_ms = 0.0
    for _ in range(args.steps):
        b.run()
        kernel_ms += b.last_ms()
    wall = time.perf_counter() - t0
    per, _, _ = b.results()
    arrived = sum(r["arrived"] for r in per)
    log(f"lens: {args.steps} runs in {wall:.3f}s; {traces * args.steps / wall / 1e9:.2f} G ray traces/s; "
        f"kernel {kernel_ms / args.steps:.2f} ms/run; arrived {arrived / traces:.3f}")
    # The grids' bin adds are the algorithmic HBM traffic: 8 B per arrived ray and plane (the
    # per-wave bin tables turn most of them into LDS adds).
    hist_bytes = 8.0 * arrived * LENS_PLANES
    k_ms = kernel_ms / args.steps
    model = {"bound": "hbm", "achieved": hist_bytes / (k_ms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
             "frac": hist_bytes / (k_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
             "traffic": measured_traffic(PROFILE_LENS_TRAFFIC, "lensKernel"),
             "kernel": "lensKernel", "algorithmic_bytes_per_launch": hist_bytes}
    roofline = valu_roofline("lensKernel", k_ms)
    if roofline is None:
        roofline = model
    else:
        roofline["traffic"] = model["traffic"]
        roofline["hbm_model"] = model
    out = {"value": traces * args.steps / wall, "unit": "ray traces/s (lens x wavelength x ray)",
           "ms_per_step": wall * 1e3 / args.steps,
           "config": {"workload": "batched lens analysis, 256 seeded cemented achromats, 7 wavelengths, 0.5 deg field",
                      "lenses": len(lenses), "traces_per_step": traces, "arrived_fraction": arrived / traces,
                      "focus_planes": LENS_PLANES, "grid_pixels": 128,
                      "kernel_ms_per_step": k_ms, "autotuned_ms": tuned,
                      "rays_min": min(p.rays for p in lenses), "rays_max": max(p.rays for p in lenses)},
           "roofline": roofline}
    b.close()
    if not args.no_cpu_baseline:
        sys.path.insert(0, str(ROOT / "oracle"))
        from pyoracle import LensOracle, Oracle

        threads = min(16, os.cpu_count() or 1)
        sample = lenses[:4]
        lo = LensOracle(Oracle(), sample)
        t0 = time.perf_counter()
        for li in range(len(sample)):
            for lam in L.FRAUNHOFER7:
                lo.moments(li, lam, threads=threads)
        el = time.perf_counter() - t0
        n = 7 * sum(p.rays for p in sample)
        out["cpu_baseline"] = {"value": n / el, "unit": "ray traces/s", "cores": threads, "kind": "port",
                               "sample": f"the first 4 lenses ({n} traces, {el:.1f} s) through the oracle's fp64 "
                                         f"sequential trace, {threads} threads"}
    return out


def bench_caustic(args):
    """Differentiable caustic design: one gradient step of L = mean((E − target)²) over a 256 × 256
    height field — 1,048,576 particles traced forward into the 512 × 512 wall grid, then replayed
    backward (path-replay adjoint) into ∂L/∂heights."""
    import torch

    from owe_hip import caustic as C

    s = C.caustic_settings()
    c = C.Caustic(s)
    H = torch.tensor(C.initial_heights(), device="cuda")
    T = torch.tensor(C.target(s), device="cuda")
    E = torch.empty((s.pixels, s.pixels), dtype=torch.float32, device="cuda")
    g = torch.empty_like(H)

    def step(k):
        c.forward_into(H, E, k)
        gE = (2.0 / E.numel()) * (E - T)
        c.backward_into(H, gE, g, k)

    for k in range(args.warmup):
        step(k)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    torch.cuda.synchronize()
    ev[0].record()
    for k in range(args.steps):
        c.forward_into(H, E, k)
    ev[1].record()
    gE = ((2.0 / E.numel()) * (E - T)).contiguous()
    for k in range(args.steps):
        c.backward_into(H, gE, g, k)
    ev[2].record()
    torch.cuda.synchronize()
    f_ms, b_ms = ev[0].elapsed_time(ev[1]) / args.steps, ev[1].elapsed_time(ev[2]) / args.steps
    t0 = time.perf_counter()
    for k in range(args.steps):
        step(k)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    n = s.particles
    log(f"caustic: {args.steps} gradient steps in {wall:.3f}s; forward {f_ms:.3f} ms, adjoint {b_ms:.3f} ms "
        f"(ratio {b_ms / f_ms:.2f}); {n * args.steps / wall / 1e6:.1f} M particles/s forward+adjoint")
    # Algorithmic HBM traffic of the adjoint: 16 float atomics of 4 B per particle into the heights'
    # gradient, ∂L/∂E read at 9 pixels (4 B each) — the heights and ∂L/∂E are L2-resident.
    adj_bytes = n * (16 * 4 + 9 * 4)
    # The metric's "adjoint grads/s": whole gradients (all 65,536 height partials, each a sum over
    # the pass's 1,048,576 particles) per second, forward pass included.
    out = {"value": args.steps / wall, "unit": "gradients/s (forward + adjoint, 65,536 heights each)",
           "ms_per_step": wall * 1e3 / args.steps,
           "config": {"workload": "caustic design gradient step, 100 mm PMMA slab, 256x256 heights, 512x512 wall",
                      "particles_per_step": n, "nodes": s.nodes, "pixels": s.pixels,
                      "height_partials_per_s": s.nodes * s.nodes * args.steps / wall,
                      "particles_per_s": n * args.steps / wall,
                      "entry_sampling": "stratified over square tiles of the entry face",
                      "forward_ms": f_ms, "adjoint_ms": b_ms, "adjoint_over_forward": b_ms / f_ms},
           "roofline": {"bound": "hbm", "achieved": adj_bytes / (b_ms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS,
                        "unit": "GB/s", "frac": adj_bytes / (b_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                        "traffic": measured_traffic(PROFILE_CAUSTIC_TRAFFIC, "causticBackward"),
                        "kernel": "causticBackward", "algorithmic_bytes_per_launch": adj_bytes}}
    c.close()
    if not args.no_cpu_baseline:
        sys.path.insert(0, str(ROOT / "oracle"))
        import numpy as np
        from pyoracle import CausticOracle, Oracle

        threads = min(16, os.cpu_count() or 1)
        co = CausticOracle(Oracle(), s)
        Hn = H.cpu().numpy().astype(np.float64)
        frames, el = 0, 0.0
        while el < 10.0 and frames < 16:
            t0 = time.perf_counter()
            co.forward(Hn, frames, threads=threads)
            el += time.perf_counter() - t0
            frames += 1
        # The CPU has no adjoint: a gradient by central differences costs 2 x 65,536 forward passes.
        out["cpu_baseline"] = {"value": frames / el / (2 * s.nodes * s.nodes), "unit": "gradients/s (central differences)",
                               "cores": threads, "kind": "port",
                               "forward_particles_per_s": frames * n / el,
                               "sample": f"{frames} forward passes of {n} particles ({el:.1f} s) through the oracle's "
                                         f"fp64 slab, {threads} threads; a finite-difference gradient costs "
                                         f"2 x {s.nodes * s.nodes} such passes"}
    return out


def telescope_variants(n: int = 64, rings: int = 200, seed: int = 5):
    from owe_hip import lens as L

    return L.gregorian_variants(n, rings, seed)


TELESCOPE_PARAMS = [(0, 0), (0, 1), (1, 0), (1, 1), (0, 2), (1, 2), (2, 0), (4, 0)]


def bench_telescope(args):
    """Differentiable telescope: 64 perturbed Gregorians, each traced with 120,601 pupil rays at the F,
    d and C lines; per variant the angular RMS of its afocal spot and its derivatives with respect
    to 8 parameters (both mirrors' curvatures and conics, both mirror gaps, two ocular curvatures)
    by forward-mode dual numbers."""
    import numpy as np

    from owe_hip import lens as L

    variants = telescope_variants()
    lams = (486.13, 587.56, 656.27)
    st = L.c_settings(field_deg=0.05, wavelengths=lams, pixels=0, reference=587.5618)
    b = L.LensBatch(variants, st)
    b.set_params(TELESCOPE_PARAMS)
    for _ in range(args.warmup):
        b.run()
    t0 = time.perf_counter()
    kernel_ms = 0.0
    for _ in range(args.steps):
        b.run()
        kernel_ms += b.last_ms()
    wall = time.perf_counter() - t0
    g = b.gradients()
    per, _, _ = b.results()
    traces = len(variants) * variants[0].rays * len(lams)
    grads = len(variants) * len(TELESCOPE_PARAMS)
    log(f"telescope: {args.steps} runs in {wall:.3f}s; {grads * args.steps / wall:.0f} derivatives/s; kernel "
        f"{kernel_ms / args.steps:.2f} ms/run; angular RMS {np.mean([r['angular_rms'] for r in per]):.3e} rad")
    out = {"value": grads * args.steps / wall, "unit": "spot-metric derivatives/s (variant x parameter)",
           "ms_per_step": wall * 1e3 / args.steps,
           "config": {"workload": "differentiable Gregorian telescope: 64 perturbed variants x 8 parameters, "
                                  "120,601 pupil rays x 3 wavelengths each, afocal angular RMS",
                      "variants": len(variants), "parameters": len(TELESCOPE_PARAMS), "traces_per_step": traces,
                      "kernel_ms_per_step": kernel_ms / args.steps,
                      "ray_traces_per_s": traces * args.steps / wall,
                      "mean_angular_rms_rad": float(np.mean([r["angular_rms"] for r in per])),
                      "gradient_norm_mean": float(np.linalg.norm(g, axis=1).mean())}}
    b.close()
    if not args.no_cpu_baseline:
        sys.path.insert(0, str(ROOT / "oracle"))
        from pyoracle import LensOracle, Oracle

        lo = LensOracle(Oracle(), variants[:1], field_deg=0.05, reference=587.5618)
        idx = np.arange(variants[0].rays, dtype=np.uint64)
        t0 = time.perf_counter()
        # One variant's gradient by central differences: two perturbed traces per parameter.
        for k in range(len(TELESCOPE_PARAMS)):
            for sgn in (1.0, -1.0):
                delta = np.zeros(len(TELESCOPE_PARAMS))
                delta[k] = sgn * 1e-6
                for lam in lams:
                    lo.trace_perturbed(0, lam, TELESCOPE_PARAMS, delta, idx)
        el = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": len(TELESCOPE_PARAMS) / el, "unit": "spot-metric derivatives/s", "cores": 1,
                               "kind": "port",
                               "sample": f"one variant's 8 derivatives by central differences of the oracle's fp64 "
                                         f"trace (16 perturbed traces of {variants[0].rays} rays x 3 wavelengths, "
                                         f"{el:.1f} s, 1 thread)"}
    return out


TELESCOPE_VARIANTS, TELESCOPE_SPP, TELESCOPE_PASS_SPP = 64, 256, 16


def bench_telescope_render(args):
    """The differentiable telescope's rendering half: the 64 perturbed Gregorians as matter, each seen
    through its ocular in a closed room with a 2,009,780-triangle statue, 512×512, depth 32, as ONE
    batch (owe_hip_batch): every pass renders all 64 variants in one persistent launch. A step is one
    pass of 16 samples per pixel of every variant (the full 256-spp workload is 16 such passes, the
    same launch repeated); the launch shape is autotuned on the first variant. Beside it, the first
    four variants rendered one renderer at a time with the same shape, the loop the batch replaces."""
    from owe_hip import bodies as B
    from owe_hip import render
    from owe_hip import scene as S

    variants = telescope_variants(n=TELESCOPE_VARIANTS)
    st = S.settings(seed=1, max_depth=32, rr_depth=4)
    size = 512
    t0 = time.perf_counter()
    scenes = [B.telescope_room_scene(v, size=size) for v in variants]
    t1 = time.perf_counter()
    batch = render.Batch(scenes, st)
    t2 = time.perf_counter()
    lead = batch.members[0]
    lead.set_mode(render.Renderer.PERSISTENT, use_graph=False)
    lead.autotune(TELESCOPE_PASS_SPP)
    shape = lead.mode()
    batch.run_pass(TELESCOPE_PASS_SPP)  # warm-up
    batch.set_timing(True)
    before = batch.stats()
    steps = max(1, min(args.steps, 2))
    for _ in range(steps):
        batch.run_pass(TELESCOPE_PASS_SPP)
    ms, launches = batch.kernel_time()
    after = batch.stats()
    paths = after["paths"] - before["paths"]
    segments = after["segments"] - before["segments"]
    # The loop the batch replaces: the first variants one renderer at a time, same shape and pass.
    loop_ms, loop_paths = 0.0, 0
    for m in batch.members[:4]:
        m.set_mode(render.Renderer.PERSISTENT, shape["blocks_per_cu"], shape["chunk"], False, shape["waves_per_simd"],
                   shape["refill"])
        m.set_timing(True)
        m.run_pass(TELESCOPE_PASS_SPP)
        loop_ms += m.kernel_times()["persistent"][0]
        loop_paths += size * size * TELESCOPE_PASS_SPP
    batch_rate = paths / (ms * 1e-3)
    loop_rate = loop_paths / (loop_ms * 1e-3)
    log(f"telescope render: {TELESCOPE_VARIANTS} variants x {size}x{size}x{TELESCOPE_PASS_SPP} spp per launch: "
        f"{batch_rate / 1e6:.1f} M paths/s batched ({ms / launches:.1f} ms/launch), {loop_rate / 1e6:.1f} M paths/s "
        f"one variant at a time; E_seg {segments / paths:.2f}; scenes built in {t1 - t0:.1f} s, batch created in "
        f"{t2 - t1:.1f} s; shape {shape}; {after}")
    out = {"value": batch_rate, "unit": "spectral paths/s",
           "ms_per_step": ms / launches,
           "config": {"workload": "64 perturbed Gregorian telescopes as matter in a 2,009,780-triangle room, "
                                  "512x512, depth 32, all 64 in one persistent launch per pass (owe_hip_batch)",
                      "variants": TELESCOPE_VARIANTS, "spp_per_launch": TELESCOPE_PASS_SPP,
                      "spp_full_workload": TELESCOPE_SPP, "paths_per_launch": paths // launches,
                      "full_workload_ms": ms / launches * TELESCOPE_SPP / TELESCOPE_PASS_SPP,
                      "segments_per_path": segments / paths, "launch": shape,
                      "one_at_a_time_paths_per_s": loop_rate, "batched_over_one_at_a_time": batch_rate / loop_rate,
                      "inconsistencies": after["inconsistencies"], "leaks": after["leaks"],
                      "host_build_s": {"scenes": t1 - t0, "batch": t2 - t1}}}
    batch.close()
    if not args.no_cpu_baseline:
        sys.path.insert(0, str(ROOT / "oracle"))
        from pyoracle import Oracle, World

        threads = min(16, os.cpu_count() or 1)
        sc = B.telescope_room_scene(variants[0], size=128)
        w = World(Oracle(), sc)
        el, n, passes = 0.0, 0, 0
        while el < 8.0 and passes < 64:
            t0 = time.perf_counter()
            _, ost = w.render(st, 16, threads=threads, first_pass=passes)
            el += time.perf_counter() - t0
            n += int(ost[0])
            passes += 1
        out["cpu_baseline"] = {"value": n / el, "unit": "spectral paths/s", "cores": threads, "kind": "port",
                               "sample": f"{passes} passes of 128x128x16 spp of the first variant's scene (the same "
                                         f"room and statue) through the oracle's fp64 Tracer::radiance, {threads} "
                                         f"threads, {el:.1f} s"}
    return out


def bench_scenes(args):
    """The scene-file suite (owe bench): the reference's five benchmark scenes and the rippled pool read
    from .owe files, one pass of 16 camera paths per pixel at 256x192 each in the persistent form,
    best of 3 fresh renderers; the aggregate paths/s and each scene's digest (FNV-1a of its raw XYZ
    image)."""
    from owe_hip import render, suite

    res = suite.run(spp=16, resolution=(256, 192), repeat=3, mode=render.Renderer.PERSISTENT)
    for r in res["rows"]:
        print(f"scenes: {Path(r['scene']).name:<20s} {r['mpaths_per_s']:9.1f} M paths/s  {r['msegs_per_s']:9.1f} "
              f"M segments/s  mean Y {r['mean_Y']:.6g}  digest {r['digest']}", file=sys.stderr, flush=True)
    return {"value": res["mpaths_per_s"] * 1e6, "unit": "paths/s (suite aggregate)",
            "config": {"workload": "scene files through the loader: the reference's benchmarks dispersion_prism, "
                                   "caustic_seen_through_glass, caustic_seen_in_mirror, double_refraction_slabs, "
                                   "telescope_exit_pupil_walk, and a rippled pool",
                       "resolution": [256, 192], "spp": 16, "form": "persistent", "timing": "best of 3 passes"},
            "rows": [{k: r[k] for k in ("scene", "detector", "mpaths_per_s", "msegs_per_s", "mean_Y", "digest",
                                        "inconsistencies", "leaks")} | {"scene": Path(r["scene"]).name}
                     for r in res["rows"]]}


SPPM_SIZE, SPPM_ITERATIONS = 256, 16


def cpu_baseline_sppm(ls, sc, st, seconds_budget: float = 10.0):
    sys.path.insert(0, str(ROOT / "oracle"))
    from pyoracle import Oracle, World

    threads = min(16, os.cpu_count() or 1)
    w = World(Oracle(), sc)
    state, elapsed = None, 0.0
    while elapsed < seconds_budget and (state is None or state["iterations"] < 256):
        t0 = time.perf_counter()
        state = w.sppm(st, 2, state=state, threads=threads)
        elapsed += time.perf_counter() - t0
    n = int(state["stats"][0])
    return {"value": n / elapsed, "unit": "camera paths + photons/s", "cores": threads, "kind": "port",
            "sample": f"{state['iterations']} SPPM iterations of the same scene ({n} camera paths and photons, "
                      f"{elapsed:.1f} s) through the oracle's fp64 restatement, {threads} threads"}


def bench_sppm(args):
    """SPPM (stochastic progressive photon mapping) of a lens caustic seen through a glass plate: per
    iteration one camera path per pixel to its visible point, the hash fill, one photon per pixel from
    the sun, the progressive update; a step is a pass of 16 iterations at 256x256."""
    import torch

    from owe_hip import caustic_scenes as C
    from owe_hip import render

    ls = C.caustic_cases(resolution=SPPM_SIZE)["plate"]
    sc = ls.use()
    st = ls.settings(seed=1)
    r = render.Renderer(sc, st)
    for _ in range(args.warmup):
        r.run_pass(SPPM_ITERATIONS)
    r.reset()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        r.run_pass(SPPM_ITERATIONS)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    stats = r.stats()
    n = stats["paths"]
    its = args.steps * SPPM_ITERATIONS
    log(f"sppm: {its} iterations in {wall:.3f}s; {n / wall / 1e6:.1f} M camera paths + photons/s; "
        f"{wall * 1e3 / its:.3f} ms/iteration")
    out = {"value": n / wall, "unit": "camera paths + photons/s", "ms_per_step": wall * 1e3 / args.steps,
           "config": {"workload": "SPPM, lens caustic seen through a glass plate", "resolution": [SPPM_SIZE, SPPM_SIZE],
                      "iterations_per_step": SPPM_ITERATIONS, "photons_per_iteration": SPPM_SIZE * SPPM_SIZE,
                      "ms_per_iteration": wall * 1e3 / its, "segments": stats["segments"],
                      "mean_Y": float(r.resolve()[..., 1].mean())},
           "roofline": None}
    out["cpu_baseline"] = None if args.no_cpu_baseline else cpu_baseline_sppm(ls, sc, st)
    r.close()
    return out


VCM_SIZE, VCM_ITERATIONS = 256, 16


def cpu_baseline_vcm(ls, sc, st, seconds_budget: float = 10.0):
    sys.path.insert(0, str(ROOT / "oracle"))
    from pyoracle import Oracle, World

    threads = min(16, os.cpu_count() or 1)
    w = World(Oracle(), sc)
    n, its, elapsed = 0, 0, 0.0
    while elapsed < seconds_budget and its < 256:
        t0 = time.perf_counter()
        r = w.vcm(st, 2, merging=True, start=its, threads=threads)
        elapsed += time.perf_counter() - t0
        n += int(r["stats"][0])
        its += 2
    return {"value": n / elapsed, "unit": "camera paths + light subpaths/s", "cores": threads, "kind": "port",
            "sample": f"{its} vcm iterations of the same scene ({n} camera paths and light subpaths, {elapsed:.1f} s) "
                      f"through the oracle's fp64 restatement (merging by a sweep over the group's vertices sorted by x), "
                      f"{threads} threads"}


def bench_vcm(args):
    """Vertex connection and merging of a lens caustic seen through a glass plate: per iteration one
    light subpath per pixel (cached, splatted to the eye), the cache's scan and compaction, the merging
    hash, one camera path per pixel with next-event estimation, a connection and merging; a step is a
    pass of 16 iterations at 256x256."""
    import torch

    from owe_hip import caustic_scenes as C
    from owe_hip import render

    ls = C.caustic_cases(resolution=VCM_SIZE)["plate"]
    sc = ls.use()
    st = ls.settings(seed=1, integrator="vcm")
    r = render.Renderer(sc, st)
    for _ in range(args.warmup):
        r.run_pass(VCM_ITERATIONS)
    r.reset()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        r.run_pass(VCM_ITERATIONS)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    stats = r.stats()
    n = stats["paths"]
    its = args.steps * VCM_ITERATIONS
    img = r.resolve()
    # The kernels' own times, with HIP events on the renderer's stream, over as many passes again:
    # the camera kernel (counted as shade) and the light-subpath kernel (counted as generate).
    r.set_timing(True)
    for _ in range(args.steps):
        r.run_pass(VCM_ITERATIONS)
    times = r.kernel_times()
    r.set_timing(False)
    cam_ms, cam_n = times["shade"]
    light_ms, light_n = times["generate"]
    cam_launch = cam_ms / max(1, cam_n)
    log(f"vcm: {its} iterations in {wall:.3f}s; {n / wall / 1e6:.1f} M camera paths + light subpaths/s; "
        f"{wall * 1e3 / its:.3f} ms/iteration; camera kernel {cam_launch:.3f} ms, light kernel "
        f"{light_ms / max(1, light_n):.3f} ms per launch of {VCM_ITERATIONS} iterations")
    roofline = valu_roofline("vcmCameraKernel", cam_launch)
    if roofline:
        roofline["traffic"] = measured_traffic(PROFILE_VCM, "vcmCameraKernel")
    out = {"value": n / wall, "unit": "camera paths + light subpaths/s", "ms_per_step": wall * 1e3 / args.steps,
           "config": {"workload": "vcm, lens caustic seen through a glass plate", "resolution": [VCM_SIZE, VCM_SIZE],
                      "iterations_per_step": VCM_ITERATIONS, "light_subpaths_per_iteration": VCM_SIZE * VCM_SIZE,
                      "iterations_in_flight": r.vcm_in_flight()[0],
                      "ms_per_iteration": wall * 1e3 / its, "segments": stats["segments"],
                      "camera_kernel_ms_per_launch": cam_launch,
                      "light_kernel_ms_per_launch": light_ms / max(1, light_n),
                      "mean_Y": float(img[..., 1].mean())},
           "roofline": roofline}
    out["cpu_baseline"] = None if args.no_cpu_baseline else cpu_baseline_vcm(ls, sc, st)
    r.close()
    return out


def main() -> int:
    import torch  # noqa: F401  (first: its HIP runtime is the process's; see owe_hip/_lib.py)

    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="prism", choices=["prism", "rayquery", "light", "lens", "caustic",
                                                             "telescope", "scenes", "sppm", "vcm"])
    ap.add_argument("--no-rayquery", action="store_true", help="prism: skip the ray-query record beside it")
    ap.add_argument("--no-light", action="store_true", help="prism: skip the light-tracing record beside it")
    ap.add_argument("--no-lens", action="store_true", help="prism: skip the lens-analysis record beside it")
    ap.add_argument("--no-caustic", action="store_true", help="prism: skip the caustic-gradient record beside it")
    ap.add_argument("--no-telescope", action="store_true", help="prism: skip the telescope-gradient record beside it")
    ap.add_argument("--no-scenes", action="store_true", help="prism: skip the scene-file suite record")
    ap.add_argument("--no-sppm", action="store_true", help="prism: skip the SPPM record")
    ap.add_argument("--no-vcm", action="store_true", help="prism: skip the vertex connection and merging record")
    ap.add_argument("--no-telescope-render", action="store_true",
                    help="telescope: skip the rendering of the telescopes in the room")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only copy a 1 GiB device buffer 3 times (the known-size kernel that calibrates "
                         "rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    ap.add_argument("--no-cpu-baseline", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-quality", action="store_true", help="prism: skip the image-quality ladder")
    ap.add_argument("--no-lifted", action="store_true", help="prism: skip the lifted-prism variant")
    ap.add_argument("--solo", action="store_true",
                    help="the chosen workload alone: no records beside it, no quality ladder, no lifted variant")
    args = ap.parse_args()
    if args.solo:
        for k in ("no_rayquery", "no_light", "no_lens", "no_caustic", "no_telescope", "no_scenes", "no_sppm",
                  "no_vcm", "no_quality", "no_lifted", "no_telescope_render"):
            setattr(args, k, True)
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print(f"bench.py measures one GPU only (asked for {args.gpus}); every workload here runs on a single "
              "MI355X and no multi-GPU figure is produced.", file=sys.stderr)
        return 2
    if args.calibrate_copy:
        import torch

        src = torch.ones(256 * 1024 * 1024, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"calibration": "copy", "bytes_read_per_copy": src.numel() * 4,
                          "bytes_written_per_copy": dst.numel() * 4}), flush=True)
        return 0
    if args.workload == "prism":
        result = bench_prism(args)
        if not args.no_rayquery:
            rq = bench_rayquery(argparse.Namespace(steps=args.steps, warmup=args.warmup,
                                                   no_cpu_baseline=args.no_cpu_baseline))
            result["rayquery"] = {k: rq[k] for k in ("value", "unit", "ms_per_step", "config", "roofline", "cpu_baseline")}
        if not args.no_light:
            result["light"] = bench_light(argparse.Namespace(steps=args.steps, warmup=args.warmup,
                                                             no_cpu_baseline=args.no_cpu_baseline))
        if not args.no_lens:
            result["lens"] = bench_lens(argparse.Namespace(steps=args.steps, warmup=args.warmup,
                                                           no_cpu_baseline=args.no_cpu_baseline))
        if not args.no_caustic:
            result["caustic"] = bench_caustic(argparse.Namespace(steps=args.steps, warmup=args.warmup,
                                                                 no_cpu_baseline=args.no_cpu_baseline))
        if not args.no_telescope:
            ns = argparse.Namespace(steps=args.steps, warmup=args.warmup, no_cpu_baseline=args.no_cpu_baseline)
            result["telescope"] = bench_telescope(ns)
            if not args.no_telescope_render:
                result["telescope"]["render"] = bench_telescope_render(ns)
        if not args.no_scenes:
            result["scenes"] = bench_scenes(args)
        if not args.no_sppm:
            result["sppm"] = bench_sppm(argparse.Namespace(steps=min(args.steps, 8), warmup=1,
                                                           no_cpu_baseline=args.no_cpu_baseline))
        if not args.no_vcm:
            result["vcm"] = bench_vcm(argparse.Namespace(steps=min(args.steps, 8), warmup=1,
                                                         no_cpu_baseline=args.no_cpu_baseline))
    elif args.workload == "sppm":
        result = bench_sppm(args)
    elif args.workload == "vcm":
        result = bench_vcm(args)
    elif args.workload == "scenes":
        result = bench_scenes(args)
    elif args.workload == "caustic":
        result = bench_caustic(args)
    elif args.workload == "telescope":
        result = bench_telescope(args)
        if not args.no_telescope_render:
            result["render"] = bench_telescope_render(args)
    elif args.workload == "light":
        result = bench_light(args)
    elif args.workload == "lens":
        result = bench_lens(args)
    else:
        result = bench_rayquery(args)
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
