"""The scene-file throughput benchmark with result digests, as ``owe bench`` runs it (apps/owe.cpp:588-650).

Each case is a scene file and a detector: the scene is loaded (with edits), the detector set to the
benchmark resolution, and one pass of ``spp`` camera paths per pixel rendered with the path
integrator — the best of ``repeat`` fresh renderers is kept. A row reports paths/s and segments/s
over the pass's own wall time, mean Y, and the digest: FNV-1a 64 over the raw XYZ doubles of the
image, so that a change that moves any result bit shows at once (owe_hip.output.image_digest).

``python -m owe_hip.suite [scene.owe[:Detector] …] [--spp 16] [--resolution 256x192] [--repeat 1]
[--set Block.key=value …] [--json file]`` prints the reference's table and, with --json, its records.
Without scene arguments it runs the reference's benchmark scenes and the rippled pool.
"""
from __future__ import annotations

import argparse
import json
import sys
import time
from pathlib import Path

from . import loader as LD
from . import output as O

SCENES = Path(__file__).parent / "scenes"
# The reference's five benchmark scenes (owe_hip.reference_cases) and the package's rippled pool.
DEFAULT_CASES = [(str(SCENES / f"{n}.owe"), "") for n in
                 ("dispersion_prism", "caustic_seen_through_glass", "caustic_seen_in_mirror", "double_refraction_slabs",
                  "telescope_exit_pupil_walk", "pool")]


def split_spec(spec: str):
    """``file.owe:Detector`` → (file, detector); a bare file names no detector."""
    head, sep, tail = spec.rpartition(":")
    if sep and head.endswith(".owe") and tail:
        return head, tail
    return spec, ""


def run(cases=None, spp: int = 16, resolution=(256, 192), repeat: int = 1, edits=(), mode: int | None = None,
        device: int = 0) -> dict:
    """Renders each case and returns {"rows": [...], "total_seconds", "total_paths", "mpaths_per_s"}."""
    from . import render

    cases = DEFAULT_CASES if not cases else cases
    rows, total_paths, total_seconds = [], 0, 0.0
    for path, det in cases:
        ls = LD.load_scene(path, edits)
        i = ls.detector_index(det or None)
        sc = ls.use(i, resolution)
        st = ls.settings(integrator="path")
        best, img, stats = float("inf"), None, None
        for _ in range(max(1, repeat)):
            r = render.Renderer(sc, st, device)
            if mode is not None:
                r.set_mode(mode, use_graph=mode == render.Renderer.PERSISTENT)
            t0 = time.perf_counter()
            r.run_pass(spp)
            dt = time.perf_counter() - t0
            if dt < best:
                best, img, stats = dt, r.resolve(), r.stats()
            r.close()
        row = dict(scene=path, detector=ls.detectors[i]["name"], resolution=f"{resolution[0]}x{resolution[1]}",
                   spp=spp, integrator="path", seconds=best, paths=int(stats["paths"]), segments=int(stats["segments"]),
                   mean_Y=float(img[..., 1].mean()), digest=O.image_digest(img),
                   inconsistencies=int(stats["inconsistencies"]), leaks=int(stats["leaks"]))
        row["mpaths_per_s"] = row["paths"] / best * 1e-6
        row["msegs_per_s"] = row["segments"] / best * 1e-6
        rows.append(row)
        total_paths += row["paths"]
        total_seconds += best
    return dict(rows=rows, total_seconds=total_seconds, total_paths=total_paths,
                mpaths_per_s=total_paths / total_seconds * 1e-6 if total_seconds > 0 else 0.0)


def format_table(result: dict) -> str:
    head = f"{'scene':<26s} {'detector':<15s} {'res':>9s} {'spp':>4s} {'Mpaths/s':>9s} {'Msegs/s':>9s} " \
           f"{'seconds':>8s} {'mean Y':>12s} {'digest':>16s}"
    lines = [head]
    for r in result["rows"]:
        lines.append(f"{Path(r['scene']).name:<26s} {r['detector']:<15s} {r['resolution']:>9s} {r['spp']:4d} "
                     f"{r['mpaths_per_s']:9.3f} {r['msegs_per_s']:9.3f} {r['seconds']:8.3f} {r['mean_Y']:12.6g} "
                     f"{r['digest']:>16s}")
    lines.append(f"total: {result['total_seconds']:.3f} s, {result['mpaths_per_s']:.3f} Mpaths/s")
    return "\n".join(lines)


def main(argv=None) -> int:
    import torch  # noqa: F401  (the process's HIP runtime is torch's)

    ap = argparse.ArgumentParser(description="scene-file benchmark with digests (owe bench)")
    ap.add_argument("scenes", nargs="*", help="scene.owe[:Detector]")
    ap.add_argument("--spp", type=int, default=16)
    ap.add_argument("--resolution", default="256x192")
    ap.add_argument("--repeat", type=int, default=1)
    ap.add_argument("--set", action="append", default=[], help="Block.key=value")
    ap.add_argument("--json", help="write the records here")
    a = ap.parse_args(argv)
    w, h = (int(x) for x in a.resolution.lower().split("x"))
    result = run([split_spec(s) for s in a.scenes], a.spp, (w, h), a.repeat, tuple(a.set))
    print(format_table(result), flush=True)
    if a.json:
        with open(a.json, "w") as f:
            json.dump(result["rows"], f, indent=2)
            f.write("\n")
    return 0


if __name__ == "__main__":
    sys.exit(main())
