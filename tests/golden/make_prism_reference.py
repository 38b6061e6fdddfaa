"""Writes prism_reference.npz: the fp64 oracle's image of the dispersion-prism workload (the
equilateral Sellmeier prism standing on the 1 m² Lambertian floor in sunlight, 256×256 pinhole,
one wavelength per path uniform over 380–780 nm, max depth 16, roulette from depth 4) at 16,384
samples per pixel — 256 times the 64 spp a timed pass takes — from seed 1000, a stream set no
benchmark or test draws. The quality half of the prism benchmark measures its images against it
(relMSE as docs/BENCHMARKS.md:94-101 defines it).

    python tests/golden/make_prism_reference.py   # about 2 minutes on 8 cores
"""
from __future__ import annotations

import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path[:0] = [str(ROOT / "optical-pathtracer_amd"), str(ROOT / "oracle")]

from owe_hip import scene as S  # noqa: E402
from pyoracle import Oracle, World  # noqa: E402

SPP, PASSES, SEED = 1024, 16, 1000
OUT = Path(__file__).resolve().parent / "prism_reference.npz"


def main() -> None:
    sc = S.dispersion_prism_scene(256, 256, lift=0.0)
    st = S.settings(seed=SEED, max_depth=16, rr_depth=4, spectral=S.SPECTRAL_UNIFORM1, lambda_min=380.0,
                    lambda_max=780.0)
    t0 = time.time()
    img, stats = World(Oracle(), sc).render(st, SPP, passes=PASSES, layout=1, threads=os.cpu_count() or 1)
    print(f"{SPP * PASSES} spp in {time.time() - t0:.1f} s, stats {stats}, mean Y {img[..., 1].mean():.6f}")
    np.savez_compressed(OUT, xyz=img.astype(np.float32), spp=SPP * PASSES, seed=SEED,
                        stats=np.asarray(stats, dtype=np.uint64))


if __name__ == "__main__":
    main()
