"""The reference's benchmark scenes on the GPU, against the closed forms they are built around.

* telescope_exit_pupil_walk: an eye of 1.2 mm pupil walked sideways off the afocal refractor's exit
  pupil sees the image dim as the overlap of its pupil with the emerging beam, loses it once the two
  no longer overlap, and sees it again bit for bit when walked back (tests/test_navigation.cpp:220-275:
  the reference measured 0.693, 0.386, 0.139, 0 against the overlap 0.692, 0.395, 0.146, 0);
* double_refraction_slabs: a pinhole eye scanned across the beam beyond the slabs finds it moved by
  the central ray's Snell-law shift (``beam_shift``), and the scan's profile is the fp64 oracle's.

Reference: benchmarks/telescope_exit_pupil_walk.owe, benchmarks/double_refraction_slabs.owe.
"""
from __future__ import annotations

import numpy as np
import pytest

from owe_hip import reference_cases as RC
from owe_hip import render
from pyoracle import World
from test_reference_cases import N_BK7_D, OFFSETS, _half_max_midpoint, check_walk, exit_pupil_walk, slab_profile

pytestmark = pytest.mark.gpu


def _gpu_mean(spp):
    def mean(sc, st):
        r = render.Renderer(sc, st)
        r.run_pass(spp)
        y = r.resolve()[..., 1].mean()
        s = r.stats()
        r.close()
        assert s["leaks"] == 0, s
        return y
    return mean


def test_exit_pupil_walk_follows_the_disc_overlap(cuda):
    print("exit-pupil walk: 32x32, 256 spp per eye position", flush=True)
    check_walk(*exit_pupil_walk(_gpu_mean(256), resolution=32), tight=0.02)


def test_slabs_move_the_beam_by_the_snell_shift(cuda, oracle):
    print(f"slab beam scan: {len(OFFSETS)} eye positions, 16x16, 16 spp", flush=True)
    bare, moved = slab_profile(_gpu_mean(16), False), slab_profile(_gpu_mean(16), True)
    c0, c1 = _half_max_midpoint(OFFSETS, bare), _half_max_midpoint(OFFSETS, moved)
    want = RC.beam_shift_across(N_BK7_D)
    print(f"beam centre without slabs {c0 * 1e3:.3f} mm, with {c1 * 1e3:.3f} mm; central ray {want * 1e3:.3f} mm",
          flush=True)
    assert abs(c0) < 0.05e-3
    assert c1 - c0 == pytest.approx(want, abs=0.12e-3)

    def ora(sc, st):
        img, _ = World(oracle, sc).render(st, 16, layout=1, threads=8)
        return img[..., 1].mean()

    ref = slab_profile(ora, True)
    dev = np.abs(moved - ref).max() / ref.max()
    print(f"scan profile: largest difference from the oracle's {dev:.4f} of the peak", flush=True)
    assert dev < 0.05
