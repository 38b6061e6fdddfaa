"""Statistical comparison of two renderers of one measurement (src/owe/backends/compare.cpp:24-135).

Both renderers are unbiased estimators, so their images differ by noise only: each renders `runs`
independent estimates (seeds 1000 + k and 2000 + k), block means of X, Y, Z are formed per run, and
the difference of the means is compared with its standard error estimated from the runs. The
verdict is the reference's ``consistent()``: image |z| < 4, at most max(1, tests/100) block tests
beyond |z| > 4, and Σz²/tests < 2.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Callable

import numpy as np


@dataclass
class ComparisonReport:
    runs: int = 0
    blocks: int = 0
    tests: int = 0
    mean_a: float = 0.0
    se_a: float = 0.0
    mean_b: float = 0.0
    se_b: float = 0.0
    z_image: float = 0.0
    max_abs_z: float = 0.0
    over3: int = 0
    over4: int = 0
    chi2_per_test: float = 0.0
    worst: list = field(default_factory=list)

    def consistent(self) -> bool:
        return abs(self.z_image) < 4 and self.over4 <= max(1, self.tests // 100) and self.chi2_per_test < 2

    def text(self) -> str:
        ratio = self.mean_b / self.mean_a if self.mean_a else 0.0
        return (f"A mean Y {self.mean_a:.6g} ± {self.se_a:.2g}; B mean Y {self.mean_b:.6g} ± {self.se_b:.2g}; "
                f"B/A {ratio:.5f}, z {self.z_image:+.2f}; {self.tests} block tests, Σz²/n {self.chi2_per_test:.3f}, "
                f"max |z| {self.max_abs_z:.2f}, |z|>3 {self.over3}, |z|>4 {self.over4}: "
                f"{'consistent' if self.consistent() else 'INCONSISTENT'}")


def block_means(img: np.ndarray, B: int) -> np.ndarray:
    """Per-block mean XYZ, blocks of B×B pixels (the last row/column of blocks may be partial)."""
    h, w, _ = img.shape
    by, bx = (h + B - 1) // B, (w + B - 1) // B
    out = np.zeros((by, bx, 3))
    cnt = np.zeros((by, bx, 1))
    yy, xx = np.divmod(np.arange(h * w), w)
    np.add.at(out, (yy // B, xx // B), img.reshape(-1, 3))
    np.add.at(cnt, (yy // B, xx // B), 1)
    return (out / np.maximum(cnt, 1)).reshape(-1)


def _mean_se(xs: np.ndarray):
    n = xs.shape[0]
    m = xs.mean(axis=0)
    q = ((xs - m) ** 2).sum(axis=0) / (n - 1)
    return m, np.sqrt(q / n)


def compare_images(runs_a: list[np.ndarray], runs_b: list[np.ndarray], B: int = 8) -> ComparisonReport:
    """compareMeasurements on already rendered runs: images [H, W, 3] of each renderer."""
    if len(runs_a) < 3 or len(runs_b) < 3:
        raise ValueError("compare: need at least 3 runs per renderer")
    rep = ComparisonReport(runs=len(runs_a))
    ya = np.array([r[..., 1].mean() for r in runs_a])
    yb = np.array([r[..., 1].mean() for r in runs_b])
    rep.mean_a, rep.se_a = (float(v) for v in _mean_se(ya))
    rep.mean_b, rep.se_b = (float(v) for v in _mean_se(yb))
    se = math.hypot(rep.se_a, rep.se_b)
    rep.z_image = (rep.mean_b - rep.mean_a) / se if se > 0 else (0.0 if rep.mean_b == rep.mean_a else math.inf)
    ba = np.array([block_means(r, B) for r in runs_a])
    bb = np.array([block_means(r, B) for r in runs_b])
    rep.blocks = ba.shape[1] // 3
    ma, sa = _mean_se(ba)
    mb, sb = _mean_se(bb)
    s = np.hypot(sa, sb)
    zero = ~(s > 0)
    differ = zero & (ma != mb)
    rep.tests = int((~zero).sum() + differ.sum())
    z = np.where(zero, 0.0, (mb - ma) / np.where(zero, 1.0, s))
    rep.over3 = int((np.abs(z[~zero]) > 3).sum() + differ.sum())
    rep.over4 = int((np.abs(z[~zero]) > 4).sum() + differ.sum())
    rep.max_abs_z = float(np.abs(z[~zero]).max()) if (~zero).any() else 0.0
    if differ.any():
        rep.max_abs_z = math.inf
    n_var = int((~zero).sum())
    rep.chi2_per_test = float((z[~zero] ** 2).sum() / rep.tests) if rep.tests else 0.0
    order = np.argsort(-np.abs(z))[:8]
    rep.worst = [(int(i // 3), int(i % 3), float(ma[i]), float(mb[i]), float(z[i])) for i in order]
    del n_var
    return rep


def compare_renderers(render_a: Callable[[int], np.ndarray], render_b: Callable[[int], np.ndarray], runs: int = 4,
                      B: int = 8) -> ComparisonReport:
    """Renders `runs` estimates with each callable (given seeds 1000 + k and 2000 + k) and compares."""
    a = [render_a(1000 + k) for k in range(runs)]
    b = [render_b(2000 + k) for k in range(runs)]
    return compare_images(a, b, B)


def rel_rmse(img: np.ndarray, ref: np.ndarray) -> float:
    """Relative RMSE of Y against a reference image: sqrt(mean((y − r)²) ) / mean(r)."""
    y, r = img[..., 1], ref[..., 1]
    return float(np.sqrt(np.mean((y - r) ** 2)) / np.mean(r))


def rel_mse(img: np.ndarray, ref: np.ndarray) -> float:
    """The reference benchmark metric: mean((img − ref)² / (ref² + 0.01 · mean(ref)²)) over X, Y, Z
    (docs/BENCHMARKS.md:95)."""
    m = ref.mean()
    return float(np.mean((img - ref) ** 2 / (ref ** 2 + 0.01 * m * m)))
