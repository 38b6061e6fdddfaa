"""A brute-force random walk in a homogeneous sphere, written directly from the radiative transfer
equation and independent of the oracle's transport code (no next-event estimation, no MIS, no
wavelength sharing): the check that pins the oracle's medium transport where no closed form does.

The eye sits at the centre of a unit sphere filled with a medium of extinction σt = σs + σa and
Henyey–Greenstein asymmetry g; the sphere's inner wall emits unit radiance and reflects diffusely
with albedo ρ. Following a ray backwards from the eye, every collision keeps the fraction
ϖ = σs/σt of the weight and turns the direction by an angle drawn from the HG phase; at the wall
the walk collects the emitted 1 and continues, weighted by ρ, in a cosine-distributed direction.
By symmetry every viewing direction sees the same radiance, E[Σ weights collected].
"""
from __future__ import annotations

import numpy as np


def _frame(d):
    """Two unit vectors orthogonal to each row of d (and to each other)."""
    a = np.where(np.abs(d[:, :1]) > 0.9, np.array([[0.0, 1.0, 0.0]]), np.array([[1.0, 0.0, 0.0]]))
    u = np.cross(d, a)
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    return u, np.cross(d, u)


def _turn(d, cos_t, phi):
    s = np.sqrt(np.maximum(0.0, 1.0 - cos_t * cos_t))
    u, v = _frame(d)
    return (u * (s * np.cos(phi))[:, None] + v * (s * np.sin(phi))[:, None] + d * cos_t[:, None])


def hg_cosine(g: float, xi: np.ndarray) -> np.ndarray:
    """Cosine between the old and the new propagation direction, drawn from the HG phase by
    inverting its cumulative distribution (Henyey & Greenstein 1941; the standard inversion)."""
    if abs(g) < 1e-6:
        return 1.0 - 2.0 * xi
    f = (1.0 - g * g) / (1.0 - g + 2.0 * g * xi)
    return (1.0 + g * g - f * f) / (2.0 * g)


def sphere_radiance(sigma_s: float, sigma_a: float, g: float, rho: float = 0.0, walks: int = 400_000,
                    seed: int = 0, max_events: int = 100_000) -> tuple[float, float]:
    """(mean, standard error) of the radiance seen from the centre."""
    rng = np.random.default_rng(seed)
    sigma_t = sigma_s + sigma_a
    albedo = sigma_s / sigma_t if sigma_t > 0 else 1.0
    d = rng.normal(size=(walks, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    p = np.zeros((walks, 3))
    w = np.ones(walks)
    total = np.zeros(walks)
    alive = np.arange(walks)
    for _ in range(max_events):
        if alive.size == 0:
            break
        pa, da = p[alive], d[alive]
        b = np.einsum("ij,ij->i", pa, da)
        c = np.einsum("ij,ij->i", pa, pa) - 1.0
        wall = -b + np.sqrt(np.maximum(0.0, b * b - c))
        free = -np.log1p(-rng.random(alive.size)) / sigma_t if sigma_t > 0 else np.full(alive.size, np.inf)
        hit = free >= wall
        # Collisions inside the medium.
        col = alive[~hit]
        p[col] = pa[~hit] + da[~hit] * free[~hit, None]
        w[col] *= albedo
        d[col] = _turn(da[~hit], hg_cosine(g, rng.random(col.size)), 2 * np.pi * rng.random(col.size))
        # The wall: collect the emission, then reflect diffusely about the inward normal.
        at = alive[hit]
        q = pa[hit] + da[hit] * wall[hit, None]
        total[at] += w[at]
        w[at] *= rho
        n = -q / np.linalg.norm(q, axis=1, keepdims=True)
        xi1, xi2 = rng.random(at.size), rng.random(at.size)
        p[at] = q * (1 - 1e-12)
        d[at] = _turn(n, np.sqrt(1 - xi1), 2 * np.pi * xi2)
        # Russian roulette on small weights keeps the expectation.
        low = w < 0.05
        keep = rng.random(walks) < w / 0.05
        w = np.where(low & keep, 0.05, np.where(low, 0.0, w))
        alive = np.nonzero(w > 0)[0]
    return float(total.mean()), float(total.std(ddof=1) / np.sqrt(walks))
