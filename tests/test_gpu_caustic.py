"""The caustic slab's forward pass and path-replay adjoint on the GPU.

* forward against the fp64 oracle on the same particles: the same irradiance up to float rounding
  of each particle's wall point; identical passes are bit-identical (integer bins);
* adjoint against central finite differences of L = mean((E − target)²) on 64 random heights, on
  the GPU itself and in fp64 on the CPU, with a step in the linear regime of the loss (the particles
  move a few thousandths of a pixel); the 1e-4 mm step is also reported — there each particle moves
  about a third of a pixel, so the difference quotient is a secant a few percent from the tangent,
  the same for both precisions;
* the torch.autograd.Function gives the same gradient, and gradient steps lower the loss;
* the adjoint costs at most three forward passes.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

from owe_hip import caustic as C
from pyoracle import CausticOracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def setup(cuda):
    import torch

    s = C.caustic_settings(particles=1 << 18)
    c = C.Caustic(s)
    H = torch.tensor(C.initial_heights(), device="cuda")
    T = torch.tensor(C.target(s), device="cuda")
    yield s, c, H, T
    c.close()


def forward(c, s, H, pass_index=0):
    import torch

    out = torch.empty((s.pixels, s.pixels), dtype=torch.float32, device="cuda")
    c.forward_into(H.contiguous(), out, pass_index)
    torch.cuda.synchronize()
    return out


def loss_of(E, T):
    return float(((E.double() - T.double()) ** 2).mean())


def adjoint(c, s, H, T, pass_index=0):
    import torch

    E = forward(c, s, H, pass_index)
    gE = (2.0 * (E.double() - T.double()) / E.numel()).float().contiguous()
    g = torch.empty_like(H)
    c.backward_into(H.contiguous(), gE, g, pass_index)
    torch.cuda.synchronize()
    return g


def test_forward_matches_oracle(setup, oracle):
    s, c, H, T = setup
    E = forward(c, s, H).cpu().numpy().astype(np.float64)
    Eo = CausticOracle(oracle, s).forward(H.cpu().numpy().astype(np.float64))
    pix = 2 * s.wall_half / s.pixels
    tot = abs(E.sum() - Eo.sum()) / Eo.sum()
    diff = np.abs(E - Eo)
    print(f"forward: flux {E.sum() * pix * pix:.6e} vs {Eo.sum() * pix * pix:.6e} (rel {tot:.2e}), "
          f"max |ΔE| {diff.max():.3e} of max E {Eo.max():.3f}", flush=True)
    assert tot < 1e-5
    assert diff.max() < 1e-3 * Eo.max()
    assert np.array_equal(E, forward(c, s, H).cpu().numpy().astype(np.float64))


def test_adjoint_matches_finite_differences(setup, oracle):
    import torch

    s, c, H, T = setup
    g = adjoint(c, s, H, T).cpu().numpy().ravel().astype(np.float64)
    nodes = np.random.default_rng(17).integers(0, s.nodes * s.nodes, 64)
    co = CausticOracle(oracle, s)
    Hn = H.cpu().numpy().astype(np.float64)
    Tn = T.cpu().numpy().astype(np.float64)
    fd_cpu = co.fd(Hn, Tn, nodes, 1e-9)
    secant_cpu = co.fd(Hn, Tn, nodes, 1e-7)

    def fd_gpu(h):
        out = []
        flat = H.clone().view(-1)
        for k in nodes:
            v = float(flat[k])
            flat[k] = v + h
            lp = loss_of(forward(c, s, flat.view(H.shape)), T)
            flat[k] = v - h
            lm = loss_of(forward(c, s, flat.view(H.shape)), T)
            flat[k] = v
            out.append((lp - lm) / (2 * h))
        return np.array(out)

    fd = fd_gpu(3e-9)
    secant = fd_gpu(1e-7)
    a = g[nodes]
    results = {}
    for name, ref in (("gpu fd", fd), ("fp64 cpu fd", fd_cpu)):
        rel = np.linalg.norm(a - ref) / np.linalg.norm(ref)
        worst = (np.abs(a - ref) / (np.abs(ref) + 0.02 * np.abs(ref).max())).max()
        results[name] = (rel, worst)
        print(f"adjoint vs {name}: relative error {rel:.2e} (norm over 64 heights), worst component {worst:.2e}",
              flush=True)
    sec = np.linalg.norm(a - secant) / np.linalg.norm(secant)
    agree = np.linalg.norm(secant - secant_cpu) / np.linalg.norm(secant_cpu)
    print(f"1e-4 mm step: adjoint vs secant {sec:.2e}; gpu secant vs fp64 secant {agree:.2e}", flush=True)
    # Within 1 % of the GPU's own finite differences over the 64 heights, each component within 5 %
    # (the float forward's rounding makes single small differences noisy); the fp64 differences are
    # free of that noise and pin the adjoint ten times tighter.
    rel, worst = results["gpu fd"]
    assert rel < 0.01 and worst < 0.05
    rel, worst = results["fp64 cpu fd"]
    assert rel < 1e-3 and worst < 5e-3
    assert agree < 0.01


def test_autograd_function_and_descent(setup):
    import torch

    s, c, H, T = setup
    g_direct = adjoint(c, s, H, T)
    Hp = H.clone().requires_grad_(True)
    E = c.irradiance(Hp)
    L = ((E.double() - T.double()) ** 2).mean()
    L.backward()
    torch.cuda.synchronize()
    assert torch.allclose(Hp.grad, g_direct, rtol=1e-5, atol=1e-6 * float(g_direct.abs().max()))
    # A few normalised gradient steps on fresh particle passes lower the loss.
    Hs = H.clone()
    losses = []
    for it in range(6):
        Hv = Hs.clone().requires_grad_(True)
        L = ((c.irradiance(Hv, it).double() - T.double()) ** 2).mean()
        L.backward()
        losses.append(float(L.detach()))
        with torch.no_grad():
            Hs = Hs - 2e-6 * Hv.grad / Hv.grad.abs().max()
    print(f"losses: {losses}", flush=True)
    assert losses[-1] < losses[0]


def test_adjoint_costs_at_most_three_forwards(cuda):
    import torch

    s = C.caustic_settings()
    c = C.Caustic(s)
    H = torch.tensor(C.initial_heights(), device="cuda")
    T = torch.tensor(C.target(s), device="cuda")
    E = torch.empty((s.pixels, s.pixels), dtype=torch.float32, device="cuda")
    gE = torch.zeros_like(E)
    g = torch.empty_like(H)
    for _ in range(2):
        c.forward_into(H, E, 0)
        c.backward_into(H, gE + 1e-3, g, 0)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    reps = 5
    ev[0].record()
    for k in range(reps):
        c.forward_into(H, E, k)
    ev[1].record()
    gE = (2.0 * (E - T) / E.numel()).contiguous()
    for k in range(reps):
        c.backward_into(H, gE, g, k)
    ev[2].record()
    torch.cuda.synchronize()
    f, b = ev[0].elapsed_time(ev[1]) / reps, ev[1].elapsed_time(ev[2]) / reps
    print(f"forward {f:.3f} ms, adjoint {b:.3f} ms, ratio {b / f:.2f} (1,048,576 particles)", flush=True)
    assert b <= 3.0 * f
    c.close()
