"""The adjoint dJ/dr of J = sum of effective resistances.

CPU: the oracle's FP64 gradient against central finite differences of its own
forward solves and against a dense pseudo-inverse (R_ij = P_ii + P_jj - 2 P_ij).
GPU: the HIP adjoint kernel against the oracle, and the torch autograd path.
"""

from __future__ import annotations

import numpy as np
import pytest


def _surface(shape, seed):
    rng = np.random.default_rng(seed)
    return np.exp(rng.normal(0.0, 1.0, shape))


def _dense_reff(r, pairs, four=False):
    ny, nx = r.shape
    g = 1 / r
    n = nx * ny
    L = np.zeros((n, n))
    steps = [(0, 1), (1, 0)] + ([] if four else [(1, 1), (1, -1)])
    for y in range(ny):
        for x in range(nx):
            for dy, dx in steps:
                yy, xx = y + dy, x + dx
                if 0 <= yy < ny and 0 <= xx < nx:
                    w = (g[y, x] + g[yy, xx]) / (2 * np.sqrt(2) if dy and dx else 2)
                    a, b = y * nx + x, yy * nx + xx
                    L[a, b] -= w
                    L[b, a] -= w
                    L[a, a] += w
                    L[b, b] += w
    P = np.linalg.inv(L + 1.0 / n)
    return np.array([P[s, s] + P[t, t] - 2 * P[s, t] for s, t in pairs])


def test_oracle_resistances_match_pseudo_inverse(oracle):
    r = _surface((7, 9), 1)
    pairs = [(0, 62), (10, 40), (5, 33)]
    J, R, _ = oracle.adjoint(r, [p[0] for p in pairs], [p[1] for p in pairs])
    np.testing.assert_allclose(R, _dense_reff(r, pairs), rtol=1e-10)


@pytest.mark.parametrize("four", [False, True])
def test_oracle_gradient_matches_central_differences(oracle, four):
    r = _surface((12, 14), 2)
    src, dst = [3, 20, 100], [160, 150, 7]
    J, R, grad = oracle.adjoint(r, src, dst, four)
    rng = np.random.default_rng(3)
    for c in rng.choice(r.size, 12, replace=False):
        h = 1e-4 * r.flat[c]
        rp, rm = r.copy(), r.copy()
        rp.flat[c] += h
        rm.flat[c] -= h
        fd = (oracle.adjoint(rp, src, dst, four)[0] - oracle.adjoint(rm, src, dst, four)[0]) / (2 * h)
        np.testing.assert_allclose(grad.flat[c], fd, rtol=1e-6, atol=1e-12 * np.abs(grad).max())
    assert np.all(grad >= 0)  # Rayleigh monotonicity


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_gpu_adjoint_matches_oracle(oracle, precision):
    from landschaft_amd.circuit import _backend

    _backend.settings.precision = precision
    _backend.settings.tol = 1e-14
    r = _surface((64, 80), 4)
    rng = np.random.default_rng(5)
    src = rng.choice(r.size, 8, replace=False)
    dst = rng.choice(r.size, 8, replace=False)
    J, R, grad = _backend.effective_resistance_grad(r, src, dst)
    J0, R0, g0 = oracle.adjoint(r, src, dst)
    np.testing.assert_allclose(R, R0, rtol=1e-9)
    np.testing.assert_allclose(J, J0, rtol=1e-9)
    np.testing.assert_allclose(grad, g0, rtol=1e-8, atol=1e-10 * np.abs(g0).max())


@pytest.mark.gpu
def test_autograd_effective_resistance(oracle):
    import torch

    from landschaft_amd.circuit import _backend
    from landschaft_amd.circuit.adjoint import effective_resistance

    _backend.settings.precision = "mixed"
    _backend.settings.tol = 1e-14
    r0 = _surface((40, 50), 6)
    src, dst = [0, 77, 1500], [1999, 1200, 49]
    r = torch.tensor(r0, device="cuda", requires_grad=True)
    R = effective_resistance(r, src, dst)
    w = torch.tensor([1.0, 2.0, -0.5], dtype=torch.float64, device="cuda")
    (R * w).sum().backward()
    _, R0, _ = oracle.adjoint(r0, src, dst)
    np.testing.assert_allclose(R.detach().cpu().numpy(), R0, rtol=1e-9)
    expect = sum(wk * oracle.adjoint(r0, [s], [t])[2] for wk, s, t in zip([1.0, 2.0, -0.5], src, dst))
    np.testing.assert_allclose(r.grad.cpu().numpy(), expect, rtol=1e-7, atol=1e-10 * np.abs(expect).max())
