"""Differentiable circuits: potentials and effective resistances with gradients w.r.t. the resistance surface.

The forward pass factorises the grid circuit of conductance 1 / r once on the
GPU and solves every pair's right-hand side; the backward pass reuses that
factor to solve L lambda = dJ/dV and forms dJ/dr per cell from the edge
differences of the potentials and the adjoint (a HIP kernel), for calibrating
resistance surfaces. There is no reference counterpart of this module.
"""

from __future__ import annotations

import ctypes

import numpy as np
import torch

from .. import _native as N
from ._backend import _opts


class _Circuit:
    """A factorised grid circuit held by the library (lsc_grid_circuit)."""

    def __init__(self, r: np.ndarray, four: bool):
        r = N.f64(r)
        self.shape = r.shape
        self.r = r.copy()
        self.version = 0
        self.h = ctypes.c_void_p()
        o = _opts()
        N.check(N.lib().lsc_grid_circuit_create(N.ptr(r), ctypes.c_int64(r.shape[1]), ctypes.c_int64(r.shape[0]),
                                                int(bool(four)), ctypes.byref(o), ctypes.byref(self.h)))

    def update(self, r: np.ndarray) -> None:
        """Factorise again for new resistances on the same cells (the symbolic plans are kept)."""
        r = N.f64(r)
        if r.shape != self.shape:
            raise ValueError("the new resistance must have the circuit's shape")
        N.check(N.lib().lsc_grid_circuit_update(self.h, N.ptr(r)))
        self.r = r.copy()
        self.version += 1

    def ensure(self, r: np.ndarray, version: int) -> None:
        """The factor of r again if the circuit was refactorised since `version` was taken."""
        if self.version != version:
            self.update(r)

    def solve(self, rhs: torch.Tensor) -> torch.Tensor:
        rhs = rhs.to(torch.float64).contiguous()
        out = torch.empty_like(rhs)
        N.check(N.lib().lsc_grid_circuit_solve(self.h, int(rhs.shape[0]), ctypes.c_void_p(rhs.data_ptr()),
                                               ctypes.c_void_p(out.data_ptr())))
        return out

    def adjoint(self, dJdV: torch.Tensor, V: torch.Tensor) -> torch.Tensor:
        dJdV = dJdV.to(torch.float64).contiguous()
        grad = torch.empty(self.shape, dtype=torch.float64, device=V.device)
        N.check(N.lib().lsc_grid_circuit_adjoint(self.h, int(V.shape[0]), ctypes.c_void_p(dJdV.data_ptr()),
                                                 ctypes.c_void_p(V.contiguous().data_ptr()),
                                                 ctypes.c_void_p(grad.data_ptr())))
        return grad

    def __del__(self):
        if getattr(self, "h", None) and self.h.value:
            N.lib().lsc_grid_circuit_destroy(self.h)


_cached: dict = {}


def _circuit_for(r: np.ndarray, four: bool) -> _Circuit:
    """The factorised circuit of r: the last circuit of the same shape, neighbours and cells in the
    circuit (finite r > 0) and solver settings, refactorised on its plans when the values changed, else
    a new one. A
    calibration loop that changes the values of r, not its NoData cells, pays the symbolic analysis
    once. A backward pass whose forward values were replaced since refactorises them again."""
    valid = np.isfinite(r) & (r > 0)
    o = _opts()
    key = (r.shape, bool(four), valid.tobytes(), o.precision, o.tol, o.max_refine)
    c = _cached.get("circuit")
    if c is not None and _cached.get("key") == key:
        if not np.array_equal(c.r, r, equal_nan=True):
            c.update(r)
        return c
    c = _Circuit(r, four)
    _cached.update(circuit=c, key=key)
    return c


class PairPotentials(torch.autograd.Function):
    """V[k] = potentials of the grid circuit of conductance 1 / r with 1 A entering
    at src[k] and leaving at dst[k] (each component grounded at its last node)."""

    @staticmethod
    def forward(ctx, r: torch.Tensor, src: torch.Tensor, dst: torch.Tensor, four: bool = False):
        if r.device.type != "cuda":
            raise ValueError("the resistance must be a CUDA (HIP) tensor")
        ny, nx = r.shape
        circuit = _circuit_for(r.detach().to(torch.float64).cpu().numpy(), four)
        k = src.numel()
        rhs = torch.zeros((k, ny, nx), dtype=torch.float64, device=r.device)
        idx = torch.arange(k, device=r.device)
        rhs.view(k, -1)[idx, dst.to(r.device)] += 1.0
        rhs.view(k, -1)[idx, src.to(r.device)] -= 1.0
        V = circuit.solve(rhs)
        ctx.circuit, ctx.version, ctx.r = circuit, circuit.version, circuit.r
        ctx.save_for_backward(V)
        return V

    @staticmethod
    def backward(ctx, gV: torch.Tensor):
        (V,) = ctx.saved_tensors
        ctx.circuit.ensure(ctx.r, ctx.version)
        grad = ctx.circuit.adjoint(gV, V)
        return grad, None, None, None


def pair_potentials(r: torch.Tensor, src_cells, dst_cells, four_neighbors: bool = False) -> torch.Tensor:
    """Potentials (k, ny, nx) of every pair, differentiable w.r.t. r through the adjoint kernel."""
    src = torch.as_tensor(np.asarray(src_cells, dtype=np.int64))
    dst = torch.as_tensor(np.asarray(dst_cells, dtype=np.int64))
    return PairPotentials.apply(r, src, dst, bool(four_neighbors))


def effective_resistance(r: torch.Tensor, src_cells, dst_cells, four_neighbors: bool = False) -> torch.Tensor:
    """R_eff between every pair of cells (a differentiable function of the resistance surface)."""
    V = pair_potentials(r, src_cells, dst_cells, four_neighbors)
    k = V.shape[0]
    flat = V.reshape(k, -1)
    idx = torch.arange(k, device=V.device)
    s = torch.as_tensor(np.asarray(src_cells, dtype=np.int64), device=V.device)
    t = torch.as_tensor(np.asarray(dst_cells, dtype=np.int64), device=V.device)
    return flat[idx, t] - flat[idx, s]
