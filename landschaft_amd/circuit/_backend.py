"""The compiled circuit functions, on the HIP library.

``core()`` plays the role of the reference's backend switch
(src/landschaft/circuit/_backend.py:12-23), which returns the pybind11 module
``landschaft._core``; here it returns this module, whose functions keep the
binding's names, arguments and returned dicts (src/cpp/circuit/python.cpp:527-559)
and run on the GPU through ``include/landschaft_circuit.h``.
"""

from __future__ import annotations

import ctypes
import sys
from dataclasses import dataclass

import numpy as np

from .. import _native as N


@dataclass
class SolverSettings:
    """How circuits are solved: ``precision`` "mixed" (FP32 MFMA fronts + FP64
    refinement) or "fp64"; ``tol`` the relative residual refinement reaches;
    ``batch`` circuits per batch."""

    precision: str = "mixed"
    tol: float = 1e-12
    max_refine: int = 30
    batch: int = 4096


settings = SolverSettings()
last_stats: dict = {}


def core():
    """The module behind every API function (the GPU shim)."""
    return sys.modules[__name__]


def _opts(timing=False):
    return N.options(settings.precision, settings.tol, settings.max_refine, settings.batch, timing)


def _record(stats: N.SolveStats):
    last_stats.clear()
    last_stats.update(stats.as_dict())


def circuit_omniscape(resistance, source, radius, buffer, block_size, four_neighbors, resistance_is_conductance,
                      correct_artifacts, flow_potential, n_threads=0, conditions=None, memory_limit=0.0):
    """Omniscape's cumulative current (python.cpp:90-137). ``n_threads`` and
    ``memory_limit`` size the reference's CPU threads and are accepted and
    ignored; conditional connectivity is not available on this path."""
    if conditions and any(v is not None for k, v in conditions.items() if k.startswith("condition")):
        raise ValueError("conditional connectivity is not supported by the GPU circuit path")
    r, s = N.f64(resistance), N.f64(source)
    if r.ndim != 2 or s.shape != r.shape:
        raise ValueError("expected resistance and source rasters of the same shape")
    ny, nx = r.shape
    cur = np.empty((ny, nx))
    fp = np.empty((ny, nx)) if flow_potential else None
    nt = ctypes.c_int64(0)
    st = N.SolveStats()
    o = _opts()
    N.check(N.lib().lsc_circuit_omniscape(N.ptr(r), N.ptr(s), ctypes.c_int64(nx), ctypes.c_int64(ny), int(radius),
                                          int(buffer), int(block_size), int(bool(four_neighbors)),
                                          int(bool(resistance_is_conductance)), int(bool(correct_artifacts)),
                                          int(bool(flow_potential)), ctypes.byref(o), N.ptr(cur), N.ptr(fp),
                                          ctypes.byref(nt), ctypes.byref(st)))
    _record(st)
    out = {"current": cur, "n_targets": nt.value, "threads": 1}
    if flow_potential:
        out["flow_potential"] = fp
    return out


def circuit_omniscape_correction(radius, buffer, block_size, four_neighbors, nx, ny):
    """Omniscape's block-artifact correction factors (python.cpp:139-153)."""
    w = 2 * (int(radius) + int(buffer)) + 1
    out = np.empty((w, w))
    o = _opts()
    N.check(N.lib().lsc_circuit_omniscape_correction(int(radius), int(buffer), int(block_size),
                                                     int(bool(four_neighbors)), ctypes.c_int64(nx),
                                                     ctypes.c_int64(ny), ctypes.byref(o), N.ptr(out)))
    return out
