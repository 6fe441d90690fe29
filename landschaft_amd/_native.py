"""ctypes binding of liblandschaft_circuit.so (the C-ABI in include/landschaft_circuit.h).

The library is the only compute path: when it is missing this module raises,
and no entry point falls back to a CPU solver.
"""

from __future__ import annotations

import ctypes
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "_lib" / "liblandschaft_circuit.so"

PRECISION_MIXED = 0
PRECISION_FP64 = 1
PRECISION_AUTO = 2

# Every symbol declared in include/landschaft_circuit.h.
EXPORTS = ("lsc_last_error", "lsc_device_info", "lsc_circuit_omniscape", "lsc_circuit_omniscape_device",
           "lsc_circuit_omniscape_device_static",
           "lsc_circuit_omniscape_correction", "lsc_probe_mfma", "lsc_window_plan_stats", "lsc_circuit_advanced",
           "lsc_circuitscape_advanced", "lsc_circuitscape_pairwise", "lsc_circuitscape_multi_advanced",
           "lsc_circuitscape_multi_advanced_sum", "lsc_network_advanced", "lsc_network_pairwise",
           "lsc_effective_resistance_grad", "lsc_grid_circuit_create", "lsc_grid_circuit_update",
           "lsc_grid_circuit_solve", "lsc_grid_circuit_adjoint", "lsc_grid_circuit_stats",
           "lsc_grid_circuit_destroy", "lsc_ascii_grid_rows", "lsc_circuitscape_analyze", "lsc_network_analyze")

class SolveOptions(ctypes.Structure):
    _fields_ = [("precision", ctypes.c_int32), ("max_refine", ctypes.c_int32), ("tol", ctypes.c_double),
                ("batch", ctypes.c_int64), ("timing", ctypes.c_int32), ("reserved", ctypes.c_int32),
                ("stream", ctypes.c_void_p)]


class OmniscapeCondition(ctypes.Structure):
    _fields_ = [("present", ctypes.POINTER(ctypes.c_double)), ("future", ctypes.POINTER(ctypes.c_double)),
                ("equal", ctypes.c_int32), ("reserved", ctypes.c_int32), ("lower", ctypes.c_double),
                ("upper", ctypes.c_double)]


class SolveStats(ctypes.Structure):
    _fields_ = [("n_targets", ctypes.c_int64), ("n_solved", ctypes.c_int64), ("n_unconverged", ctypes.c_int64),
                ("n_over_1e10", ctypes.c_int64), ("refine_max", ctypes.c_int32), ("n_batches", ctypes.c_int32),
                ("n_solve_launches", ctypes.c_int64), ("n_factored", ctypes.c_int64), ("refine_mean", ctypes.c_double), ("resid_max", ctypes.c_double), ("resid_mean", ctypes.c_double),
                ("ms_fill", ctypes.c_double), ("ms_flood", ctypes.c_double), ("ms_factor", ctypes.c_double),
                ("ms_solve", ctypes.c_double), ("ms_residual", ctypes.c_double),
                ("ms_currents", ctypes.c_double), ("ms_accumulate", ctypes.c_double),
                ("n_streams", ctypes.c_int32), ("n_fallback", ctypes.c_int32), ("resid_median", ctypes.c_double)]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_ if name != "reserved"}


class Analysis(ctypes.Structure):
    _fields_ = [("n_nodes", ctypes.c_int64), ("n_components", ctypes.c_int64), ("n_entries", ctypes.c_int64),
                ("n_fronts", ctypes.c_int64), ("factor_entries", ctypes.c_int64), ("largest_front", ctypes.c_int64),
                ("factor_flops", ctypes.c_double), ("graph_ms", ctypes.c_double), ("plan_ms", ctypes.c_double),
                ("graph_digest", ctypes.c_uint64), ("plan_digest", ctypes.c_uint64)]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


def _signatures() -> dict:
    """restype and argtypes of every export, as include/landschaft_circuit.h declares them: ctypes then
    converts (and range-checks) each argument instead of passing Python ints as C ints."""
    c = ctypes
    i32, i64, f64, vp = c.c_int32, c.c_int64, c.c_double, c.c_void_p
    pi32, pi64, pf64, pu8 = c.POINTER(i32), c.POINTER(i64), c.POINTER(f64), c.POINTER(c.c_uint8)
    ci = c.c_int
    opt, st = c.POINTER(SolveOptions), c.POINTER(SolveStats)
    cond = c.POINTER(OmniscapeCondition)
    omni = [vp, vp, i64, i64, ci, ci, ci, ci, ci, ci, ci, cond, opt, vp, vp, pi64, st]
    return {
        "lsc_last_error": (c.c_char_p, []),
        "lsc_device_info": (ci, [c.c_char_p, i64, pi32, pi64]),
        "lsc_circuit_omniscape": (ci, omni),
        "lsc_circuit_omniscape_device": (ci, omni),
        "lsc_circuit_omniscape_device_static": (ci, omni),
        "lsc_circuit_omniscape_correction": (ci, [ci, ci, ci, ci, i64, i64, opt, pf64]),
        "lsc_probe_mfma": (ci, [ci, pf64]),
        "lsc_window_plan_stats": (ci, [ci, ci, ci, ci, pf64]),
        "lsc_circuit_advanced": (ci, [pf64, pf64, pf64, i64, i64, ci, opt, pf64, pf64, st]),
        "lsc_circuitscape_advanced": (ci, [pf64, pi64, pf64, pf64, i64, i64, ci, ci, ci, opt, pi32, pf64, pf64,
                                           pi64, pi32, pi32, st]),
        "lsc_circuitscape_pairwise": (ci, [pf64, pi64, i64, i64, pi64, i64, pu8, ci, ci, ci, vp, vp, opt, pf64,
                                           pf64, pf64, st]),
        "lsc_circuitscape_multi_advanced": (ci, [pf64, pi64, i64, i64, ci, ci, ci, i64, pi64, pi64, pf64, pi64,
                                                 pi64, pf64, pi64, opt, pi32, pf64, pf64, pu8, pi32, st]),
        "lsc_circuitscape_multi_advanced_sum": (ci, [pf64, pi64, i64, i64, ci, ci, ci, i64, pi64, pi64, pf64,
                                                     pi64, pi64, pf64, pi64, ci, opt, pi32, pf64, pf64, pf64, pi32,
                                                     st]),
        "lsc_network_advanced": (ci, [pi32, pi32, pf64, i64, i32, pf64, pf64, ci, opt, pf64, pf64, pi32, pi32,
                                      pf64, pi64, pi32, pi32, st]),
        "lsc_network_pairwise": (ci, [pi32, pi32, pf64, i64, i32, pi32, i64, pu8, ci, vp, vp, opt, pf64, pf64,
                                      pi32, pi32, pf64, pi64, pi32, st]),
        "lsc_effective_resistance_grad": (ci, [pf64, i64, i64, ci, pi64, pi64, i64, opt, pf64, pf64, pf64, st]),
        "lsc_grid_circuit_create": (ci, [pf64, i64, i64, ci, opt, c.POINTER(vp)]),
        "lsc_grid_circuit_update": (ci, [vp, pf64]),
        "lsc_grid_circuit_solve": (ci, [vp, ci, vp, vp]),
        "lsc_grid_circuit_adjoint": (ci, [vp, ci, vp, vp, vp]),
        "lsc_grid_circuit_stats": (ci, [vp, st]),
        "lsc_grid_circuit_destroy": (None, [vp]),
        "lsc_ascii_grid_rows": (ci, [pf64, i64, i64, c.c_char_p, i64, pi64]),
        "lsc_circuitscape_analyze": (ci, [pf64, pi64, i64, i64, ci, ci, c.POINTER(Analysis)]),
        "lsc_network_analyze": (ci, [pi32, pi32, pf64, i64, i32, c.POINTER(Analysis)]),
    }


_lib = None


def lib() -> ctypes.CDLL:
    """The loaded library; raises when it has not been built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(f"{LIB_PATH} is missing: build it with `python -m landschaft_amd.build` "
                              "(landschaft_amd has no CPU fallback)")
        # torch's HIP runtime first: a process that uses both then shares one
        # runtime (torch cannot initialise the GPU after another runtime has)
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        handle = ctypes.CDLL(str(LIB_PATH))
        for name, (res, args) in _signatures().items():
            fn = getattr(handle, name)
            fn.restype, fn.argtypes = res, args
        _lib = handle
    return _lib


def check(rc: int) -> None:
    if rc == 0:
        return
    msg = lib().lsc_last_error().decode()
    raise (ValueError if rc == 1 else RuntimeError)(msg)


def options(precision="mixed", tol=None, max_refine=None, batch=None, timing=False, stream=None) -> SolveOptions:
    o = SolveOptions()
    o.precision = {"mixed": PRECISION_MIXED, "fp64": PRECISION_FP64, "auto": PRECISION_AUTO}[precision]
    o.tol = float(tol) if tol else 0.0
    o.max_refine = int(max_refine) if max_refine else 0
    o.batch = int(batch) if batch else 0
    o.timing = 1 if timing else 0
    o.stream = stream
    return o


def f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def ptr(a, t=ctypes.c_double):
    if a is None:
        return None
    return a.ctypes.data_as(ctypes.POINTER(t))


def device_info() -> dict:
    name = ctypes.create_string_buffer(64)
    cus = ctypes.c_int32(0)
    hbm = ctypes.c_int64(0)
    check(lib().lsc_device_info(name, ctypes.c_int64(64), ctypes.byref(cus), ctypes.byref(hbm)))
    return {"arch": name.value.decode(), "compute_units": cus.value, "hbm_bytes": hbm.value}


def probe_mfma(precision: int = PRECISION_MIXED) -> np.ndarray:
    out = np.empty(256)
    check(lib().lsc_probe_mfma(int(precision), ptr(out)))
    return out.reshape(16, 16)


def circuitscape_analyze(gmap, polymap=None, avg_resistances: bool = False, four_neighbors: bool = False) -> dict:
    """lsc_circuitscape_analyze: the node graph and every component's plan of a raster circuit (host only)."""
    g = f64(gmap)
    poly = None if polymap is None else np.ascontiguousarray(polymap, dtype=np.int64)
    out = Analysis()
    check(lib().lsc_circuitscape_analyze(ptr(g), ptr(poly, ctypes.c_int64), g.shape[1], g.shape[0],
                                         int(bool(avg_resistances)), int(bool(four_neighbors)), ctypes.byref(out)))
    return out.as_dict()


def network_analyze(a, b, w, n_nodes: int) -> dict:
    """lsc_network_analyze: the node graph and every component's plan of an edge list (host only)."""
    a = np.ascontiguousarray(a, dtype=np.int32)
    b = np.ascontiguousarray(b, dtype=np.int32)
    w = f64(w)
    if not (a.shape == b.shape == w.shape):
        raise ValueError("the edge arrays must have one length")
    out = Analysis()
    check(lib().lsc_network_analyze(ptr(a, ctypes.c_int32), ptr(b, ctypes.c_int32), ptr(w), a.size, int(n_nodes),
                                    ctypes.byref(out)))
    return out.as_dict()


def window_plan_stats(radius: int, buffer: int = 0, four_neighbors: bool = False, disk: bool = True) -> dict:
    out = np.empty(7)
    check(lib().lsc_window_plan_stats(int(radius), int(buffer), int(bool(four_neighbors)), int(bool(disk)),
                                      ptr(out)))
    keys = ("unknowns", "fronts", "l_entries", "factor_flops", "max_front", "levels", "l_unpacked")
    return dict(zip(keys, out.tolist()))
