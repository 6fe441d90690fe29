"""ctypes front end of the FP64 CPU oracle — TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this
module, as the checker. The product path (landschaft_amd) never imports it.
Each function mirrors the reference entry point it restates (paths relative to
the reference repository):

  advanced ................ src/cpp/circuit/python.cpp:65-79 (grid.cpp:73-210)
  omniscape ............... src/cpp/circuit/python.cpp:90-137 (omniscape.cpp:394-608)
  omniscape_correction .... src/cpp/circuit/python.cpp:139-153 (omniscape.cpp:326-392)
  cs_advanced ............. src/cpp/circuit/python.cpp:170-204 (circuitscape.cpp:393-468)
  cs_pairwise ............. src/cpp/circuit/python.cpp:214-277 (circuitscape.cpp:474-743)
  cs_multi_advanced ....... src/cpp/circuit/python.cpp:404-482 (circuitscape.cpp:904-1083)
  adjoint ................. new (no reference counterpart)
"""

from __future__ import annotations

import ctypes
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "_build" / "liboracle.so"
POLICIES = {"keepall": 0, "rmvsrc": 1, "rmvgnd": 2, "rmvall": 3}

_lib = None
_EMIT = ctypes.CFUNCTYPE(None, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                         ctypes.POINTER(ctypes.c_double), ctypes.c_void_p)


def build() -> Path:
    """Compile the oracle with its Makefile (g++ + OpenMP)."""
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        _lib = ctypes.CDLL(str(LIB))
        _lib.orc_last_error.restype = ctypes.c_char_p
    return _lib


def _p(a, t=ctypes.c_double):
    return None if a is None else a.ctypes.data_as(ctypes.POINTER(t))


def _check(rc):
    if rc == 0:
        return
    msg = lib().orc_last_error().decode()
    raise (ValueError if rc == 1 else RuntimeError)(msg)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def advanced(conductance, source, ground, four_neighbors=False, n_threads=0):
    g, s, gr = _f64(conductance), _f64(source), _f64(ground)
    ny, nx = g.shape
    cur, volt = np.empty((ny, nx)), np.empty((ny, nx))
    _check(lib().orc_advanced(_p(g), _p(s), _p(gr), ctypes.c_int64(nx), ctypes.c_int64(ny), int(four_neighbors),
                              int(n_threads), _p(cur), _p(volt)))
    return {"current": cur, "voltage": volt}


def omniscape(resistance, source, radius, buffer, block_size, four_neighbors, resistance_is_conductance,
              correct_artifacts, flow_potential, n_threads=0, limit=0):
    r, s = _f64(resistance), _f64(source)
    ny, nx = r.shape
    cur = np.empty((ny, nx))
    fp = np.empty((ny, nx)) if flow_potential else None
    nt = ctypes.c_int64(0)
    _check(lib().orc_omniscape(_p(r), _p(s), ctypes.c_int64(nx), ctypes.c_int64(ny), int(radius), int(buffer),
                               int(block_size), int(four_neighbors), int(resistance_is_conductance),
                               int(correct_artifacts), int(flow_potential), int(n_threads), ctypes.c_int64(limit),
                               _p(cur), _p(fp), ctypes.byref(nt)))
    out = {"current": cur, "n_targets": nt.value}
    if flow_potential:
        out["flow_potential"] = fp
    return out


def omniscape_conditions(resistance, source, radius, buffer, block_size, four_neighbors, resistance_is_conductance,
                         correct_artifacts, flow_potential, conditions, n_threads=0):
    """Omniscape with conditional connectivity; ``conditions`` as the binding's dict (condition1,
    condition1_future, comparison1, condition1_lower, condition1_upper, and the same for 2)."""
    r, s = _f64(resistance), _f64(source)
    ny, nx = r.shape
    keep, present, future, equal, lower, upper = [], [], [], [], [], []
    for k in (1, 2):
        p = conditions.get(f"condition{k}")
        f = conditions.get(f"condition{k}_future")
        p = None if p is None else _f64(p)
        f = None if f is None else _f64(f)
        keep += [p, f]
        present.append(_p(p))
        future.append(_p(f))
        equal.append(1 if conditions.get(f"comparison{k}", "within") == "equal" else 0)
        lower.append(float(conditions.get(f"condition{k}_lower", 0.0)))
        upper.append(float(conditions.get(f"condition{k}_upper", 0.0)))
    dp = ctypes.POINTER(ctypes.c_double)
    cur = np.empty((ny, nx))
    fp = np.empty((ny, nx)) if flow_potential else None
    nt = ctypes.c_int64(0)
    _check(lib().orc_omniscape_cond(_p(r), _p(s), ctypes.c_int64(nx), ctypes.c_int64(ny), int(radius), int(buffer),
                                    int(block_size), int(four_neighbors), int(resistance_is_conductance),
                                    int(correct_artifacts), int(flow_potential), int(n_threads), (dp * 2)(*present),
                                    (dp * 2)(*future), (ctypes.c_int * 2)(*equal), (ctypes.c_double * 2)(*lower),
                                    (ctypes.c_double * 2)(*upper), _p(cur), _p(fp), ctypes.byref(nt)))
    out = {"current": cur, "n_targets": nt.value}
    if flow_potential:
        out["flow_potential"] = fp
    return out


def omniscape_correction(radius, buffer, block_size, four_neighbors, nx, ny):
    w = 2 * (radius + buffer) + 1
    out = np.empty((w, w))
    _check(lib().orc_omniscape_correction(int(radius), int(buffer), int(block_size), int(four_neighbors),
                                          ctypes.c_int64(nx), ctypes.c_int64(ny), _p(out)))
    return out


def cs_advanced(gmap, polymap, source, ground, avg_resistances, four_neighbors, policy):
    g, s, gr = _f64(gmap), _f64(source), _f64(ground)
    ny, nx = g.shape
    poly = None if polymap is None else np.ascontiguousarray(polymap, dtype=np.int64)
    nm = np.empty((ny, nx), dtype=np.int32)
    volt, cur = np.empty((ny, nx)), np.empty((ny, nx))
    solved, ncomp = ctypes.c_int(0), ctypes.c_int(0)
    _check(lib().orc_cs_advanced(_p(g), _p(poly, ctypes.c_int64), _p(s), _p(gr), ctypes.c_int64(nx),
                                 ctypes.c_int64(ny), int(avg_resistances), int(four_neighbors), POLICIES[policy],
                                 _p(nm, ctypes.c_int32), _p(volt), _p(cur), ctypes.byref(solved),
                                 ctypes.byref(ncomp)))
    return {"nodemap": nm, "voltage": volt, "current": cur, "solved": bool(solved.value),
            "n_components": ncomp.value}


PAIR_FLAGS = ("volt_maps", "cur_maps", "cum", "max", "focal_zero", "null_currents", "null_voltages",
              "log_transform")


def cs_pairwise(gmap, polymap, point_cells, excluded, avg_resistances, four_neighbors, options, emit=None):
    g = _f64(gmap)
    ny, nx = g.shape
    poly = None if polymap is None else np.ascontiguousarray(polymap, dtype=np.int64)
    pts = np.ascontiguousarray(point_cells, dtype=np.int64)
    n = pts.size
    ex = np.ascontiguousarray(excluded, dtype=np.uint8).reshape(n, n)
    flags = sum(1 << k for k, f in enumerate(PAIR_FLAGS) if options.get(f))
    R = np.empty((n, n))
    cum = np.empty((ny, nx)) if options.get("cum") else None
    mx = np.empty((ny, nx)) if options.get("max") else None

    def _cb(ci, cj, v, c, _ctx):
        if emit is None:
            return
        vv = None if not v else np.ctypeslib.as_array(v, shape=(ny, nx)).copy()
        cc = None if not c else np.ctypeslib.as_array(c, shape=(ny, nx)).copy()
        emit(ci, cj, vv, cc)

    cb = _EMIT(_cb)
    _check(lib().orc_cs_pairwise(_p(g), _p(poly, ctypes.c_int64), ctypes.c_int64(nx), ctypes.c_int64(ny),
                                 _p(pts, ctypes.c_int64), ctypes.c_int64(n), _p(ex, ctypes.c_uint8),
                                 int(avg_resistances), int(four_neighbors), flags, cb, None, _p(R), _p(cum),
                                 _p(mx)))
    out = {"resistances": R}
    if cum is not None:
        out["cum"] = cum
    if mx is not None:
        out["max"] = mx
    return out


def cs_multi_advanced(gmap, polymap, avg_resistances, four_neighbors, problems, policy, schur=True):
    g = _f64(gmap)
    ny, nx = g.shape
    poly = None if polymap is None else np.ascontiguousarray(polymap, dtype=np.int64)
    npb = len(problems)

    def cat(key_c, key_v):
        ptr = np.zeros(npb + 1, dtype=np.int64)
        for k, q in enumerate(problems):
            ptr[k + 1] = ptr[k] + len(q[key_c])
        cells = np.ascontiguousarray(np.concatenate([np.asarray(q[key_c], dtype=np.int64) for q in problems])
                                     if npb else np.zeros(0, np.int64))
        vals = np.ascontiguousarray(np.concatenate([np.asarray(q[key_v], dtype=np.float64) for q in problems])
                                    if npb else np.zeros(0))
        return ptr, cells, vals

    sp, sc, sv = cat("src_cells", "src_vals")
    gp, gc, gv = cat("gnd_cells", "gnd_vals")
    check = np.array([q["check_cell"] for q in problems], dtype=np.int64)
    nm = np.empty((ny, nx), dtype=np.int32)
    volt = np.empty((npb, ny, nx))
    cur = np.empty((npb, ny, nx))
    act = np.empty((npb, ny, nx), dtype=np.uint8)
    solved = np.empty(npb, dtype=np.int32)
    i64p = ctypes.c_int64
    _check(lib().orc_cs_multi_advanced(_p(g), _p(poly, i64p), i64p(nx), i64p(ny), int(avg_resistances),
                                       int(four_neighbors), POLICIES[policy], i64p(npb), _p(sp, i64p),
                                       _p(sc, i64p), _p(sv), _p(gp, i64p), _p(gc, i64p), _p(gv),
                                       _p(check, i64p), int(schur), _p(nm, ctypes.c_int32), _p(volt), _p(cur),
                                       _p(act, ctypes.c_uint8), _p(solved, ctypes.c_int32)))
    return {"nodemap": nm, "voltage": volt, "current": cur, "active": act.astype(bool),
            "solved": solved.astype(bool)}


def adjoint(resistance, src_cells, gnd_cells, four_neighbors=False, n_threads=0):
    r = _f64(resistance)
    ny, nx = r.shape
    s = np.ascontiguousarray(src_cells, dtype=np.int64)
    t = np.ascontiguousarray(gnd_cells, dtype=np.int64)
    J = ctypes.c_double(0)
    R = np.empty(s.size)
    grad = np.empty((ny, nx))
    _check(lib().orc_adjoint(_p(r), ctypes.c_int64(nx), ctypes.c_int64(ny), int(four_neighbors),
                             _p(s, ctypes.c_int64), _p(t, ctypes.c_int64), ctypes.c_int64(s.size),
                             int(n_threads), ctypes.byref(J), _p(R), _p(grad)))
    return J.value, R, grad


def window_stats(radius, buffer=0, four_neighbors=False):
    out = np.empty(5)
    _check(lib().orc_window_stats(int(radius), int(buffer), int(four_neighbors), _p(out)))
    keys = ("unknowns", "fronts", "l_entries", "factor_flops", "max_front")
    return dict(zip(keys, out.tolist()))


def threads_default() -> int:
    return os.cpu_count() or 1
