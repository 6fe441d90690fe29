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
    """How circuits are solved: ``precision`` "auto" (the default: Omniscape
    windows mixed; a graph factorisation FP64 when it serves at least 8
    right-hand sides, mixed otherwise), "mixed" (FP32 MFMA fronts + FP64
    refinement) or "fp64"; ``tol`` the relative residual refinement reaches;
    ``batch`` circuits per batch."""

    precision: str = "auto"
    tol: float = 1e-12
    max_refine: int = 30
    batch: int = 16384


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
    """Omniscape's cumulative current (python.cpp:90-137), conditions as the binding's dict
    (condition1, condition1_future, comparison1, condition1_lower, condition1_upper; the same for 2).
    ``n_threads`` and ``memory_limit`` size the reference's CPU threads and are accepted and ignored."""
    r, s = N.f64(resistance), N.f64(source)
    if r.ndim != 2 or s.shape != r.shape:
        raise ValueError("expected resistance and source rasters of the same shape")
    ny, nx = r.shape
    conds = None
    keep = []
    if conditions and conditions.get("condition1") is not None:
        conds = (N.OmniscapeCondition * 2)()
        for i, k in enumerate((1, 2)):
            present = conditions.get(f"condition{k}")
            if present is None:
                continue
            cmp = conditions.get(f"comparison{k}", "within")
            if cmp not in ("within", "equal"):
                raise ValueError("comparisons must be 'within' or 'equal'")
            p = N.f64(present)
            if p.shape != r.shape:
                raise ValueError(f"condition{k} must have the resistance's shape")
            f = conditions.get(f"condition{k}_future")
            f = None if f is None else N.f64(f)
            if f is not None and f.shape != r.shape:
                raise ValueError(f"condition{k}_future must have the resistance's shape")
            keep += [p, f]
            conds[i].present = N.ptr(p)
            conds[i].future = N.ptr(f)
            conds[i].equal = 1 if cmp == "equal" else 0
            conds[i].lower = float(conditions.get(f"condition{k}_lower", 0.0))
            conds[i].upper = float(conditions.get(f"condition{k}_upper", 0.0))
    cur = np.empty((ny, nx))
    fp = np.empty((ny, nx)) if flow_potential else None
    nt = ctypes.c_int64(0)
    st = N.SolveStats()
    o = _opts()
    N.check(N.lib().lsc_circuit_omniscape(N.ptr(r), N.ptr(s), ctypes.c_int64(nx), ctypes.c_int64(ny), int(radius),
                                          int(buffer), int(block_size), int(bool(four_neighbors)),
                                          int(bool(resistance_is_conductance)), int(bool(correct_artifacts)),
                                          int(bool(flow_potential)), conds, ctypes.byref(o), N.ptr(cur), N.ptr(fp),
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


_POLICIES = {"keepall": 0, "rmvsrc": 1, "rmvgnd": 2, "rmvall": 3}
_PAIR_FLAGS = ("volt_maps", "cur_maps", "cum", "max", "focal_zero", "null_currents", "null_voltages",
               "log_transform")
_EMIT = ctypes.CFUNCTYPE(None, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                         ctypes.POINTER(ctypes.c_double), ctypes.c_void_p)


def circuit_advanced(conductance, source, ground, four_neighbors=False, n_threads=0):
    """Circuitscape's advanced mode on a raster (python.cpp:65-79): node currents and voltages."""
    g, s, gr = N.f64(conductance), N.f64(source), N.f64(ground)
    if g.ndim != 2 or s.shape != g.shape or gr.shape != g.shape:
        raise ValueError("expected conductance, source and ground rasters of the same shape")
    ny, nx = g.shape
    cur, volt = np.empty((ny, nx)), np.empty((ny, nx))
    st = N.SolveStats()
    o = _opts()
    N.check(N.lib().lsc_circuit_advanced(N.ptr(g), N.ptr(s), N.ptr(gr), ctypes.c_int64(nx), ctypes.c_int64(ny),
                                         int(bool(four_neighbors)), ctypes.byref(o), N.ptr(cur), N.ptr(volt),
                                         ctypes.byref(st)))
    _record(st)
    return {"current": cur, "voltage": volt}


def _poly(polymap, shape):
    if polymap is None:
        return None
    p = np.ascontiguousarray(polymap, dtype=np.int64)
    if p.shape != shape:
        raise ValueError("the polygon raster must have the habitat's shape")
    return p


def _policy(policy):
    if policy not in _POLICIES:
        raise ValueError("remove_src_or_gnd must be keepall, rmvsrc, rmvgnd or rmvall")
    return _POLICIES[policy]


def circuitscape_advanced(gmap, polymap, source, ground, avg_resistances, four_neighbors, policy, n_threads=0):
    """Advanced mode on the raster node graph (python.cpp:170-204): node map, per-node voltage and current."""
    g, s, gr = N.f64(gmap), N.f64(source), N.f64(ground)
    if g.ndim != 2 or s.shape != g.shape or gr.shape != g.shape:
        raise ValueError("expected habitat, source and ground rasters of the same shape")
    ny, nx = g.shape
    poly = _poly(polymap, g.shape)
    nm = np.empty((ny, nx), dtype=np.int32)
    volt, cur = np.empty(nx * ny), np.empty(nx * ny)
    nn, solved, ncomp = ctypes.c_int64(0), ctypes.c_int32(0), ctypes.c_int32(0)
    st = N.SolveStats()
    o = _opts()
    N.check(N.lib().lsc_circuitscape_advanced(N.ptr(g), N.ptr(poly, ctypes.c_int64), N.ptr(s), N.ptr(gr),
                                              ctypes.c_int64(nx), ctypes.c_int64(ny), int(bool(avg_resistances)),
                                              int(bool(four_neighbors)), _policy(policy), ctypes.byref(o),
                                              N.ptr(nm, ctypes.c_int32), N.ptr(volt), N.ptr(cur), ctypes.byref(nn),
                                              ctypes.byref(solved), ctypes.byref(ncomp), ctypes.byref(st)))
    _record(st)
    return {"nodemap": nm, "voltage": volt[:nn.value].copy(), "current": cur[:nn.value].copy(),
            "solved": bool(solved.value), "n_components": ncomp.value}


def circuitscape_pairwise(gmap, polymap, point_cells, excluded, avg_resistances, four_neighbors, options, emit=None,
                          n_threads=0):
    """Pairwise mode (python.cpp:214-277): resistances, cum / max maps; per-pair maps through `emit`.
    With ``options["region_ids"]`` (the id of every point cell, sorted) the points are focal regions,
    solved one graph per pair of ids."""
    g = N.f64(gmap)
    if g.ndim != 2:
        raise ValueError("expected a 2-D habitat raster")
    if options.get("region_ids") is not None:
        return pairwise_regions(_pairwise_points, g, polymap, point_cells, options["region_ids"], excluded,
                                avg_resistances, four_neighbors, options, emit)
    return _pairwise_points(g, polymap, point_cells, excluded, avg_resistances, four_neighbors, options, emit)


def region_polymap(polymap, ncell, cells, ids, a, b):
    """The short-circuit map of one pair of focal regions (circuitscape.cpp:749-779), flat.

    Without polygons the cells of ids ``a`` and ``b`` are polygons of those
    ids. With polygons, a region of several cells becomes one new polygon
    (numbered past the largest id) that takes in every polygon it overlaps;
    a single-cell point is left as it is.
    """
    if polymap is None:
        poly = np.zeros(ncell, dtype=np.int64)
        for pid in (a, b):
            poly[cells[ids == pid]] = pid
        return poly
    poly = np.array(polymap, dtype=np.int64).ravel()
    top = int(poly.max())
    for pid in (a, b):
        mine = cells[ids == pid]
        if len(mine) == 1:
            continue
        top += 1
        hit = np.unique(poly[mine])
        hit = hit[hit != 0]
        if hit.size:
            poly[np.isin(poly, hit)] = top
        poly[mine] = top
    return poly


def pairwise_regions(solve, gmap, polymap, point_cells, region_ids, excluded, avg_resistances, four_neighbors,
                     options, emit=None):
    """raster_pairwise_regions (circuitscape.cpp:781-855): every pair of distinct ids not excluded
    gets its own graph (``region_polymap``) and one two-point pairwise solve by ``solve`` (a
    pairwise function taking this module's arguments), between the first cells of the two ids.
    Resistances -1 where a pair is not solved, 0 on the diagonal; the cum map is the sum of the
    pairs' cum maps, the max map their maximum, both floored at -9999."""
    g = np.asarray(gmap, dtype=np.float64)
    ny, nx = g.shape
    cells = np.asarray(point_cells, dtype=np.int64).ravel()
    ids = np.asarray(region_ids, dtype=np.int64).ravel()
    if cells.size != ids.size:
        raise ValueError("region_ids needs one id per point cell")
    pts = ids[np.r_[True, ids[1:] != ids[:-1]]] if ids.size else ids
    n = len(pts)
    ex = np.asarray(excluded, dtype=np.uint8).reshape(n, n)
    R = np.full((n, n), -1.0)
    cum = np.zeros((ny, nx)) if options.get("cum") else None
    mx = np.full((ny, nx), -9999.0) if options.get("max") else None
    sub = {k: v for k, v in options.items() if k != "region_ids"}
    for i in range(n):
        for j in range(i + 1, n):
            if ex[i, j] or ex[j, i]:
                continue
            poly = region_polymap(polymap, nx * ny, cells, ids, pts[i], pts[j]).reshape(ny, nx)
            ends = np.array([cells[np.flatnonzero(ids == pts[i])[0]], cells[np.flatnonzero(ids == pts[j])[0]]],
                            dtype=np.int64)

            def remap(ci, cj, v, c, i=i, j=j):
                if emit is not None:
                    emit(i if ci == 0 else j, i if cj == 0 else j, v, c)

            r = solve(g, poly, ends, np.zeros((2, 2), dtype=np.uint8), avg_resistances, four_neighbors, sub, remap)
            R[i, j] = R[j, i] = r["resistances"][0, 1]
            if cum is not None:
                cum += r["cum"]
            if mx is not None:
                np.maximum(mx, r["max"], out=mx)
    np.fill_diagonal(R, 0.0)
    out = {"resistances": R, "nodes": None}
    if cum is not None:
        cum[cum < -9999.0] = -9999.0
        out["cum"] = cum
    if mx is not None:
        out["max"] = mx
    return out


def _pairwise_points(g, polymap, point_cells, excluded, avg_resistances, four_neighbors, options, emit=None):
    ny, nx = g.shape
    poly = _poly(polymap, g.shape)
    pts = np.ascontiguousarray(point_cells, dtype=np.int64).ravel()
    n = pts.size
    ex = np.ascontiguousarray(excluded, dtype=np.uint8).reshape(n, n)
    flags = sum(1 << k for k, f in enumerate(_PAIR_FLAGS) if options.get(f))
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
    st = N.SolveStats()
    o = _opts()
    N.check(N.lib().lsc_circuitscape_pairwise(N.ptr(g), N.ptr(poly, ctypes.c_int64), ctypes.c_int64(nx),
                                              ctypes.c_int64(ny), N.ptr(pts, ctypes.c_int64), ctypes.c_int64(n),
                                              N.ptr(ex, ctypes.c_uint8), int(bool(avg_resistances)),
                                              int(bool(four_neighbors)), flags, cb, None, ctypes.byref(o), N.ptr(R),
                                              N.ptr(cum), N.ptr(mx), ctypes.byref(st)))
    _record(st)
    out = {"resistances": R, "nodes": None}
    if cum is not None:
        out["cum"] = cum
    if mx is not None:
        out["max"] = mx
    return out


def circuitscape_multi_advanced(gmap, polymap, avg_resistances, four_neighbors, problems, policy, n_threads=0):
    """Several advanced problems on one graph (python.cpp:404-482): per problem voltage, current, active, solved."""
    g = N.f64(gmap)
    ny, nx = g.shape
    poly = _poly(polymap, g.shape)
    npb = len(problems)

    def cat(kc, kv):
        ptr = np.zeros(npb + 1, dtype=np.int64)
        for k, q in enumerate(problems):
            ptr[k + 1] = ptr[k] + len(q[kc])
        cells = np.ascontiguousarray(np.concatenate([np.asarray(q[kc], dtype=np.int64) for q in problems])
                                     if npb else np.zeros(0, np.int64))
        vals = np.ascontiguousarray(np.concatenate([np.asarray(q[kv], dtype=np.float64) for q in problems])
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
    st = N.SolveStats()
    o = _opts()
    i64p = ctypes.c_int64
    N.check(N.lib().lsc_circuitscape_multi_advanced(
        N.ptr(g), N.ptr(poly, i64p), i64p(nx), i64p(ny), int(bool(avg_resistances)), int(bool(four_neighbors)),
        _policy(policy), i64p(npb), N.ptr(sp, i64p), N.ptr(sc, i64p), N.ptr(sv), N.ptr(gp, i64p), N.ptr(gc, i64p),
        N.ptr(gv), N.ptr(check, i64p), ctypes.byref(o), N.ptr(nm, ctypes.c_int32), N.ptr(volt), N.ptr(cur),
        N.ptr(act, ctypes.c_uint8), N.ptr(solved, ctypes.c_int32), ctypes.byref(st)))
    _record(st)
    # per-node arrays, as the binding returns them (every node has a cell)
    on = nm >= 0
    n_nodes = int(nm.max()) + 1 if on.any() else 0
    results = []
    for p in range(npb):
        v = np.zeros(n_nodes)
        c = np.zeros(n_nodes)
        a = np.zeros(n_nodes, dtype=bool)
        v[nm[on]] = volt[p][on]
        c[nm[on]] = cur[p][on]
        a[nm[on]] = act[p][on].astype(bool)
        results.append({"voltage": v, "current": c, "active": np.flatnonzero(a).astype(np.int32),
                        "solved": bool(solved[p])})
    return {"nodemap": nm, "results": results}


def _problem_arrays(problems):
    npb = len(problems)

    def cat(kc, kv):
        ptr = np.zeros(npb + 1, dtype=np.int64)
        for k, q in enumerate(problems):
            ptr[k + 1] = ptr[k] + len(q[kc])
        cells = np.ascontiguousarray(np.concatenate([np.asarray(q[kc], dtype=np.int64) for q in problems])
                                     if npb else np.zeros(0, np.int64))
        vals = np.ascontiguousarray(np.concatenate([np.asarray(q[kv], dtype=np.float64) for q in problems])
                                    if npb else np.zeros(0))
        return ptr, cells, vals

    sp, sc, sv = cat("src_cells", "src_vals")
    gp, gc, gv = cat("gnd_cells", "gnd_vals")
    check = np.array([q["check_cell"] for q in problems], dtype=np.int64)
    return sp, sc, sv, gp, gc, gv, check


def circuitscape_multi_advanced_sum(gmap, polymap, avg_resistances, four_neighbors, problems, policy, cum, max_map,
                                    focal_zero=False, n_threads=0):
    """The problems of circuitscape_multi_advanced reduced to the cum / max current maps (updated in place,
    ny x nx f64; max_map may be None), as run_onetoall sums them (circuitscape.py:826-848), plus each
    problem's voltage at its check cell and whether it was solved. The per-problem maps stay on the GPU."""
    g = N.f64(gmap)
    ny, nx = g.shape
    poly = _poly(polymap, g.shape)
    npb = len(problems)
    if cum.shape != g.shape or cum.dtype != np.float64 or not cum.flags.c_contiguous:
        raise ValueError("cum must be a C-contiguous float64 raster of the habitat's shape")
    if max_map is not None and (max_map.shape != g.shape or max_map.dtype != np.float64
                                or not max_map.flags.c_contiguous):
        raise ValueError("max_map must be a C-contiguous float64 raster of the habitat's shape")
    sp, sc, sv, gp, gc, gv, check = _problem_arrays(problems)
    nm = np.empty((ny, nx), dtype=np.int32)
    vcheck = np.zeros(npb)
    solved = np.zeros(npb, dtype=np.int32)
    st = N.SolveStats()
    o = _opts()
    i64p = ctypes.c_int64
    N.check(N.lib().lsc_circuitscape_multi_advanced_sum(
        N.ptr(g), N.ptr(poly, i64p), i64p(nx), i64p(ny), int(bool(avg_resistances)), int(bool(four_neighbors)),
        _policy(policy), i64p(npb), N.ptr(sp, i64p), N.ptr(sc, i64p), N.ptr(sv), N.ptr(gp, i64p), N.ptr(gc, i64p),
        N.ptr(gv), N.ptr(check, i64p), int(bool(focal_zero)), ctypes.byref(o), N.ptr(nm, ctypes.c_int32),
        N.ptr(cum), N.ptr(max_map), N.ptr(vcheck), N.ptr(solved, ctypes.c_int32), ctypes.byref(st)))
    _record(st)
    return {"nodemap": nm, "check_voltage": vcheck, "solved": solved.astype(bool)}


def ascii_grid_rows(values) -> bytes:
    """The body of an ESRI ASCII grid (python.cpp:484-521): every value in its shortest round-trip
    text, NaN as -9999, rows ended by newlines."""
    a = N.f64(values)
    if a.ndim != 2:
        raise ValueError("expected a 2-D array")
    ny, nx = a.shape
    size = ctypes.c_int64(0)
    N.check(N.lib().lsc_ascii_grid_rows(N.ptr(a), nx, ny, None, 0, ctypes.byref(size)))
    buf = ctypes.create_string_buffer(max(size.value, 1))
    N.check(N.lib().lsc_ascii_grid_rows(N.ptr(a), nx, ny, buf, size.value, ctypes.byref(size)))
    return buf.raw[:size.value]


def effective_resistance_grad(resistance, src_cells, gnd_cells, four_neighbors=False):
    """J = sum of effective resistances of the pairs and dJ/dr per cell (the adjoint on the GPU)."""
    r = N.f64(resistance)
    ny, nx = r.shape
    s = np.ascontiguousarray(src_cells, dtype=np.int64).ravel()
    t = np.ascontiguousarray(gnd_cells, dtype=np.int64).ravel()
    J = ctypes.c_double(0)
    R = np.empty(s.size)
    grad = np.empty((ny, nx))
    st = N.SolveStats()
    o = _opts()
    N.check(N.lib().lsc_effective_resistance_grad(N.ptr(r), ctypes.c_int64(nx), ctypes.c_int64(ny),
                                                  int(bool(four_neighbors)), N.ptr(s, ctypes.c_int64),
                                                  N.ptr(t, ctypes.c_int64), ctypes.c_int64(s.size), ctypes.byref(o),
                                                  ctypes.byref(J), N.ptr(R), N.ptr(grad), ctypes.byref(st)))
    _record(st)
    return J.value, R, grad


# ---- networks

_NET_EMIT = ctypes.CFUNCTYPE(None, ctypes.c_int, ctypes.c_int, ctypes.c_int64, ctypes.POINTER(ctypes.c_int32),
                             ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double), ctypes.c_int64,
                             ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32),
                             ctypes.POINTER(ctypes.c_double), ctypes.c_void_p)
_NET_FLAGS = ("volt_maps", "cur_maps", "cum", "max")


def _network_edges(a, b, w, n_nodes):
    a = np.ascontiguousarray(a, dtype=np.int32).ravel()
    b = np.ascontiguousarray(b, dtype=np.int32).ravel()
    w = N.f64(w).ravel()
    if not a.size == b.size == w.size:
        raise ValueError("edge arrays of different lengths")
    if int(n_nodes) <= 0:
        raise ValueError("a network needs nodes")
    return a, b, w, int(n_nodes)


def _i32(n):
    return np.empty(max(int(n), 1), dtype=np.int32)


def network_advanced(a, b, w, n_nodes, source, ground, policy, n_threads=0):
    """Advanced mode on a network (python.cpp:310-346): voltage and current by node, the branch
    currents as ``(r, c, current)`` arrays (nodes r < c), solved, n_components."""
    a, b, w, m = _network_edges(a, b, w, n_nodes)
    s, g = N.f64(source).ravel(), N.f64(ground).ravel()
    if s.size != m or g.size != m:
        raise ValueError("sources and grounds must have one value per node")
    volt, cur = np.empty(m), np.empty(m)
    br, bc, bv = _i32(a.size), _i32(a.size), np.empty(max(a.size, 1))
    nb, solved, ncomp = ctypes.c_int64(0), ctypes.c_int32(0), ctypes.c_int32(0)
    st = N.SolveStats()
    o = _opts()
    i32p = ctypes.c_int32
    N.check(N.lib().lsc_network_advanced(N.ptr(a, i32p), N.ptr(b, i32p), N.ptr(w), ctypes.c_int64(a.size),
                                         ctypes.c_int32(m), N.ptr(s), N.ptr(g), _policy(policy), ctypes.byref(o),
                                         N.ptr(volt), N.ptr(cur), N.ptr(br, i32p), N.ptr(bc, i32p), N.ptr(bv),
                                         ctypes.byref(nb), ctypes.byref(solved), ctypes.byref(ncomp),
                                         ctypes.byref(st)))
    _record(st)
    k = nb.value
    return {"voltage": volt, "current": cur, "branch": (br[:k].copy(), bc[:k].copy(), bv[:k].copy()),
            "solved": bool(solved.value), "n_components": ncomp.value}


def network_pairwise(a, b, w, n_nodes, points, excluded, options, emit=None, n_threads=0):
    """Pairwise mode on a network (python.cpp:348-402): resistances, n_components; with any map option
    ``emit(ci, cj, nodes, voltage or None, current, (r, c, branch current))`` per pair, and ``cum``
    (by node) and ``cum_branch`` ((r, c, current) arrays) summed over the pairs."""
    a, b, w, m = _network_edges(a, b, w, n_nodes)
    pts = np.ascontiguousarray(points, dtype=np.int32).ravel()
    n = pts.size
    ex = np.ascontiguousarray(excluded, dtype=np.uint8).reshape(n, n)
    flags = sum(1 << k for k, f in enumerate(_NET_FLAGS) if options.get(f))
    R = np.empty((n, n))
    cum = np.empty(m)
    br, bc, bv = _i32(a.size), _i32(a.size), np.empty(max(a.size, 1))
    nb, ncomp = ctypes.c_int64(0), ctypes.c_int32(0)

    def _cb(ci, cj, count, nodes, v, c, nbr, rr, cc, vv, _ctx):
        if emit is None:
            return
        as_arr = np.ctypeslib.as_array
        nd = as_arr(nodes, shape=(count,)).copy() if count else np.zeros(0, np.int32)
        volt = as_arr(v, shape=(count,)).copy() if v and count else (np.zeros(0) if v else None)
        cur = as_arr(c, shape=(count,)).copy() if count else np.zeros(0)
        if nbr:
            branch = (as_arr(rr, shape=(nbr,)).copy(), as_arr(cc, shape=(nbr,)).copy(),
                      as_arr(vv, shape=(nbr,)).copy())
        else:
            branch = (np.zeros(0, np.int32), np.zeros(0, np.int32), np.zeros(0))
        emit(ci, cj, nd, volt, cur, branch)

    cb = _NET_EMIT(_cb)
    st = N.SolveStats()
    o = _opts()
    i32p = ctypes.c_int32
    N.check(N.lib().lsc_network_pairwise(N.ptr(a, i32p), N.ptr(b, i32p), N.ptr(w), ctypes.c_int64(a.size),
                                         ctypes.c_int32(m), N.ptr(pts, i32p), ctypes.c_int64(n),
                                         N.ptr(ex, ctypes.c_uint8), flags, cb, None, ctypes.byref(o), N.ptr(R),
                                         N.ptr(cum), N.ptr(br, i32p), N.ptr(bc, i32p), N.ptr(bv),
                                         ctypes.byref(nb), ctypes.byref(ncomp), ctypes.byref(st)))
    _record(st)
    out = {"resistances": R, "n_components": ncomp.value}
    if flags:
        k = nb.value
        out["cum"] = cum
        out["cum_branch"] = (br[:k].copy(), bc[:k].copy(), bv[:k].copy())
    return out
