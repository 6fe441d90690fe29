"""Circuitscape's raster modes: advanced, pairwise and one-to-all / all-to-one.

The Python layer of the reference's src/landschaft/circuit/circuitscape.py
(:170-353, :366-516, :519-670, :673-953): the same functions, arguments,
defaults and results, over arrays (raster files are read by the caller). The
solves run in the GPU graph engine behind ``core()``.
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from ._backend import core
from .advanced import _grid

NODATA = -9999.0
POLICIES = ("keepall", "rmvsrc", "rmvgnd", "rmvall")


@dataclass
class FocalPoints:
    """Focal points as Circuitscape reads them: 1-based rows, columns and ids, sorted by id.

    Cells sharing an id form a focal region (circuitscape.py:170-206).
    """

    rows: np.ndarray
    cols: np.ndarray
    ids: np.ndarray

    @property
    def cells(self):
        """0-based ``(rows, cols)``."""
        return np.asarray(self.rows) - 1, np.asarray(self.cols) - 1

    @property
    def has_regions(self) -> bool:
        """Whether an id covers several cells."""
        return len(self.ids) != len(np.unique(self.ids))


@dataclass
class IncludedPairs:
    """Circuitscape's include / exclude pairs (circuitscape.py:232-280).

    ``mode`` is ``"include"`` (only the marked pairs are solved) or
    ``"exclude"``; ``matrix`` marks pairs of ``ids`` with 1.
    """

    mode: str
    ids: np.ndarray
    matrix: np.ndarray

    def excluded(self) -> set:
        """The (id, id) pairs that are not solved: unmarked ones in include mode, marked ones otherwise."""
        skip = 0 if self.mode == "include" else 1
        m = np.asarray(self.matrix)
        ids = [int(v) for v in self.ids]
        return {(ids[i], ids[j]) for i in range(len(ids)) for j in range(len(ids))
                if m[i, j] == skip and m[j, i] == skip}


@dataclass
class CircuitscapeResult:
    """Outputs of a raster run (circuitscape.py:366-403).

    ``voltage`` / ``current`` per cell (advanced), ``nodemap``,
    ``resistances`` (ids in the first row and column; -1: not connected),
    ``cum_current`` / ``max_current`` and per-pair or per-point ``maps``.
    """

    voltage: np.ndarray | None = None
    current: np.ndarray | None = None
    nodemap: np.ndarray | None = None
    resistances: np.ndarray | None = None
    cum_current: np.ndarray | None = None
    max_current: np.ndarray | None = None
    maps: dict = field(default_factory=dict)


def habitat_conductance(cells, *, is_resistance: bool = True, reclass: dict | None = None) -> np.ndarray:
    """Circuitscape's read_cellmap (circuitscape.py:329-348): reclassified, then 1/r, NoData cells at 0."""
    c = np.array(cells, dtype=np.float64)
    if reclass:
        lut = {float(k): float(v) for k, v in reclass.items()}
        c = np.vectorize(lambda v: lut.get(v, v), otypes=[np.float64])(c) if c.size else c
    nodata = c == NODATA
    if is_resistance:
        if np.any(c == 0):
            raise ValueError("zero resistance values are not supported for habitat maps; use a short-circuit "
                             "region file instead")
        g = 1.0 / c
    else:
        g = c.copy()
    g[nodata] = 0.0
    return g


def _array(value, name) -> np.ndarray:
    a, _ = _grid(value, name)
    return np.where(np.isnan(a), NODATA, a)


def _cells_from_map(values, name, shape) -> np.ndarray:
    a = _array(values, name)
    if a.shape != shape:
        raise ValueError(f"the {name} must have the habitat's shape")
    return a


def _polymap(polygons, shape):
    if polygons is None:
        return None
    pm = _cells_from_map(polygons, "polygons", shape)
    pm[pm == NODATA] = 0
    if not np.all(pm == np.round(pm)):
        raise ValueError("the polygon raster must hold integer ids")
    return np.ascontiguousarray(pm.astype(np.int64))


def _mask(g, mask):
    if mask is None:
        return g
    m = _cells_from_map(mask, "mask", g.shape)
    m[m == NODATA] = 0
    g = np.where(m > 0, g, 0.0)
    if g.sum() == 0:
        raise ValueError("Mask file deleted everything!")
    return g


def _as_points(points, shape) -> FocalPoints:
    """A raster of ids as focal points, in column-major cell order, stably sorted by id."""
    if isinstance(points, FocalPoints):
        return points
    pm = _cells_from_map(points, "points", shape)
    pm[pm == NODATA] = 0
    cols, rows = np.nonzero(pm.T)
    ids = pm[rows, cols]
    keep = ids >= 0
    rows, cols, ids = rows[keep], cols[keep], ids[keep]
    order = np.argsort(ids, kind="stable")
    return FocalPoints(rows[order] + 1, cols[order] + 1, ids[order].astype(np.int64))


# ---- advanced

def circuitscape_advanced(habitat, source, ground, *, habitat_is_resistance: bool = True,
                          ground_is_resistance: bool = True, polygons=None, mask=None, four_neighbors: bool = False,
                          avg_resistances: bool = False, remove_src_or_gnd: str = "keepall",
                          unit_currents: bool = False, direct_grounds: bool = False, reclass: dict | None = None,
                          n_threads: int = 0) -> CircuitscapeResult:
    """Circuitscape's advanced mode on rasters (circuitscape.py:406-496).

    Habitat cells with a conductance > 0 are nodes (a short-circuit polygon's
    cells one node); ``source`` amps per cell, ``ground`` a resistance (or
    conductance) to ground per cell, 0 resistance a direct ground, NoData none.
    ``remove_src_or_gnd`` resolves nodes that are both. Every component with a
    source and a ground is solved.
    """
    if remove_src_or_gnd not in POLICIES:
        raise ValueError(f"remove_src_or_gnd must be one of {POLICIES}")
    g = habitat_conductance(_array(habitat, "habitat"), is_resistance=habitat_is_resistance, reclass=reclass)
    poly = _polymap(polygons, g.shape)
    g = _mask(g, mask)
    src = _cells_from_map(source, "source", g.shape)
    src[src == NODATA] = 0.0
    gnd = _cells_from_map(ground, "ground", g.shape)
    missing = gnd == NODATA
    if ground_is_resistance:
        with np.errstate(divide="ignore"):
            gnd = 1.0 / gnd
    gnd[missing] = 0.0
    if unit_currents:
        src = np.where(src != 0, 1.0, src)
    if direct_grounds:
        gnd = np.where(gnd != 0, np.inf, gnd)
    r = core().circuitscape_advanced(np.ascontiguousarray(g), poly, np.ascontiguousarray(src),
                                     np.ascontiguousarray(gnd), bool(avg_resistances), bool(four_neighbors),
                                     remove_src_or_gnd, int(n_threads))
    nm = r["nodemap"]
    on = nm >= 0
    volt = np.zeros(nm.shape)
    cur = np.zeros(nm.shape)
    volt[on] = r["voltage"][nm[on]]
    cur[on] = r["current"][nm[on]]
    res = CircuitscapeResult(voltage=volt, current=cur, nodemap=nm)
    res.maps["solved"] = bool(r["solved"])
    return res


# ---- pairwise

def _pair_options(volt_maps, cur_maps, cum_only, max_maps, focal_zero, null_currents, null_voltages, log_transform):
    # circuitscape.py:543-547: a cum map needs the current maps; they are only kept without cum_only
    return {"volt_maps": bool(volt_maps), "cur_maps": bool(cur_maps and not cum_only),
            "cum": bool(cur_maps or cum_only), "max": bool(max_maps), "focal_zero": bool(focal_zero),
            "null_currents": bool(null_currents), "null_voltages": bool(null_voltages),
            "log_transform": bool(log_transform)}


def _apply_included_pairs(points: FocalPoints, pairs: IncludedPairs | None):
    # include mode drops the focal points it does not name (circuitscape.py:694-701)
    if pairs is None:
        return points, set()
    if pairs.mode == "include":
        keep = np.isin(points.ids, pairs.ids)
        points = FocalPoints(points.rows[keep], points.cols[keep], points.ids[keep])
    return points, pairs.excluded()


def circuitscape_pairwise(habitat, points, *, habitat_is_resistance: bool = True, polygons=None, mask=None,
                          four_neighbors: bool = False, avg_resistances: bool = False, included_pairs=None,
                          voltage_maps: bool = False, current_maps: bool = False, cum_map: bool = False,
                          max_map: bool = False, set_focal_node_currents_to_zero: bool = False,
                          set_null_currents_to_nodata: bool = False, set_null_voltages_to_nodata: bool = False,
                          log_transform_maps: bool = False, reclass: dict | None = None,
                          n_threads: int = 0) -> CircuitscapeResult:
    """Circuitscape's pairwise mode on rasters (circuitscape.py:550-670).

    Effective resistances between every pair of focal points: each component
    is factorised once and its pairs solved together (without maps or
    exclusions only the first point's pairs are solved and the rest derived,
    Circuitscape's shortcut). ``cum_map`` / ``max_map`` sum / take the maximum
    of the pairs' current maps; per-pair maps land in ``result.maps``.
    """
    g = habitat_conductance(_array(habitat, "habitat"), is_resistance=habitat_is_resistance, reclass=reclass)
    poly = _polymap(polygons, g.shape)
    g = _mask(g, mask)
    points = _as_points(points, g.shape)
    if included_pairs is not None and not isinstance(included_pairs, IncludedPairs):
        raise ValueError("included_pairs must be an IncludedPairs (files are read by the caller)")
    points, excluded = _apply_included_pairs(points, included_pairs)
    if points.has_regions:
        raise ValueError("focal regions are not supported by the GPU pairwise path")
    ids = points.ids
    n = len(ids)
    maps: dict = {}

    def emit(ci, cj, volt, cur):
        key = (int(ids[ci]), int(ids[cj]))
        if volt is not None:
            maps[("voltage", *key)] = volt
        if cur is not None:
            maps[("current", *key)] = cur

    opts = _pair_options(voltage_maps, current_maps or cum_map, not current_maps and cum_map, max_map,
                         set_focal_node_currents_to_zero, set_null_currents_to_nodata, set_null_voltages_to_nodata,
                         log_transform_maps)
    rows, cols = points.cells
    cells = np.ascontiguousarray(rows * g.shape[1] + cols, dtype=np.int64)
    ex = np.zeros((n, n), dtype=np.uint8)
    for a in range(n):
        for b in range(n):
            ex[a, b] = (int(ids[a]), int(ids[b])) in excluded
    r = core().circuitscape_pairwise(np.ascontiguousarray(g), poly, cells, ex, bool(avg_resistances),
                                     bool(four_neighbors), opts, emit, int(n_threads))
    res = np.zeros((n + 1, n + 1))
    res[0, 1:] = ids
    res[1:, 0] = ids
    res[1:, 1:] = r["resistances"]
    return CircuitscapeResult(resistances=res, cum_current=r.get("cum"), max_current=r.get("max"), maps=maps)


# ---- one-to-all / all-to-one

def _point_polymap(polymap, point_map: np.ndarray, ids: np.ndarray) -> np.ndarray:
    """Focal points and regions merged into the short-circuit regions (circuitscape.py:712-734)."""
    if polymap is None:
        return point_map.astype(np.int64)
    out = polymap.copy()
    cols, rows = np.nonzero(point_map.T)  # column-major
    if len(ids) == len(np.unique(ids)):
        top = polymap.max()
        for r, c in zip(rows, cols):
            if polymap[r, c] == 0:
                out[r, c] = point_map[r, c] + top
        return out
    top = max(polymap.max(), point_map.max())
    for r, c in zip(rows, cols):
        pid, cur = point_map[r, c], out[r, c]
        if cur == 0:
            out[r, c] = top + pid
        elif pid != cur:
            out[out == cur] = pid
    return out


def _one_to_all(g, poly, points: FocalPoints, strengths, pairs, one_to_all: bool, four, avg, opts, emit, n_threads):
    """run_onetoall (circuitscape.py:737-851): one advanced problem per focal id, batched per graph."""
    ny, nx = g.shape
    if pairs is not None:
        keep = np.isin(points.ids, pairs.ids)
        points = FocalPoints(points.rows[keep], points.cols[keep], points.ids[keep])
        if strengths is not None:
            strengths = strengths[np.isin(strengths[:, 0], pairs.ids)]
    rows, cols = points.cells
    cells = rows * nx + cols
    ids = points.ids
    point_map = np.zeros((ny, nx), dtype=np.int64)
    point_map[rows, cols] = ids
    uids, first = np.unique(ids, return_index=True)
    first_map = np.zeros((ny, nx), dtype=np.int64)
    first_map[rows[first], cols[first]] = uids
    single = len(ids) == len(uids)
    mark = 0 if pairs is not None and pairs.mode == "include" else 1
    base_poly = _point_polymap(poly, point_map, ids)
    res = np.zeros(len(uids))
    cum = np.zeros((ny, nx)) if opts["cum"] else None
    mx = np.full((ny, nx), NODATA) if opts["max"] else None
    policy = "rmvgnd" if one_to_all else "rmvsrc"

    def problem(i):
        pid = int(uids[i])
        pm = point_map.copy()
        newpoly = base_poly
        amps = float(strengths[i, 1]) if strengths is not None else 1.0
        if pairs is not None:
            row = int(np.nonzero(pairs.ids == pid)[0][0])
            for j in range(len(pairs.ids)):
                if j != row and pairs.matrix[row, j] == mark:
                    pm[pm == pairs.ids[j]] = 0
            if not single:
                newpoly = _point_polymap(poly, pm, ids)
        smap = None
        if strengths is not None:
            st = strengths[:, 1].copy()
            st[pm[rows, cols] == 0] = 1.0
            smap = np.zeros((ny, nx))
            smap[rows, cols] = st[: len(rows)]
        if pm.sum() == pid:
            return None
        if one_to_all:
            src = np.where(first_map == pid, amps, 0.0)
            gnd = np.where(pm == pid, 0, pm).astype(np.float64)
            gnd[gnd > 0] = np.inf
        else:
            if smap is not None:
                src = np.where(first_map == pid, 0.0, smap)
            else:
                src = np.where(first_map != 0, 1.0, 0.0)
                src = np.where(pm == pid, 0.0, src)
            gnd = np.where(pm == pid, np.inf, 0.0)
        sc, gc = np.flatnonzero(src), np.flatnonzero(gnd)
        return {"n": pid, "strength": amps, "poly": newpoly, "src_cells": sc.astype(np.int64),
                "src_vals": src.ravel()[sc], "gnd_cells": gc.astype(np.int64), "gnd_vals": gnd.ravel()[gc],
                "check_cell": int(cells[np.nonzero(ids == pid)[0][0]])}

    probs = [(i, problem(i)) for i in range(len(uids))]
    for i, q in probs:
        if q is None:
            res[i] = -1
    todo = [(i, q) for i, q in probs if q is not None]
    # one call per graph: consecutive problems on the same polygon map share it
    groups: list = []
    for i, q in todo:
        if groups and groups[-1][0] is q["poly"]:
            groups[-1][1].append((i, q))
        else:
            groups.append((q["poly"], [(i, q)]))
    if not (opts["volt_maps"] or opts["cur_maps"]):
        # only the cum / max maps and the point voltages: reduced on the GPU,
        # in problem order, into running maps
        cum_run = cum if cum is not None else np.zeros((ny, nx))
        for pm_, items in groups:
            out = core().circuitscape_multi_advanced_sum(
                np.ascontiguousarray(g), np.ascontiguousarray(pm_, dtype=np.int64), bool(avg), bool(four),
                [q for _, q in items], policy, cum_run, mx, bool(opts["focal_zero"]), int(n_threads))
            for (i, q), v, ok in zip(items, out["check_voltage"], out["solved"]):
                if not ok:
                    res[i] = -1
                elif not one_to_all:
                    res[i] = 0
                else:
                    val = v / q["strength"]
                    res[i] = -1 if val == 0 else val
        todo = []
    solved = {}
    for pm_, items in groups if todo else []:
        out = core().circuitscape_multi_advanced(np.ascontiguousarray(g), np.ascontiguousarray(pm_, dtype=np.int64),
                                                 bool(avg), bool(four), [q for _, q in items], policy,
                                                 int(n_threads))
        for (i, _), r in zip(items, out["results"]):
            solved[i] = (out["nodemap"], r)
    for i, q in todo:
        nm, r = solved[i]
        if not r["solved"]:
            res[i] = -1
        elif not one_to_all:
            res[i] = 0
        else:
            node = nm[first_map == q["n"]][0]
            val = 0.0 if node < 0 else r["voltage"][node] / q["strength"]
            res[i] = -1 if val == 0 else val
        on = nm >= 0
        volt = np.zeros((ny, nx))
        volt[on] = r["voltage"][nm[on]]
        cur = np.zeros((ny, nx))
        cur[on] = r["current"][nm[on]]
        if opts["focal_zero"]:
            cur[np.isin(nm, r["active"])] = 0.0
        if cum is not None:
            cum += cur
        if mx is not None:
            np.maximum(mx, cur, out=mx)
        emit(q["n"], volt if opts["volt_maps"] else None, cur if opts["cur_maps"] else None)
    if cum is not None:
        cum[cum < NODATA] = NODATA
    if mx is not None:
        mx[mx < NODATA] = NODATA
    table = np.column_stack([uids.astype(np.float64), res])
    return CircuitscapeResult(resistances=table, cum_current=cum, max_current=mx)


def circuitscape_one_to_all(habitat, points, *, all_to_one: bool = False, habitat_is_resistance: bool = True,
                            polygons=None, mask=None, four_neighbors: bool = False, avg_resistances: bool = False,
                            included_pairs=None, strengths=None, voltage_maps: bool = False,
                            current_maps: bool = False, cum_map: bool = False, max_map: bool = False,
                            set_focal_node_currents_to_zero: bool = False, reclass: dict | None = None,
                            n_threads: int = 0) -> CircuitscapeResult:
    """Circuitscape's one-to-all (or all-to-one) mode on rasters (circuitscape.py:854-953).

    For every focal id in turn: one-to-all injects 1 A (or its ``strengths``
    entry, a table of id and strength) there with every other focal point a
    direct ground; all-to-one grounds it and injects at the others.
    ``resistances`` holds the ids and the point's voltage over its strength
    (one-to-all; -1 where nothing was solved) or 0 (all-to-one).
    """
    g = habitat_conductance(_array(habitat, "habitat"), is_resistance=habitat_is_resistance, reclass=reclass)
    poly = _polymap(polygons, g.shape)
    g = _mask(g, mask)
    points = _as_points(points, g.shape)
    if included_pairs is not None and not isinstance(included_pairs, IncludedPairs):
        raise ValueError("included_pairs must be an IncludedPairs (files are read by the caller)")
    maps: dict = {}

    def emit(pid, volt, cur):
        if volt is not None:
            maps[("voltage", pid)] = volt
        if cur is not None:
            maps[("current", pid)] = cur

    opts = {"volt_maps": voltage_maps, "cur_maps": current_maps, "cum": current_maps or cum_map, "max": max_map,
            "focal_zero": set_focal_node_currents_to_zero}
    r = _one_to_all(g, poly, points, None if strengths is None else np.asarray(strengths, dtype=np.float64),
                    included_pairs, not all_to_one, four_neighbors, avg_resistances, opts, emit, n_threads)
    r.maps = maps
    return r
