"""Circuitscape's network (graph) mode: pairwise and advanced on an edge list.

Mirrors the reference's src/landschaft/circuit/network.py (readers :24-113,
result :116-157, advanced :177-262, pairwise :266-374, INI :377-446); the
solves run in the GPU graph engine behind ``core()`` (the general-graph
nested dissection orders the unknowns, there being no grid).
"""

from __future__ import annotations

import os
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

from ._backend import core
from .circuitscape import (POLICIES, SCENARIOS, IncludedPairs, _flag, _open, _prefix, _resolve, _unset,
                           _write_matrix, read_included_pairs)

__all__ = ["NetworkResult", "circuitscape_network_advanced", "circuitscape_network_pairwise", "read_graph"]


# ---- reading (network.py:24-113)

def _read_table(path) -> list:
    rows = []
    with _open(path) as f:
        for line in f:
            parts = line.replace(",", " ").split()
            if parts:
                rows.append([float(t) for t in parts])
    return rows


def _as_table(value, ncol: int, name: str) -> np.ndarray:
    if isinstance(value, (str, os.PathLike)):
        rows = _read_table(value)
        if not rows or any(len(r) != len(rows[0]) for r in rows):
            raise ValueError(f"{value}: expected a table with the same number of columns on every line")
        t = np.array(rows, dtype=np.float64)
    else:
        t = np.array(value, dtype=np.float64)
    if t.size == 0:
        t = t.reshape(-1, ncol)
    if t.ndim != 2 or t.shape[1] != ncol:
        raise ValueError(f"the {name} must have {ncol} columns")
    return t


def read_graph(edges):
    """A Circuitscape edge list (path or ``(k, 3)`` array: node, node, value) as
    ``(from, to, value, shifted)``: nodes numbered from 1 (a list numbered from 0 is shifted by one,
    and ``shifted`` says so)."""
    t = _as_table(edges, 3, "graph")
    if len(t) == 0:
        raise ValueError("the graph has no edges")
    i, j = t[:, 0].astype(np.int64), t[:, 1].astype(np.int64)
    low = min(i.min(), j.min())
    if low > 1:
        raise ValueError(f"the graph numbers its nodes from {low}; Circuitscape counts from 1 (or 0)")
    if low < 0:
        raise ValueError("negative node numbers")
    if low == 0:
        i, j = i + 1, j + 1
    return i, j, t[:, 2].copy(), bool(low == 0)


def _read_focal_nodes(points) -> np.ndarray:
    # every number of the file, column after column; a list from 0 shifted to 1
    if isinstance(points, (str, os.PathLike)):
        rows = _read_table(points)
        if rows and all(len(r) == len(rows[0]) for r in rows):
            fp = np.array(rows).ravel(order="F")
        else:
            fp = np.array([v for r in rows for v in r])
    else:
        fp = np.asarray(points, dtype=np.float64).ravel()
    fp = fp.astype(np.int64)
    if len(fp) and fp.min() == 0:
        fp = fp + 1
    return fp


def _read_strengths(value, starts_from_zero: bool) -> np.ndarray:
    # (node, value) rows, shifted when they or the graph count from 0
    t = _as_table(value, 2, "node values")
    if len(t) and (t[:, 0].min() == 0 or starts_from_zero):
        t[:, 0] += 1
    return t


# ---- the result (network.py:116-157)

@dataclass
class NetworkResult:
    """Circuitscape's network outputs, nodes numbered from 1 as it writes them.

    Advanced: ``voltages`` (node, voltage), ``node_currents`` (node,
    current), ``branch_currents`` (node, node, current; above 1e-6).
    Pairwise: ``resistances`` (ids in the first row and column; -1 not
    connected), with maps ``cum_node_currents`` / ``cum_branch_currents``
    and per-pair tables in ``maps`` keyed ``(kind, id1, id2)``.
    """

    voltages: np.ndarray | None = None
    node_currents: np.ndarray | None = None
    branch_currents: np.ndarray | None = None
    resistances: np.ndarray | None = None
    cum_node_currents: np.ndarray | None = None
    cum_branch_currents: np.ndarray | None = None
    maps: dict = field(default_factory=dict)
    n_components: int = 0


def _branch_table(r, c, v) -> np.ndarray:
    # the branches Circuitscape writes: current not within 1e-6 of 0, nodes from 1
    r, c, v = np.asarray(r), np.asarray(c), np.asarray(v)
    keep = ~(np.abs(v) <= 1e-6)
    return np.column_stack([r[keep] + 1.0, c[keep] + 1.0, v[keep]])


def _edges(i, j, w, is_resistance: bool):
    w = 1 / w if is_resistance else w
    m = int(max(i.max(), j.max()))
    return (np.ascontiguousarray(i - 1, dtype=np.int32), np.ascontiguousarray(j - 1, dtype=np.int32),
            np.ascontiguousarray(w, dtype=np.float64), m)


# ---- advanced (network.py:177-262)

def circuitscape_network_advanced(graph, sources, grounds, *, graph_is_resistance: bool = True,
                                  ground_is_resistance: bool = True, remove_src_or_gnd: str = "keepall",
                                  n_threads: int = 0) -> NetworkResult:
    """Advanced mode on a network (``data_type = network``, ``scenario = advanced``).

    ``graph``: edge list (node, node, resistance or conductance), a path or
    array, nodes from 1 (or from 0: then the sources' and grounds' nodes
    shift too). Duplicate edges add up; edges whose conductances cancel are
    left out. ``sources``: (node, amps); ``grounds``: (node, resistance or
    conductance to ground), 0 resistance a direct ground. Every component
    with a source and a ground is solved.
    """
    if remove_src_or_gnd not in POLICIES:
        raise ValueError(f"remove_src_or_gnd must be one of {POLICIES}")
    i, j, w, zero = read_graph(graph)
    a, b, w, m = _edges(i, j, w, graph_is_resistance)
    src_t = _read_strengths(sources, zero)
    gnd_t = _read_strengths(grounds, zero)
    if ground_is_resistance:
        with np.errstate(divide="ignore"):
            gnd_t[:, 1] = 1 / gnd_t[:, 1]
    src, gnd = np.zeros(m), np.zeros(m)
    for table, target, name in ((src_t, src, "source"), (gnd_t, gnd, "ground")):
        ids = table[:, 0].astype(np.int64)
        if len(ids) and (ids.min() < 1 or ids.max() > m):
            raise ValueError(f"a {name} node is not in the graph (nodes 1..{m})")
        target[ids - 1] = table[:, 1]  # a node listed twice keeps its last value
    r = core().network_advanced(a, b, w, m, src, gnd, remove_src_or_gnd, int(n_threads))
    nodes = np.arange(1, m + 1, dtype=np.float64)
    return NetworkResult(voltages=np.column_stack([nodes, r["voltage"]]),
                         node_currents=np.column_stack([nodes, r["current"]]),
                         branch_currents=_branch_table(*r["branch"]), n_components=int(r["n_components"]))


# ---- pairwise (network.py:266-374)

def circuitscape_network_pairwise(graph, points, *, graph_is_resistance: bool = True, included_pairs=None,
                                  voltage_maps: bool = False, current_maps: bool = False,
                                  n_threads: int = 0) -> NetworkResult:
    """Pairwise mode on a network (``data_type = network``, ``scenario = pairwise``).

    Effective resistances between every pair of focal nodes (``points``:
    node numbers, a path or an array; from 0 they shift by one). Every
    component is factorised once and all its pairs solved (no shortcut on
    networks). ``included_pairs``: an :class:`IncludedPairs` or a path. With
    ``voltage_maps`` every pair's node voltages, and with either option its
    node and branch currents, go to ``maps``; the currents are summed over
    the pairs into ``cum_node_currents`` / ``cum_branch_currents``.
    """
    i, j, w, zero = read_graph(graph)
    a, b, w, m = _edges(i, j, w, graph_is_resistance)
    fp = _read_focal_nodes(points)
    if included_pairs is not None and not isinstance(included_pairs, IncludedPairs):
        included_pairs = read_included_pairs(included_pairs, starts_from_zero=zero)
    excluded: set = set()
    if included_pairs is not None:
        if included_pairs.mode == "include":
            fp = fp[np.isin(fp, included_pairs.ids)]
        excluded = included_pairs.excluded()
    n = len(fp)
    ex = np.zeros((n, n), dtype=np.uint8)
    if excluded:
        for p in range(n):
            for q in range(n):
                ex[p, q] = (int(fp[p]), int(fp[q])) in excluded
    maps: dict = {}

    def emit(ci, cj, nodes, volt, cur, branch):
        key = (int(fp[ci]), int(fp[cj]))
        ids = nodes + 1.0
        if volt is not None:
            maps[("voltages", *key)] = np.column_stack([ids, volt])
        maps[("node_currents", *key)] = np.column_stack([ids, cur])
        maps[("branch_currents", *key)] = _branch_table(*branch)

    opts = {"volt_maps": bool(voltage_maps), "cur_maps": bool(current_maps), "cum": bool(current_maps)}
    pts = np.ascontiguousarray(np.where((fp >= 1) & (fp <= m), fp - 1, -1), dtype=np.int32)
    r = core().network_pairwise(a, b, w, m, pts, ex, opts, emit, int(n_threads))
    res = np.zeros((n + 1, n + 1))
    res[0, 1:] = fp
    res[1:, 0] = fp
    res[1:, 1:] = r["resistances"]
    out = NetworkResult(resistances=res, maps=maps, n_components=int(r["n_components"]))
    if "cum" in r:
        out.cum_node_currents = np.column_stack([np.arange(1, m + 1, dtype=np.float64), r["cum"]])
        # one slot per input line: the first line of a node pair (either way round) takes the pair's
        # summed current, as Circuitscape's cumulative vectors are laid out
        cr, cc, cv = (np.asarray(x) for x in r["cum_branch"])
        total = {(int(p), int(q)): float(v) for p, q, v in zip(cr + 1, cc + 1, cv)}
        seen: set = set()
        vals = np.zeros(len(i))
        for k, (p, q) in enumerate(zip(i.tolist(), j.tolist())):
            key = (min(p, q), max(p, q))
            if key in seen:
                continue
            seen.add(key)
            vals[k] = total.get(key, 0.0)
        keep = ~(np.abs(vals) <= 1e-6)
        out.cum_branch_currents = np.column_stack([i[keep].astype(np.float64), j[keep].astype(np.float64),
                                                   vals[keep]])
    return out


# ---- INI files (network.py:377-446)

def _write_table(path: str, a: np.ndarray, sep: str = "\t") -> None:
    with open(path, "w", encoding="utf-8") as f:
        for row in np.atleast_2d(a):
            f.write(sep.join(repr(float(v)) for v in row) + "\n")


def network_ini(cfg: dict, base: Path, *, write_outputs: bool = True, n_threads: int = 0) -> NetworkResult:
    """``circuitscape_ini`` for ``data_type = network`` (``cfg``: the INI over Circuitscape's defaults)."""

    def need(key):
        if _unset(cfg[key]):
            raise ValueError(f"{key} is not set")
        p = _resolve(cfg[key], base)
        if not p.is_file():
            raise FileNotFoundError(f'{key} = "{cfg[key]}" does not exist')
        return p

    scenario = SCENARIOS.get(cfg["scenario"])
    if scenario is None:
        raise ValueError(f'scenario = "{cfg["scenario"]}" is not one of {sorted(set(SCENARIOS.values()))}')
    graph = need("habitat_file")
    is_res = _flag(cfg, "habitat_map_is_resistances")
    volt, cur = _flag(cfg, "write_volt_maps"), _flag(cfg, "write_cur_maps")
    cum_only = _flag(cfg, "write_cum_cur_map_only")
    if scenario == "pairwise":
        pairs = None
        if _flag(cfg, "use_included_pairs"):
            pairs = read_included_pairs(need("included_pairs_file"), starts_from_zero=read_graph(graph)[3])
        # a pair's currents are written with any map option; the cumulative tables need write_cur_maps
        any_maps = volt or cur or cum_only or _flag(cfg, "write_max_cur_maps")
        result = circuitscape_network_pairwise(graph, need("point_file"), graph_is_resistance=is_res,
                                               included_pairs=pairs, voltage_maps=volt, current_maps=any_maps,
                                               n_threads=n_threads)
        if not cur:
            result.cum_node_currents = result.cum_branch_currents = None
    else:
        # networks have no one-to-all / all-to-one: every other scenario runs as advanced
        policy = cfg["remove_src_or_gnd"]
        if policy == "not entered":
            policy = "keepall"
        result = circuitscape_network_advanced(graph, need("source_file"), need("ground_file"),
                                               graph_is_resistance=is_res,
                                               ground_is_resistance=_flag(cfg, "ground_file_is_resistances"),
                                               remove_src_or_gnd=policy, n_threads=n_threads)
    if not write_outputs:
        return result
    prefix, _ = _prefix(cfg)
    with open(cfg["output_file"], "w", encoding="utf-8") as f:
        for k, v in cfg.items():
            f.write(f"{k} = {v}\n")
    if scenario == "pairwise":
        r = result.resistances
        _write_matrix(prefix + "_resistances.out", r)
        ids = r[1:, 0]
        three = [(ids[p], ids[q], r[q + 1, p + 1]) for p in range(len(ids)) for q in range(p + 1, len(ids))]
        _write_matrix(prefix + "_resistances_3columns.out", np.array(three).reshape(-1, 3))
        for (kind, p, q), table in result.maps.items():
            _write_table(f"{prefix}_{kind}_{p}_{q}.txt", table)
        if result.cum_node_currents is not None:
            _write_table(f"{prefix}_node_currents_cum.txt", result.cum_node_currents)
            _write_table(f"{prefix}_branch_currents_cum.txt", result.cum_branch_currents)
    else:
        if volt:
            _write_table(f"{prefix}_voltages.txt", result.voltages)
        if cur or cum_only:
            _write_table(f"{prefix}_node_currents.txt", result.node_currents)
            _write_table(f"{prefix}_branch_currents.txt", result.branch_currents)
    return result
