"""Circuitscape's network (graph) mode, advanced and pairwise, on the GPU graph engine.

The public names, arguments, results and error messages are those of the reference's
src/landschaft/circuit/network.py (readers :24-113, NetworkResult :116-157, advanced :177-262,
pairwise :266-374, the INI driver :377-446), so callers can switch modules. The engine behind
``core()`` orders a graph without a grid by a general-graph nested dissection.

Inputs are normalised once into a :class:`_Network` (0-based int32 endpoints, conductances, the node
count, whether the caller counted from 0); everything downstream works on it.
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

# branch currents within this of 0 are not written (Circuitscape's tables)
_BRANCH_FLOOR = 1e-6


# ---- text tables

def _rows_of(path) -> list:
    """Every non-empty line of a (possibly gzipped) text file as a list of floats; commas count as
    blanks."""
    with _open(path) as f:
        tokens = (line.replace(",", " ").split() for line in f)
        return [[float(t) for t in toks] for toks in tokens if toks]


def _matrix(value, ncol: int, name: str) -> np.ndarray:
    """A path or array as a float64 table of `ncol` columns (an empty one as 0 x ncol)."""
    if isinstance(value, (str, os.PathLike)):
        rows = _rows_of(value)
        widths = {len(r) for r in rows}
        if len(widths) != 1:
            raise ValueError(f"{value}: expected a table with the same number of columns on every line")
        table = np.asarray(rows, dtype=np.float64)
    else:
        table = np.asarray(value, dtype=np.float64).copy()
    if table.size == 0:
        table = np.zeros((0, ncol))
    if table.ndim != 2 or table.shape[1] != ncol:
        raise ValueError(f"the {name} must have {ncol} columns")
    return table


def read_graph(edges):
    """A Circuitscape edge list (path or ``(k, 3)`` array of node, node, value) as
    ``(from, to, value, shifted)``: node numbers from 1 (a list counting from 0 is moved up by one,
    and ``shifted`` says so) and the edge values."""
    table = _matrix(edges, 3, "graph")
    if table.shape[0] == 0:
        raise ValueError("the graph has no edges")
    ends = table[:, :2].astype(np.int64)
    base = int(ends.min())
    if base < 0:
        raise ValueError("negative node numbers")
    if base > 1:
        raise ValueError(f"the graph numbers its nodes from {base}; Circuitscape counts from 1 (or 0)")
    ends += 1 - base
    return ends[:, 0].copy(), ends[:, 1].copy(), table[:, 2].copy(), base == 0


def _focal_nodes(points) -> np.ndarray:
    """Focal node numbers from a file (all its numbers, read down the columns when every line has
    as many, else line by line) or an array; a list holding 0 counts from 0 and moves up by one."""
    if isinstance(points, (str, os.PathLike)):
        rows = _rows_of(points)
        if rows and len({len(r) for r in rows}) == 1:
            nums = np.asarray(rows).T.ravel()
        else:
            nums = np.asarray([v for r in rows for v in r])
    else:
        nums = np.asarray(points, dtype=np.float64).ravel()
    nodes = nums.astype(np.int64)
    return nodes + 1 if nodes.size and nodes.min() == 0 else nodes


def _node_table(value, graph_from_zero: bool) -> np.ndarray:
    """(node, value) rows; moved up by one when they count from 0 or the graph does."""
    table = _matrix(value, 2, "node values")
    if table.shape[0] and (graph_from_zero or table[:, 0].min() == 0):
        table[:, 0] += 1
    return table


# ---- the graph handed to the library

@dataclass
class _Network:
    frm: np.ndarray  # edge endpoints as given, from 1
    to: np.ndarray
    a: np.ndarray  # the same from 0, int32
    b: np.ndarray
    w: np.ndarray  # conductances
    m: int  # nodes 1..m
    from_zero: bool

    @classmethod
    def of(cls, graph, is_resistance: bool) -> "_Network":
        frm, to, value, from_zero = read_graph(graph)
        w = np.ascontiguousarray(1.0 / value if is_resistance else value, dtype=np.float64)
        return cls(frm, to, np.ascontiguousarray(frm - 1, dtype=np.int32), np.ascontiguousarray(to - 1, dtype=np.int32),
                   w, int(max(frm.max(), to.max())), from_zero)

    def per_node(self, table: np.ndarray, what: str) -> np.ndarray:
        """Node values as a dense array over the nodes (a node listed twice: its last value)."""
        out = np.zeros(self.m)
        nodes = table[:, 0].astype(np.int64)
        bad = (nodes < 1) | (nodes > self.m)
        if bad.any():
            raise ValueError(f"a {what} node is not in the graph (nodes 1..{self.m})")
        out[nodes - 1] = table[:, 1]
        return out

    def numbered(self, values: np.ndarray) -> np.ndarray:
        """(node, value) rows over every node."""
        return np.column_stack([np.arange(1.0, self.m + 1.0), values])


# ---- results

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


def _written_branches(r, c, v) -> np.ndarray:
    """Branches whose current is more than 1e-6 from 0 (NaN kept), nodes from 1."""
    r, c, v = (np.asarray(x) for x in (r, c, v))
    shown = ~(np.abs(v) <= _BRANCH_FLOOR)
    return np.stack([r[shown] + 1.0, c[shown] + 1.0, v[shown]], axis=1) if shown.any() else np.zeros((0, 3))


# ---- advanced

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
    net = _Network.of(graph, graph_is_resistance)
    src = net.per_node(_node_table(sources, net.from_zero), "source")
    gnd_t = _node_table(grounds, net.from_zero)
    if ground_is_resistance:
        with np.errstate(divide="ignore"):
            gnd_t[:, 1] = np.reciprocal(gnd_t[:, 1])  # a 0 resistance: an infinite conductance (direct)
    gnd = net.per_node(gnd_t, "ground")
    out = core().network_advanced(net.a, net.b, net.w, net.m, src, gnd, remove_src_or_gnd, int(n_threads))
    return NetworkResult(voltages=net.numbered(out["voltage"]), node_currents=net.numbered(out["current"]),
                         branch_currents=_written_branches(*out["branch"]), n_components=int(out["n_components"]))


# ---- pairwise

def _excluded_matrix(nodes: np.ndarray, excluded: set) -> np.ndarray:
    """uint8 n x n: 1 where the ordered pair of focal nodes is excluded."""
    n = nodes.size
    ex = np.zeros((n, n), dtype=np.uint8)
    if excluded:
        pos = {}
        for k, v in enumerate(nodes.tolist()):
            pos.setdefault(v, []).append(k)
        for p, q in excluded:
            for i in pos.get(p, ()):
                ex[i, pos.get(q, [])] = 1
    return ex


def _cumulative_branches(net: _Network, cum_branch) -> np.ndarray:
    """The pairs' summed branch currents laid out on the input lines, as Circuitscape writes its
    cumulative vector: every unordered node pair's current on its first input line, 0 on repeats;
    lines more than 1e-6 from 0 kept."""
    cr, cc, cv = (np.asarray(x) for x in cum_branch)
    lo = np.minimum(net.frm, net.to)
    hi = np.maximum(net.frm, net.to)
    key = lo * (net.m + 1) + hi
    sums = dict(zip(((np.minimum(cr, cc) + 1) * (net.m + 1) + np.maximum(cr, cc) + 1).tolist(), cv.tolist()))
    _, first = np.unique(key, return_index=True)
    line_value = np.zeros(key.size)
    line_value[first] = [sums.get(k, 0.0) for k in key[first].tolist()]
    shown = ~(np.abs(line_value) <= _BRANCH_FLOOR)
    return np.stack([net.frm[shown].astype(np.float64), net.to[shown].astype(np.float64), line_value[shown]], axis=1)


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
    net = _Network.of(graph, graph_is_resistance)
    focal = _focal_nodes(points)
    pairs = included_pairs
    if pairs is not None and not isinstance(pairs, IncludedPairs):
        pairs = read_included_pairs(pairs, starts_from_zero=net.from_zero)
    excluded = set()
    if pairs is not None:
        if pairs.mode == "include":
            focal = focal[np.isin(focal, pairs.ids)]
        excluded = pairs.excluded()
    ex = _excluded_matrix(focal, excluded)
    maps: dict = {}

    def record(ci, cj, nodes, volt, cur, branch):
        pair = (int(focal[ci]), int(focal[cj]))
        numbered = nodes + 1.0
        if volt is not None:
            maps[("voltages", *pair)] = np.stack([numbered, volt], axis=1)
        maps[("node_currents", *pair)] = np.stack([numbered, cur], axis=1)
        maps[("branch_currents", *pair)] = _written_branches(*branch)

    options = {"volt_maps": bool(voltage_maps), "cur_maps": bool(current_maps), "cum": bool(current_maps)}
    on_graph = (focal >= 1) & (focal <= net.m)
    library_points = np.ascontiguousarray(np.where(on_graph, focal - 1, -1), dtype=np.int32)
    out = core().network_pairwise(net.a, net.b, net.w, net.m, library_points, ex, options, record, int(n_threads))
    table = np.zeros((focal.size + 1, focal.size + 1))
    table[0, 1:] = table[1:, 0] = focal
    table[1:, 1:] = out["resistances"]
    result = NetworkResult(resistances=table, maps=maps, n_components=int(out["n_components"]))
    if "cum" in out:
        result.cum_node_currents = net.numbered(out["cum"])
        result.cum_branch_currents = _cumulative_branches(net, out["cum_branch"])
    return result


# ---- network INI files

def _write_table(path: str, a: np.ndarray, sep: str = "\t") -> None:
    lines = (sep.join(repr(float(v)) for v in row) for row in np.atleast_2d(a))
    Path(path).write_text("".join(line + "\n" for line in lines), encoding="utf-8")


def _ini_file(cfg: dict, base: Path, key: str) -> Path:
    """The file an INI key names (relative paths: the working directory, then the INI's)."""
    if _unset(cfg[key]):
        raise ValueError(f"{key} is not set")
    path = _resolve(cfg[key], base)
    if not path.is_file():
        raise FileNotFoundError(f'{key} = "{cfg[key]}" does not exist')
    return path


def _write_network_outputs(cfg: dict, scenario: str, result: NetworkResult) -> None:
    prefix, _ = _prefix(cfg)
    Path(cfg["output_file"]).write_text("".join(f"{k} = {v}\n" for k, v in cfg.items()), encoding="utf-8")
    if scenario != "pairwise":
        if _flag(cfg, "write_volt_maps"):
            _write_table(f"{prefix}_voltages.txt", result.voltages)
        if _flag(cfg, "write_cur_maps") or _flag(cfg, "write_cum_cur_map_only"):
            _write_table(f"{prefix}_node_currents.txt", result.node_currents)
            _write_table(f"{prefix}_branch_currents.txt", result.branch_currents)
        return
    res = result.resistances
    _write_matrix(prefix + "_resistances.out", res)
    ids = res[1:, 0]
    iu, ju = np.triu_indices(ids.size, 1)
    _write_matrix(prefix + "_resistances_3columns.out",
                  np.stack([ids[iu], ids[ju], res[ju + 1, iu + 1]], axis=1).reshape(-1, 3))
    for (kind, p, q), table in result.maps.items():
        _write_table(f"{prefix}_{kind}_{p}_{q}.txt", table)
    if result.cum_node_currents is not None:
        _write_table(f"{prefix}_node_currents_cum.txt", result.cum_node_currents)
        _write_table(f"{prefix}_branch_currents_cum.txt", result.cum_branch_currents)


def network_ini(cfg: dict, base: Path, *, write_outputs: bool = True, n_threads: int = 0) -> NetworkResult:
    """``circuitscape_ini`` for ``data_type = network`` (``cfg``: the INI over Circuitscape's defaults)."""
    scenario = SCENARIOS.get(cfg["scenario"])
    if scenario is None:
        raise ValueError(f'scenario = "{cfg["scenario"]}" is not one of {sorted(set(SCENARIOS.values()))}')
    graph = _ini_file(cfg, base, "habitat_file")
    is_res = _flag(cfg, "habitat_map_is_resistances")
    if scenario == "pairwise":
        pairs = None
        if _flag(cfg, "use_included_pairs"):
            pairs = read_included_pairs(_ini_file(cfg, base, "included_pairs_file"),
                                        starts_from_zero=read_graph(graph)[3])
        # any map option writes each pair's currents; the cumulative tables come with write_cur_maps
        any_map = any(_flag(cfg, k) for k in ("write_volt_maps", "write_cur_maps", "write_cum_cur_map_only",
                                               "write_max_cur_maps"))
        result = circuitscape_network_pairwise(graph, _ini_file(cfg, base, "point_file"), graph_is_resistance=is_res,
                                               included_pairs=pairs, voltage_maps=_flag(cfg, "write_volt_maps"),
                                               current_maps=any_map, n_threads=n_threads)
        if not _flag(cfg, "write_cur_maps"):
            result.cum_node_currents = result.cum_branch_currents = None
    else:
        # a network has no one-to-all or all-to-one: every other scenario is solved as advanced
        policy = "keepall" if cfg["remove_src_or_gnd"] == "not entered" else cfg["remove_src_or_gnd"]
        result = circuitscape_network_advanced(graph, _ini_file(cfg, base, "source_file"),
                                               _ini_file(cfg, base, "ground_file"), graph_is_resistance=is_res,
                                               ground_is_resistance=_flag(cfg, "ground_file_is_resistances"),
                                               remove_src_or_gnd=policy, n_threads=n_threads)
    if write_outputs:
        _write_network_outputs(cfg, scenario, result)
    return result
