"""Omniscape: omnidirectional circuit connectivity over moving windows.

A numpy restatement of the reference's Python front end
(src/landschaft/circuit/omniscape.py:102-404) over the GPU backend. Results
are an ``xarray.Dataset`` when xarray is importable, else a light
:class:`Dataset` with the same variable names and ``attrs``.
"""

from __future__ import annotations

import warnings

import numpy as np

from ._backend import core
from .advanced import _grid

COMPARISONS = ("within", "equal")
COMPARE_TO_FUTURE = ("none", "1", "2", "both")
PRECISIONS = ("double", "single")


class Variable(np.ndarray):
    """A map of a :class:`Dataset` (an ndarray with ``.values``)."""

    @property
    def values(self) -> np.ndarray:
        return np.asarray(self)


class Dataset(dict):
    """Named maps plus ``attrs``; ``ds["cum_currmap"]`` and ``ds.cum_currmap`` both work."""

    def __init__(self, data: dict, attrs: dict):
        super().__init__({k: np.asarray(v).view(Variable) for k, v in data.items()})
        self.attrs = attrs

    def __getattr__(self, name):
        try:
            return self[name]
        except KeyError as e:
            raise AttributeError(name) from e

    @property
    def data_vars(self):
        return list(self.keys())


def _to_int(value, name: str) -> int:
    # omniscape.py:43-48: Int64(round(parse(Float64, value))), halves to even.
    v = float(value)
    if not np.isfinite(v):
        raise ValueError(f"{name} must be finite")
    return int(round(v))


def read_reclass_table(path) -> np.ndarray:
    """Omniscape's reclass table: (value, new value) rows, ``missing`` = NaN (omniscape.py:51-82)."""
    rows = []
    with open(path, encoding="utf-8") as f:
        for line in f:
            parts = line.split()
            if not parts:
                continue
            if len(parts) != 2:
                raise ValueError(f"reclass table lines need two columns: {line.strip()!r}")
            rows.append([np.nan if p.strip('"') == "missing" else float(p) for p in parts])
    return np.array(rows, dtype=np.float64).reshape(-1, 2)


def _reclass_table(table) -> np.ndarray:
    if isinstance(table, (str, bytes)) or hasattr(table, "__fspath__"):
        return read_reclass_table(table)
    if isinstance(table, dict):
        table = [[k, np.nan if v is None else v] for k, v in table.items()]
    t = np.array(table, dtype=np.float64)
    if t.ndim != 2 or t.shape[1] != 2:
        raise ValueError("reclass_table must have two columns (value, new value)")
    return t


def reclassify(resistance: np.ndarray, table) -> np.ndarray:
    """reclassify_resistance!: rows applied in turn, matched on the original values (omniscape.py:102-133)."""
    t = _reclass_table(table)
    out = resistance.copy()
    for old, new in t:
        out[resistance == old] = new
    return out


def source_from_resistance(resistance, *, r_cutoff: float = np.inf, resistance_is_conductance: bool = False,
                           reclass_table=None):
    """Sources from the resistance: 1 / r (or r), 0 below 1 / r_cutoff and on NoData (omniscape.py:136-178).

    >>> import numpy as np
    >>> source_from_resistance(np.array([[1.0, 2.0], [4.0, np.nan]]))
    array([[1.  , 0.5 ],
           [0.25, 0.  ]])
    """
    r, _ = _grid(resistance, "resistance")
    if reclass_table is not None:
        r = reclassify(r, reclass_table)
    with np.errstate(divide="ignore"):
        s = r if resistance_is_conductance else 1 / r
    return np.where(np.isnan(s) | (s < 1 / float(r_cutoff)), 0.0, s)


def _condition(value, shape, name, dtype):
    if value is None:
        return None
    v, _ = _grid(value, name)
    if v.shape != shape:
        raise ValueError(f"{name} must have the resistance's shape")
    if dtype is np.float32:
        v = v.astype(np.float32).astype(np.float64)
    return np.ascontiguousarray(v)


def omniscape(resistance, radius, *, source=None, block_size=1, buffer=0, source_threshold: float = 0.0,
              r_cutoff: float = np.inf, resistance_is_conductance: bool = False, four_neighbors: bool = False,
              flow_potential: bool = False, normalized_current: bool = False, correct_artifacts: bool = True,
              mask_nodata: bool = True, reclass_table=None, condition1=None, condition2=None,
              condition1_future=None, condition2_future=None, compare_to_future: str = "none",
              comparison1: str = "within", comparison2: str = "within", condition1_lower: float = 0.0,
              condition1_upper: float = 0.0, condition2_lower: float = 0.0, condition2_upper: float = 0.0,
              precision: str = "double", memory_limit: float | None = None, n_threads: int = 0):
    """Omnidirectional connectivity with Omniscape's moving-window circuits.

    The arguments and outputs are those of the reference's ``omniscape``
    (src/landschaft/circuit/omniscape.py:192-404): every block of
    ``block_size`` cells with a positive source sum is a target, its window of
    ``radius + buffer`` is solved with the target grounded, and the windows'
    currents add up to ``cum_currmap`` (plus ``flow_potential`` and
    ``normalized_cum_currmap`` when asked for). With ``condition1`` (and
    ``condition2``), a window's sources whose condition does not match the
    target block's (future) value are dropped after scaling: ``within`` keeps
    lower <= condition - median(block) <= upper, ``equal`` keeps condition ==
    mode(block). Windows are solved on the GPU in batches.
    """
    if precision not in PRECISIONS:
        raise ValueError(f"precision must be one of {PRECISIONS}")
    dtype = np.float32 if precision == "single" else np.float64
    r, _ = _grid(resistance, "resistance")
    if dtype is np.float32:
        r = r.astype(np.float32).astype(np.float64)
    finite = r[~np.isnan(r)]
    if finite.size and finite.min() <= 0:
        raise ValueError("resistance (or conductance) values must be > 0")
    radius = _to_int(radius, "radius")
    buffer = _to_int(buffer, "buffer")
    block_size = _to_int(block_size, "block_size")
    if radius < 0 or buffer < 0 or block_size < 1:
        raise ValueError("radius and buffer must be >= 0 and block_size >= 1")
    if block_size % 2 == 0:
        warnings.warn("block_size is even, but must be odd. Using block_size + 1.", stacklevel=2)
        block_size += 1
    if source is None:
        s = source_from_resistance(r, r_cutoff=r_cutoff, resistance_is_conductance=resistance_is_conductance,
                                   reclass_table=reclass_table)
    else:
        s, _ = _grid(source, "source")
        if s.shape != r.shape:
            raise ValueError("the source raster must have the resistance's shape")
        if dtype is np.float32:
            s = s.astype(np.float32).astype(np.float64)
        s = np.where(np.isnan(s), 0.0, s)
        s = np.where(s < source_threshold, 0.0, s)
    s = np.ascontiguousarray(s, dtype=np.float64)
    if reclass_table is not None:
        r = reclassify(r, reclass_table)
        if dtype is np.float32:
            r = r.astype(np.float32).astype(np.float64)
        finite = r[~np.isnan(r)]
        if finite.size and finite.min() <= 0:
            raise ValueError("the reclassified resistance must be > 0")
    r = np.ascontiguousarray(r)
    # conditional connectivity (omniscape.py:344-368 of the reference)
    conditions = {}
    if condition2 is not None and condition1 is None:
        raise ValueError("condition2 needs condition1")
    if condition1 is not None:
        if compare_to_future not in COMPARE_TO_FUTURE:
            raise ValueError(f"compare_to_future must be one of {COMPARE_TO_FUTURE}")
        for k, cmp in ((1, comparison1), (2, comparison2)):
            if cmp not in COMPARISONS:
                raise ValueError(f"comparison{k} must be one of {COMPARISONS}")
        future1 = condition1_future if compare_to_future in ("1", "both") else None
        future2 = condition2_future if compare_to_future in ("2", "both") else None
        if compare_to_future in ("1", "both") and condition1_future is None:
            raise ValueError("compare_to_future needs condition1_future")
        if compare_to_future in ("2", "both") and condition2 is not None and condition2_future is None:
            raise ValueError("compare_to_future needs condition2_future")
        conditions = {
            "condition1": _condition(condition1, r.shape, "condition1", dtype),
            "condition1_future": _condition(future1, r.shape, "condition1_future", dtype),
            "condition2": _condition(condition2, r.shape, "condition2", dtype),
            "condition2_future": _condition(future2, r.shape, "condition2_future", dtype),
            "comparison1": comparison1, "comparison2": comparison2,
            "condition1_lower": float(condition1_lower), "condition1_upper": float(condition1_upper),
            "condition2_lower": float(condition2_lower), "condition2_upper": float(condition2_upper),
        }
    elif any(c is not None for c in (condition1_future, condition2_future)):
        raise ValueError("future conditions need condition1")
    compute_fp = bool(flow_potential or normalized_current)
    out = core().circuit_omniscape(r, s, radius, buffer, block_size, bool(four_neighbors),
                                   bool(resistance_is_conductance), bool(correct_artifacts), compute_fp,
                                   int(n_threads), conditions, 0.0 if memory_limit is None else float(memory_limit))
    cum = out["current"]
    data = {"cum_currmap": cum}
    nodata = np.isnan(r)
    if compute_fp:
        fp = out["flow_potential"]
        if normalized_current:
            with np.errstate(divide="ignore", invalid="ignore"):
                norm = cum / fp
            norm[np.isnan(norm)] = 0.0
            data["normalized_cum_currmap"] = norm
        if flow_potential:
            data["flow_potential"] = fp
    if mask_nodata:
        for v in data.values():
            v[nodata] = np.nan
    if dtype is np.float32:
        data = {k: v.astype(np.float32) for k, v in data.items()}
    attrs = {"radius": radius, "block_size": block_size, "buffer": buffer, "n_targets": int(out["n_targets"]),
             "four_neighbors": int(bool(four_neighbors)), "correct_artifacts": int(bool(correct_artifacts)),
             "precision": precision, "conditional": int(condition1 is not None), "threads": int(out["threads"])}
    order = ["cum_currmap", "flow_potential", "normalized_cum_currmap"]
    data = {k: data[k] for k in order if k in data}
    try:
        import xarray as xr
    except ImportError:
        return Dataset(data, attrs)
    return xr.Dataset({k: (("y", "x"), v) for k, v in data.items()}, attrs=attrs)
