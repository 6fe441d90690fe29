"""Omniscape: omnidirectional circuit connectivity over moving windows, on the GPU window engine.

The public function, its arguments, outputs (``cum_currmap``, ``flow_potential``,
``normalized_cum_currmap`` and ``attrs``) and error messages are those of the reference's
src/landschaft/circuit/omniscape.py:102-404, so callers can switch modules. The call is prepared in
three steps here: the rasters (precision, reclassification, sources), the conditions, then the maps
from ``core().circuit_omniscape``. Results are an ``xarray.Dataset`` when xarray is importable, else a
light :class:`Dataset` with the same variable names and ``attrs``.
"""

from __future__ import annotations

import warnings

import numpy as np

from ._backend import core
from .advanced import _grid

COMPARISONS = ("within", "equal")
COMPARE_TO_FUTURE = ("none", "1", "2", "both")
PRECISIONS = ("double", "single")
_OUTPUT_ORDER = ("cum_currmap", "flow_potential", "normalized_cum_currmap")


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
    """An integer option as Omniscape reads one (omniscape.py:43-48): the nearest integer, halves to
    even."""
    v = float(value)
    if v != v or v in (np.inf, -np.inf):
        raise ValueError(f"{name} must be finite")
    return round(v)


def read_reclass_table(path) -> np.ndarray:
    """Omniscape's reclass table: (value, new value) rows, ``missing`` = NaN (omniscape.py:51-82)."""

    def number(tok: str) -> float:
        return np.nan if tok.strip('"') == "missing" else float(tok)

    pairs = []
    with open(path, encoding="utf-8") as f:
        for raw in f:
            cols = raw.split()
            if cols and len(cols) != 2:
                raise ValueError(f"reclass table lines need two columns: {raw.strip()!r}")
            if cols:
                pairs.append((number(cols[0]), number(cols[1])))
    return np.asarray(pairs, dtype=np.float64).reshape(-1, 2)


def _reclass_pairs(table) -> np.ndarray:
    """A reclass table given as a path, a {value: new} dict (None = NoData) or (value, new) rows."""
    if isinstance(table, (str, bytes)) or hasattr(table, "__fspath__"):
        return read_reclass_table(table)
    if isinstance(table, dict):
        table = [(k, np.nan if v is None else v) for k, v in table.items()]
    pairs = np.asarray(table, dtype=np.float64)
    if pairs.ndim != 2 or pairs.shape[1] != 2:
        raise ValueError("reclass_table must have two columns (value, new value)")
    return pairs


def reclassify(resistance: np.ndarray, table) -> np.ndarray:
    """reclassify_resistance!: each (value, new value) row applied in turn to the cells whose
    original value it names (omniscape.py:102-133), so chains do not cascade."""
    original = np.asarray(resistance)
    out = original.copy()
    for old, new in _reclass_pairs(table):
        np.copyto(out, new, where=original == old)
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
    if resistance_is_conductance:
        strength = r
    else:
        with np.errstate(divide="ignore"):
            strength = np.reciprocal(r)
    low = 1.0 / float(r_cutoff)
    return np.where(np.isnan(strength) | (strength < low), 0.0, strength)


def _rounded(a: np.ndarray, single: bool) -> np.ndarray:
    """float64, through float32 first for the single-precision option (which rounds the inputs only)."""
    return a.astype(np.float32).astype(np.float64) if single else a


def _require_positive(r: np.ndarray, message: str) -> None:
    data = r[r == r]
    if data.size and data.min() <= 0:
        raise ValueError(message)


def _window_options(radius, buffer, block_size):
    radius, buffer, block_size = (_to_int(v, k) for v, k in ((radius, "radius"), (buffer, "buffer"),
                                                              (block_size, "block_size")))
    if min(radius, buffer) < 0 or block_size < 1:
        raise ValueError("radius and buffer must be >= 0 and block_size >= 1")
    if not block_size & 1:
        warnings.warn("block_size is even, but must be odd. Using block_size + 1.", stacklevel=3)
        block_size += 1
    return radius, buffer, block_size


def _source_map(r, source, single, source_threshold, r_cutoff, resistance_is_conductance, reclass_table):
    if source is None:
        s = source_from_resistance(r, r_cutoff=r_cutoff, resistance_is_conductance=resistance_is_conductance,
                                   reclass_table=reclass_table)
    else:
        s, _ = _grid(source, "source")
        if s.shape != r.shape:
            raise ValueError("the source raster must have the resistance's shape")
        s = _rounded(s, single)
        s = np.where((s != s) | (s < source_threshold), 0.0, s)
    return np.ascontiguousarray(s, dtype=np.float64)


def _condition_layers(shape, single, c1, c2, f1, f2, compare_to_future, comparison1, comparison2, bounds):
    """The binding's conditions dict (empty without condition1), validated."""
    if c1 is None:
        if c2 is not None:
            raise ValueError("condition2 needs condition1")
        if f1 is not None or f2 is not None:
            raise ValueError("future conditions need condition1")
        return {}
    if compare_to_future not in COMPARE_TO_FUTURE:
        raise ValueError(f"compare_to_future must be one of {COMPARE_TO_FUTURE}")
    for k, cmp in enumerate((comparison1, comparison2), start=1):
        if cmp not in COMPARISONS:
            raise ValueError(f"comparison{k} must be one of {COMPARISONS}")
    use1 = compare_to_future in ("1", "both")
    use2 = compare_to_future in ("2", "both")
    if use1 and f1 is None:
        raise ValueError("compare_to_future needs condition1_future")
    if use2 and c2 is not None and f2 is None:
        raise ValueError("compare_to_future needs condition2_future")

    def layer(value, name):
        if value is None:
            return None
        v, _ = _grid(value, name)
        if v.shape != shape:
            raise ValueError(f"{name} must have the resistance's shape")
        return np.ascontiguousarray(_rounded(v, single))

    out = {"condition1": layer(c1, "condition1"), "condition1_future": layer(f1 if use1 else None, "condition1_future"),
           "condition2": layer(c2, "condition2"), "condition2_future": layer(f2 if use2 else None, "condition2_future"),
           "comparison1": comparison1, "comparison2": comparison2}
    out.update({k: float(v) for k, v in bounds.items()})
    return out


def _maps(out, nodata, flow_potential, normalized_current, mask_nodata, single) -> dict:
    maps = {"cum_currmap": out["current"]}
    if flow_potential or normalized_current:
        fp = out["flow_potential"]
        if flow_potential:
            maps["flow_potential"] = fp
        if normalized_current:
            with np.errstate(divide="ignore", invalid="ignore"):
                ratio = out["current"] / fp
            maps["normalized_cum_currmap"] = np.nan_to_num(ratio, nan=0.0, posinf=np.inf, neginf=-np.inf)
    for m in maps.values():
        if mask_nodata:
            m[nodata] = np.nan
    return {k: (maps[k].astype(np.float32) if single else maps[k]) for k in _OUTPUT_ORDER if k in maps}


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
    single = precision == "single"
    r = _rounded(_grid(resistance, "resistance")[0], single)
    _require_positive(r, "resistance (or conductance) values must be > 0")
    radius, buffer, block_size = _window_options(radius, buffer, block_size)
    s = _source_map(r, source, single, source_threshold, r_cutoff, resistance_is_conductance, reclass_table)
    if reclass_table is not None:
        r = _rounded(reclassify(r, reclass_table), single)
        _require_positive(r, "the reclassified resistance must be > 0")
    r = np.ascontiguousarray(r)
    conditions = _condition_layers(r.shape, single, condition1, condition2, condition1_future, condition2_future,
                                   compare_to_future, comparison1, comparison2,
                                   {"condition1_lower": condition1_lower, "condition1_upper": condition1_upper,
                                    "condition2_lower": condition2_lower, "condition2_upper": condition2_upper})
    out = core().circuit_omniscape(r, s, radius, buffer, block_size, bool(four_neighbors),
                                   bool(resistance_is_conductance), bool(correct_artifacts),
                                   bool(flow_potential or normalized_current), int(n_threads), conditions,
                                   float(memory_limit or 0.0))
    data = _maps(out, np.isnan(r), flow_potential, normalized_current, mask_nodata, single)
    attrs = {"radius": radius, "block_size": block_size, "buffer": buffer, "n_targets": int(out["n_targets"]),
             "four_neighbors": int(bool(four_neighbors)), "correct_artifacts": int(bool(correct_artifacts)),
             "precision": precision, "conditional": int(condition1 is not None), "threads": int(out["threads"])}
    try:
        import xarray as xr
    except ImportError:
        return Dataset(data, attrs)
    return xr.Dataset({k: (("y", "x"), v) for k, v in data.items()}, attrs=attrs)
