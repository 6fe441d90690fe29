"""Circuitscape's advanced mode on a raster, and the raster inputs of the circuit API.

Mirrors the reference's src/landschaft/circuit/advanced.py.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from ._backend import core


def _grid(values, name: str, *, band=None):
    """A raster (array or DataArray) as contiguous float64 with NaN, and the DataArray it came from.

    The reference also reads paths through rasterio (advanced.py:34-54); this
    path takes arrays (and DataArrays when xarray is importable).
    """
    like = None
    if hasattr(values, "dims") and hasattr(values, "values"):
        da = values.squeeze(drop=True) if "band" in values.dims else values
        v = np.array(da.values, dtype=np.float64)
        nodata = da.rio.nodata if hasattr(da, "rio") else None
        if nodata is not None and not np.isnan(nodata):
            v[v == nodata] = np.nan
        like = da
    elif isinstance(values, (str, bytes)) or hasattr(values, "__fspath__"):
        raise ValueError(f"the {name} must be an array (raster files are read by the caller)")
    else:
        v = np.array(np.ma.filled(np.ma.asarray(values, dtype=np.float64), np.nan), dtype=np.float64)
    if v.ndim != 2:
        raise ValueError(f"the {name} must be a 2-D raster")
    return np.ascontiguousarray(v), like


@dataclass
class AdvancedResult:
    """Node currents and voltages of :func:`advanced` (the reference's advanced.py:15-31).

    DataArrays when the conductance was one, arrays otherwise: ``current``
    (max of in- and outflow per cell, 0 where nothing was solved) and
    ``voltage`` (0 in components without a source and a ground, NaN on cells
    without a node).
    """

    current: object
    voltage: object


def _like(values: np.ndarray, like, name: str):
    if like is None:
        return values
    out = like.copy(data=values)
    out.name = name
    return out


def advanced(conductance, source, ground, *, four_neighbors: bool = False, n_threads: int = 0) -> AdvancedResult:
    """Currents and voltages of a raster circuit with sources and grounds (advanced.py:69-137).

    Every cell with a conductance > 0 is a node joined to its 8 (or 4)
    neighbours by the mean conductance of the two cells (over sqrt 2 on the
    diagonals); ``source`` injects amps, ``ground`` is a conductance to ground
    (``inf``: a direct ground). A cell that is both loses its source. Every
    component with a source and a ground is solved on the GPU.
    """
    g, like = _grid(conductance, "conductance")
    s, _ = _grid(source, "source")
    gr, _ = _grid(ground, "ground")
    if s.shape != g.shape or gr.shape != g.shape:
        raise ValueError("conductance, source and ground must have the same shape")
    if np.any(g == np.inf):
        raise ValueError("conductances must be finite")
    if np.any(gr[np.isfinite(gr)] < 0):
        raise ValueError("ground conductances cannot be negative")
    r = core().circuit_advanced(np.nan_to_num(g, nan=0.0, posinf=np.inf), np.where(np.isnan(s), 0.0, s),
                                np.where(np.isnan(gr), 0.0, gr), bool(four_neighbors), int(n_threads))
    return AdvancedResult(_like(r["current"], like, "current"), _like(r["voltage"], like, "voltage"))
