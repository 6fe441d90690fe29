"""Circuitscape's advanced mode on a raster, and the raster inputs of the circuit API.

Mirrors the reference's src/landschaft/circuit/advanced.py.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from ._backend import core


def _is_path(value) -> bool:
    return isinstance(value, (str, bytes)) or hasattr(value, "__fspath__")


def _dataarray_values(da) -> np.ndarray:
    """A DataArray's cells as float64 (a band dimension squeezed away), its NoData value as NaN."""
    if "band" in da.dims:
        da = da.squeeze(drop=True)
    cells = np.asarray(da.values, dtype=np.float64).copy()
    marker = getattr(getattr(da, "rio", None), "nodata", None)
    if marker is not None and marker == marker:
        cells[cells == marker] = np.nan
    return cells, da


def _grid(values, name: str, *, band=None):
    """A raster as contiguous float64 with NaN for NoData, and the DataArray it came from (None for an
    array). Arrays (masked cells as NaN) and DataArrays (with their NoData value) are taken; the
    reference also opens paths with rasterio (advanced.py:34-54), which this path does not use."""
    if _is_path(values):
        raise ValueError(f"the {name} must be an array (raster files are read by the caller)")
    if hasattr(values, "dims") and hasattr(values, "values"):
        cells, like = _dataarray_values(values)
    else:
        cells, like = np.ma.asarray(values, dtype=np.float64).filled(np.nan).astype(np.float64), None
    if cells.ndim != 2:
        raise ValueError(f"the {name} must be a 2-D raster")
    return np.ascontiguousarray(cells), like


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
