"""Raster inputs of the circuit API (the reference's src/landschaft/circuit/advanced.py:34-66)."""

from __future__ import annotations

import numpy as np


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
