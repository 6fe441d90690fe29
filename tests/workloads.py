"""Inputs of the benchmark workloads for the parity tests, and summaries of large output maps.

The map summaries (seeded samples, row and column sums, focal values) are what the golden file
``tests/golden/workloads.npz`` holds for maps too large to keep whole; ``make_workload_vectors.py``
writes it with the FP64 oracle (``oracle_core.OracleCore`` in place of the library for the front end).
"""

from __future__ import annotations

import hashlib
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

import bench  # noqa: E402

GOLDEN = ROOT / "tests" / "golden" / "workloads.npz"


def digest(*arrays) -> str:
    """SHA-256 of the arrays' bytes: the golden file records it for the inputs it was made from."""
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def summary(a: np.ndarray, seed: int, count: int = 4096) -> dict:
    """A seeded sample of the map's cells and its row and column sums (NaN cells left out of the sums)."""
    idx = np.random.default_rng(seed).choice(a.size, min(count, a.size), replace=False)
    return {"idx": idx, "sample": a.ravel()[idx], "rows": np.nansum(a, axis=1), "cols": np.nansum(a, axis=0)}


def assert_summary_close(got: np.ndarray, z, key: str, rtol: float, atol_frac: float) -> float:
    """Compares a map with its golden summary: cells at rtol / atol_frac * max|expected|, the row and
    column sums at rtol / atol_frac * max|expected| * the sum's length. Returns the largest relative
    difference of the sampled cells (over max|expected|)."""
    idx, sample = z[f"{key}/idx"], z[f"{key}/sample"]
    top = float(np.nanmax(np.abs(sample)))
    top = max(top, float(np.nanmax(np.abs(z[f"{key}/rows"]))) / got.shape[1])
    g = got.ravel()[idx]
    np.testing.assert_allclose(g, sample, rtol=rtol, atol=atol_frac * top)
    np.testing.assert_allclose(np.nansum(got, axis=1), z[f"{key}/rows"], rtol=rtol, atol=atol_frac * top * got.shape[1])
    np.testing.assert_allclose(np.nansum(got, axis=0), z[f"{key}/cols"], rtol=rtol, atol=atol_frac * top * got.shape[0])
    return float(np.max(np.abs(g - sample)) / top) if top > 0 else 0.0


# ---- the workloads

def pairwise_2048():
    """Pairwise: 2048 x 2048, r = exp(2 z) in float32, 4 neighbours, 16 focal cells (120 pairs)."""
    return bench.pairwise_inputs(2048, 102, 16, 2.0)


def adjoint_1024():
    """Gradient: 1024 x 1024 lognormal sigma 1 in float32, 8 neighbours, 32 (source, ground) pairs."""
    return bench.adjoint_inputs(1024, 32, 104)


def large_graph_crop(r_full: np.ndarray | None = None, size: int = 1024):
    """A size x size crop (seeded offset) of the 8192 x 8192 large-graph map, with 64 seeded focal cells
    among its data cells, as a raster of point ids 1..64."""
    if r_full is None:
        r_full, _ = bench.large_graph_inputs()
    rng = np.random.default_rng(1105)
    y0, x0 = (int(v) for v in rng.integers(0, r_full.shape[0] - size, 2))
    r = np.ascontiguousarray(r_full[y0:y0 + size, x0:x0 + size])
    cells = rng.choice(np.flatnonzero(~np.isnan(r)), 64, replace=False)
    return r, points_map(r.shape, cells)


def points_map(shape, cells) -> np.ndarray:
    pm = np.zeros(shape)
    pm.flat[np.asarray(cells)] = np.arange(1, len(cells) + 1)
    return pm


def omniscape_crop(size: int = 448):
    """A size x size crop (seeded offset) of the 4096 x 4096 Omniscape benchmark map: resistance and
    sources 1 / r."""
    r, s = bench.workload(4096)
    rng = np.random.default_rng(1103)
    y0, x0 = (int(v) for v in rng.integers(0, 4096 - size, 2))
    return np.ascontiguousarray(r[y0:y0 + size, x0:x0 + size]), np.ascontiguousarray(s[y0:y0 + size, x0:x0 + size])


def all_to_one(r, points, core=None):
    """Circuitscape's all-to-one mode on the resistance raster r (4 neighbours), cumulative map only;
    ``core`` replaces the library (``oracle_core.OracleCore``)."""
    from landschaft_amd.circuit import circuitscape as cs

    if core is None:
        return cs.circuitscape_one_to_all(np.where(np.isnan(r), cs.NODATA, r), points, all_to_one=True,
                                          four_neighbors=True, cum_map=True)
    saved = cs.core
    cs.core = lambda: core
    try:
        return cs.circuitscape_one_to_all(np.where(np.isnan(r), cs.NODATA, r), points, all_to_one=True,
                                          four_neighbors=True, cum_map=True)
    finally:
        cs.core = saved
