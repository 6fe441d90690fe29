"""Circuitscape's raster modes: advanced, pairwise and one-to-all / all-to-one.

The Python layer of the reference's src/landschaft/circuit/circuitscape.py:
the readers of Circuitscape's file formats (:28-315), the array API
(:366-953) and the INI driver (:956-1286), with the same functions,
arguments, defaults and results. The solves run in the GPU graph engine
behind ``core()``.

Rasters are ESRI ASCII grids (parsed in double, as Circuitscape does) or
GeoTIFF files read by the small reader below (strips or tiles, uncompressed
or deflate, integer or float samples, band 1); the reference reads the
latter through rasterio, which this package does not use.
"""

from __future__ import annotations

import gzip
import math
import os
import struct
import warnings
import zlib
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

from ._backend import core
from .advanced import _grid

NODATA = -9999.0
POLICIES = ("keepall", "rmvsrc", "rmvgnd", "rmvall")
# one-to-all / all-to-one problems built on full rasters, as the reference
# does (circuitscape.py:771-806), instead of from the focal cells alone
DENSE_PROBLEMS = False
SCENARIOS = {"pairwise": "pairwise", "Pairwise": "pairwise", "advanced": "advanced", "Advanced": "advanced",
             "one-to-all": "one-to-all", "one_to_all": "one-to-all", "all-to-one": "all-to-one",
             "all_to_one": "all-to-one"}
TRUELIST = ("True", "true", "1")
FALSELIST = ("False", "false", "0")


# ---- reading Circuitscape's inputs (circuitscape.py:28-315)

@dataclass
class RasterMeta:
    """The habitat's grid: size, lower-left corner, cell size, GDAL-order transform
    ``(x0, dx, rot, y0, rot, -dy)``, projection WKT and, for GeoTIFF inputs, the raw GeoKey tags
    (``crs``), which written GeoTIFF outputs carry over."""

    nrows: int
    ncols: int
    xllcorner: float
    yllcorner: float
    cellsize: float
    transform: tuple = (0.0, 1.0, 0.0, 0.0, 0.0, -1.0)
    wkt: str = ""
    crs: object = None


def _open(path):
    p = os.fspath(path)
    return gzip.open(p, "rt", encoding="utf-8") if p.lower().endswith("gz") else open(p, encoding="utf-8")


def _read_bytes(path) -> bytes:
    p = os.fspath(path)
    with (gzip.open if p.lower().endswith("gz") else open)(p, "rb") as f:
        return f.read()


def file_type(path) -> str:
    """Circuitscape's file-type guess (circuitscape.py:48-64): ``geotiff`` (TIFF magic), ``npy``,
    ``aagrid`` (``ncols`` first), ``pairs_aagrid`` (``min``), ``pairs`` (``mode``) or ``list``."""
    p = os.fspath(path)
    with (gzip.open if p.lower().endswith("gz") else open)(p, "rb") as f:
        head = f.read(4)
    if head[2:4] == b"\x2a\x00":
        return "geotiff"
    with _open(p) as f:
        first = f.readline()
    if first.startswith("\x93NUMPY"):
        return "npy"
    if first.lower().startswith("ncols"):
        return "aagrid"
    if first.startswith("min"):
        return "pairs_aagrid"
    if first.startswith("mode"):
        return "pairs"
    return "list"


_ASC_KEYS = ("ncols", "nrows", "xllcorner", "yllcorner", "xllcenter", "yllcenter", "cellsize", "dx", "dy",
             "nodata_value")


def _read_asc(path):
    """An ESRI ASCII grid in double (circuitscape.py:70-93): header keys until the first data row,
    NoData and NaN cells -9999, corners from ``*llcenter`` when given so."""
    head: dict = {}
    values: list = []
    nrow_tokens = 0
    with _open(path) as f:
        for line in f:
            parts = line.split()
            if not parts:
                continue
            if not values and len(parts) == 2 and parts[0].lower() in _ASC_KEYS:
                head[parts[0].lower()] = float(parts[1])
                continue
            values.extend(float(t) for t in parts)
            nrow_tokens += 1
    ny, nx = int(head["nrows"]), int(head["ncols"])
    if nrow_tokens != ny or len(values) != ny * nx:
        raise ValueError(f"{path}: the ASCII grid does not hold {ny} rows of {nx} values")
    a = np.array(values, dtype=np.float64).reshape(ny, nx)
    dx = head.get("dx", head.get("cellsize", math.nan))
    dy = head.get("dy", dx)
    x0 = head["xllcorner"] if "xllcorner" in head else head["xllcenter"] - dx / 2
    y0 = head["yllcorner"] if "yllcorner" in head else head["yllcenter"] - dy / 2
    a[(a == head.get("nodata_value", NODATA)) | np.isnan(a)] = NODATA
    p = os.fspath(path)
    prj = Path((p[:-3] if p.endswith(".gz") else p)[:-4] + ".prj")
    wkt = prj.read_text().strip() if prj.exists() else ""
    return a, RasterMeta(ny, nx, x0, y0, dx, (x0, dx, 0.0, y0 + ny * dy, 0.0, -dy), wkt)


# TIFF field types: struct code and size
_TIFF_TYPES = {1: ("B", 1), 2: ("s", 1), 3: ("H", 2), 4: ("I", 4), 5: ("I", 8), 6: ("b", 1), 7: ("B", 1),
               8: ("h", 2), 9: ("i", 4), 10: ("i", 8), 11: ("f", 4), 12: ("d", 8), 16: ("Q", 8), 17: ("q", 8)}
_GEOKEY_TAGS = (34735, 34736, 34737)


def _tiff_tags(buf: bytes):
    """The first IFD of a classic or Big TIFF as ``(byte order, {tag: values})``."""
    order = {b"II": "<", b"MM": ">"}.get(buf[:2])
    if order is None:
        raise ValueError("not a TIFF file")
    magic = struct.unpack_from(order + "H", buf, 2)[0]
    if magic not in (42, 43):
        raise ValueError("not a TIFF file")
    big = magic == 43
    word = "Q" if big else "I"
    pos = struct.unpack_from(order + word, buf, 8 if big else 4)[0]
    n = struct.unpack_from(order + ("Q" if big else "H"), buf, pos)[0]
    pos += 8 if big else 2
    tags: dict = {}
    for _ in range(n):
        tag, typ = struct.unpack_from(order + "HH", buf, pos)
        count = struct.unpack_from(order + word, buf, pos + 4)[0]
        at = pos + (12 if big else 8)
        pos += 20 if big else 12
        if typ not in _TIFF_TYPES:
            continue
        code, size = _TIFF_TYPES[typ]
        nbytes = count * size
        if nbytes > (8 if big else 4):
            at = struct.unpack_from(order + word, buf, at)[0]
        raw = buf[at:at + nbytes]
        if typ == 2:
            tags[tag] = raw.split(b"\0")[0].decode("latin-1")
        elif typ in (5, 10):
            v = struct.unpack(order + f"{2 * count}{code}", raw)
            tags[tag] = tuple(v[k] / v[k + 1] for k in range(0, len(v), 2))
        else:
            tags[tag] = struct.unpack(order + f"{count}{code}", raw)
    return order, tags


def _read_tiff(path):
    """Band 1 of a GeoTIFF as ``(samples in their own dtype, GDAL nodata text or None, RasterMeta)``."""
    buf = _read_bytes(path)
    order, t = _tiff_tags(buf)
    nx, ny = t[256][0], t[257][0]
    bits, comp = t.get(258, (1,))[0], t.get(259, (1,))[0]
    spp, planar = t.get(277, (1,))[0], t.get(284, (1,))[0]
    fmt, pred = t.get(339, (1,))[0], t.get(317, (1,))[0]
    if comp not in (1, 8, 32946) or pred != 1 or fmt not in (1, 2, 3) or bits % 8:
        raise ValueError(f"{path}: only uncompressed or deflate GeoTIFF without a predictor is read here")
    dt = np.dtype(f"{order}{'uif'[fmt - 1]}{bits // 8}")
    per = spp if planar == 1 else 1

    def chunk(off, cnt, n):
        raw = buf[off:off + cnt]
        raw = zlib.decompress(raw) if comp != 1 else raw
        return np.frombuffer(raw, dt, count=n * per).reshape(-1, per)[:, 0]

    if 322 in t:
        tw, th = t[322][0], t[323][0]
        tx, ty = -(-nx // tw), -(-ny // th)
        band = np.empty((ty * th, tx * tw), dt)
        for k in range(tx * ty):  # band 1's tiles come first
            r0, c0 = (k // tx) * th, (k % tx) * tw
            band[r0:r0 + th, c0:c0 + tw] = chunk(t[324][k], t[325][k], tw * th).reshape(th, tw)
        band = band[:ny, :nx]
    else:
        rps = min(t.get(278, (ny,))[0], ny)
        strips = []
        for k in range(-(-ny // rps)):
            rows = min(rps, ny - k * rps)
            strips.append(chunk(t[273][k], t[279][k], rows * nx).reshape(rows, nx))
        band = np.concatenate(strips)
    if 34264 in t:
        m = t[34264]
        a_, b_, c_, d_, e_, f_ = m[0], m[1], m[3], m[4], m[5], m[7]
    elif 33550 in t and 33922 in t:
        sx, sy = t[33550][0], t[33550][1]
        i, j, _, x, y, _ = t[33922][:6]
        a_, b_, c_, d_, e_, f_ = sx, 0.0, x - i * sx, 0.0, -sy, y + j * sy
    else:
        a_, b_, c_, d_, e_, f_ = 1.0, 0.0, 0.0, 0.0, 1.0, 0.0
    nodata = t.get(42113)
    keys = {k: t[k] for k in _GEOKEY_TAGS if k in t}
    meta = RasterMeta(ny, nx, c_, f_ - ny * a_, a_, (c_, a_, b_, f_, d_, e_), "", keys or None)
    return band, nodata, meta


def read_raster(path):
    """Circuitscape's raster read (circuitscape.py:96-120): an ESRI ASCII grid parsed in double, or a
    GeoTIFF, with its nodata value and NaN set to -9999."""
    p = os.fspath(path)
    if not os.path.isfile(p):
        raise FileNotFoundError(f'the file "{p}" does not exist')
    if p.lower().removesuffix(".gz").endswith(".asc"):
        return _read_asc(p)
    if file_type(p) != "geotiff":
        raise ValueError(f"{p}: only ESRI ASCII grids and GeoTIFF rasters are read")
    band, nodata, meta = _read_tiff(p)
    a = band.astype(np.float64)
    if nodata is not None and nodata.strip():
        nd = float(nodata.strip())
        if np.issubdtype(band.dtype, np.integer):
            nd = float(np.int64(nd)) if math.isfinite(nd) else nd
        else:
            nd = float(np.array(nd).astype(band.dtype))
        a[a == nd] = NODATA
    a[np.isnan(a)] = NODATA
    return a, meta


def _read_rows(path) -> np.ndarray:
    rows = []
    with _open(path) as f:
        for line in f:
            parts = line.split()
            if parts:
                rows.append([float(t) for t in parts])
    return rows


def _read_list(path) -> np.ndarray:
    """A point list: ``id x y`` per line."""
    return np.array(_read_rows(path), dtype=np.float64).reshape(-1, 3)


def _list_cells(points: np.ndarray, meta: RasterMeta):
    # 1-based row and column of a point (circuitscape.py:133-137)
    i = np.ceil(meta.nrows - (points[:, 2] - meta.yllcorner) / meta.cellsize).astype(np.int64)
    j = np.ceil((points[:, 1] - meta.xllcorner) / meta.cellsize).astype(np.int64)
    return i, j


def _check_meta(meta: RasterMeta, habitat: RasterMeta) -> None:
    for name in ("cellsize", "ncols", "nrows", "yllcorner", "xllcorner"):
        if getattr(meta, name) != getattr(habitat, name):
            warnings.warn(f"{name} is not the same", stacklevel=4)
            return


def read_polymap(path, habitat: RasterMeta, *, integer: bool = False, nodata_as=0) -> np.ndarray:
    """A raster on the habitat's grid (circuitscape.py:147-156), NoData set to ``nodata_as``
    (kept as -9999 with None); ``integer`` rejects non-integer values."""
    a, meta = read_raster(path)
    if integer and not np.all(a == np.round(a)):
        raise ValueError(f"{path} contains non-integer values")
    if nodata_as is not None:
        a[a == NODATA] = nodata_as
    _check_meta(meta, habitat)
    return a


def _read_list_n(path, ncol: int) -> np.ndarray:
    t = np.array(_read_rows(path), dtype=np.float64)
    if t.ndim != 2 or t.shape[1] != ncol:
        raise ValueError(f"{path} must have exactly {ncol} columns")
    return t


def read_reclass_table(path) -> dict:
    """A reclass table (two columns, old and new value) as a lookup (circuitscape.py:159-163)."""
    return {float(a): float(b) for a, b in _read_list_n(path, 2)}


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


def read_point_map(path, habitat: RasterMeta) -> FocalPoints:
    """Focal points (circuitscape.py:209-229): a raster of ids, cells of one id a focal region, or a
    text list of ``id x y``; negative ids dropped, stably sorted by id."""
    if file_type(path) == "list":
        pts = _read_list(path)
        rows, cols = _list_cells(pts, habitat)
        ids = pts[:, 0]
    else:
        a = read_polymap(path, habitat, integer=True)
        c0, r0 = np.nonzero(a.T)  # column-major cell order
        rows, cols, ids = r0 + 1, c0 + 1, a[r0, c0]
    keep = ids >= 0
    order = np.argsort(ids[keep], kind="stable")
    rows = rows[keep][order].astype(np.int64)
    cols = cols[keep][order].astype(np.int64)
    ids = ids[keep][order].astype(np.int64)
    if len(ids) and (rows.min() < 0 or cols.min() < 0 or rows.max() > habitat.nrows or cols.max() > habitat.ncols):
        raise ValueError("At least one focal node location falls outside of habitat map")
    if len(np.unique(ids)) < 2:
        raise ValueError("Less than two valid focal nodes found. Please check focal node location file.")
    return FocalPoints(rows, cols, ids)


def read_included_pairs(path, *, starts_from_zero: bool = False) -> IncludedPairs:
    """Include / exclude pairs (circuitscape.py:283-311): a ``min``/``max`` matrix of ids (a pair is
    included when min <= value <= max) or a ``mode include|exclude`` list of id pairs.
    ``starts_from_zero`` shifts the ids of a network numbered from 0 by one."""
    kind = file_type(path)
    with _open(path) as f:
        lines = [ln.split() for ln in f.read().splitlines()]
    shift = 1 if starts_from_zero else 0
    if kind == "pairs_aagrid":
        lo, hi = float(lines[0][1]), float(lines[1][1])
        a = np.array([[float(t) for t in ln] for ln in lines[2:] if ln])
        m = a[1:, 1:]
        m = np.where(m > hi, 0, m)
        return IncludedPairs("include", a[0, 1:].astype(np.int64) + shift, (m >= lo).astype(np.int64))
    if kind == "pairs":
        mode = lines[0][1]
        pairs = np.array([[float(t) for t in ln] for ln in lines[1:] if ln]).reshape(-1, 2) + shift
        ids = np.unique(pairs).astype(np.int64)
        ids = ids[ids != 0]
        where = {int(v): k for k, v in enumerate(ids)}
        m = np.zeros((len(ids), len(ids)), dtype=np.int64)
        for a, b in pairs.astype(np.int64):
            if int(a) in where and int(b) in where:
                m[where[int(a)], where[int(b)]] = m[where[int(b)], where[int(a)]] = 1
        return IncludedPairs(mode, ids, m)
    raise ValueError("Error reading focal node include/exclude pairs file.")


def read_point_strengths(path) -> np.ndarray:
    """Source strengths (circuitscape.py:704-709): columns id and strength; ids numbered from 0
    are shifted by one."""
    t = _read_list_n(path, 2)
    if t[:, 0].min() == 0:
        t[:, 0] += 1
    return t


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
    return _advanced(g, poly, src, gnd, four_neighbors, avg_resistances, remove_src_or_gnd, unit_currents,
                     direct_grounds, n_threads)


def _on_cells(nodemap: np.ndarray, per_node: np.ndarray) -> np.ndarray:
    """Values per node spread over each node's cells; 0 on cells without a node."""
    if per_node.size == 0:
        return np.zeros(nodemap.shape)
    return np.where(nodemap >= 0, per_node[np.maximum(nodemap, 0)], 0.0)


def _advanced(g, poly, src, gnd, four, avg, policy, unit, direct, n_threads) -> CircuitscapeResult:
    """One advanced solve (circuitscape.py:499-516): amps per cell (1 A with ``unit``), ground
    conductances per cell (every ground direct with ``direct``), the maps back on the cells."""
    amps = np.where(src != 0, 1.0, src) if unit else src
    to_ground = np.where(gnd != 0, np.inf, gnd) if direct else gnd
    as_f64 = [np.ascontiguousarray(x, dtype=np.float64) for x in (g, amps, to_ground)]
    out = core().circuitscape_advanced(as_f64[0], poly, as_f64[1], as_f64[2], bool(avg), bool(four), policy,
                                       int(n_threads))
    nm = out["nodemap"]
    result = CircuitscapeResult(voltage=_on_cells(nm, out["voltage"]), current=_on_cells(nm, out["current"]),
                                nodemap=nm)
    result.maps["solved"] = bool(out["solved"])
    return result


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
    maps: dict = {}

    def emit(a, b, volt, cur):
        if volt is not None:
            maps[("voltage", a, b)] = volt
        if cur is not None:
            maps[("current", a, b)] = cur

    opts = _pair_options(voltage_maps, current_maps or cum_map, not current_maps and cum_map, max_map,
                         set_focal_node_currents_to_zero, set_null_currents_to_nodata, set_null_voltages_to_nodata,
                         log_transform_maps)
    res, r = _pairwise(g, poly, points, excluded, four_neighbors, avg_resistances, opts, emit, n_threads)
    return CircuitscapeResult(resistances=res, cum_current=r.get("cum"), max_current=r.get("max"), maps=maps)


def _pairwise(g, poly, points: FocalPoints, excluded: set, four, avg, opts: dict, emit, n_threads):
    """circuitscape.py:519-540: the resistance table (ids in the first row and column) and the
    binding's result; ``emit(id_a, id_b, volt, cur)`` receives each pair's maps. Focal regions
    (ids covering several cells) are solved one graph per pair of ids."""
    rows, cols = points.cells
    cells = np.ascontiguousarray(rows * g.shape[1] + cols, dtype=np.int64)
    if points.has_regions:
        opts = {**opts, "region_ids": np.ascontiguousarray(points.ids, dtype=np.int64)}
        ids = np.unique(points.ids)
    else:
        ids = points.ids
    n = len(ids)
    ex = np.zeros((n, n), dtype=np.uint8)
    for a in range(n):
        for b in range(n):
            ex[a, b] = (int(ids[a]), int(ids[b])) in excluded

    def named(ci, cj, volt, cur):
        emit(int(ids[ci]), int(ids[cj]), volt, cur)

    r = core().circuitscape_pairwise(np.ascontiguousarray(g, dtype=np.float64), poly, cells, ex, bool(avg),
                                     bool(four), opts, named, int(n_threads))
    res = np.zeros((n + 1, n + 1))
    res[0, 1:] = ids
    res[1:, 0] = ids
    res[1:, 1:] = r["resistances"]
    return res, r


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

    def problem_dense(i):
        # the reference's construction on full rasters
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

    # The same problems from the focal cells alone: every raster above is
    # zero off the focal cells, so its nonzero cells are found among them
    # (sorted, as flatnonzero gives them). An id 0 would match every empty
    # cell of the rasters, so such point sets take the full-raster path.
    pcells = np.unique(cells)
    pvals = point_map.ravel()[pcells]  # a cell listed twice keeps its last id, as the raster does
    fcells = np.unique(rows[first] * nx + cols[first])
    fvals = first_map.ravel()[fcells]
    fpos = np.searchsorted(pcells, fcells)

    def problem_sparse(i):
        pid = int(uids[i])
        pv = pvals.copy()
        newpoly = base_poly
        amps = float(strengths[i, 1]) if strengths is not None else 1.0
        if pairs is not None:
            row = int(np.nonzero(pairs.ids == pid)[0][0])
            drop = [pairs.ids[j] for j in range(len(pairs.ids)) if j != row and pairs.matrix[row, j] == mark]
            pv[np.isin(pv, drop)] = 0
            if not single:
                pm = point_map.copy()
                pm.ravel()[pcells] = pv
                newpoly = _point_polymap(poly, pm, ids)
        st = None
        if strengths is not None:
            # by position, as the reference maps them (it fails alike when the table and the points differ)
            st = strengths[:, 1].copy()
            st[pv[np.searchsorted(pcells, cells)] == 0] = 1.0
        if pv.sum() == pid:
            return None
        fpm = pv[fpos]  # the point map at the first cells
        if one_to_all:
            sc = fcells[fvals == pid]
            sv = np.full(len(sc), amps)
            gmask = (pv > 0) & (pv != pid)
            gc, gv = pcells[gmask], np.full(int(gmask.sum()), np.inf)
        else:
            if st is not None:
                sval = dict(zip(cells.tolist(), st[: len(rows)].tolist()))  # a repeated cell: the last value
                sv_all = np.array([sval[c] for c in pcells.tolist()])
                sv_all[np.isin(pcells, fcells[fvals == pid])] = 0.0
                sc, sv = pcells, sv_all
            else:
                keep = (fvals != 0) & (fpm != pid)
                sc, sv = fcells[keep], np.ones(int(keep.sum()))
            gmask = pv == pid
            gc, gv = pcells[gmask], np.full(int(gmask.sum()), np.inf)
        nz = sv != 0
        return {"n": pid, "strength": amps, "poly": newpoly, "src_cells": sc[nz].astype(np.int64),
                "src_vals": sv[nz].astype(np.float64), "gnd_cells": gc.astype(np.int64),
                "gnd_vals": gv.astype(np.float64), "check_cell": int(cells[np.nonzero(ids == pid)[0][0]])}

    problem = problem_dense if (DENSE_PROBLEMS or np.any(ids == 0)) else problem_sparse

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


# ---- INI files (circuitscape.py:971-1286)

# Circuitscape's option defaults (its default.ini), which every INI file overrides
DEFAULTS = {
    "version": "unknown", "connect_four_neighbors_only": "False", "connect_using_avg_resistances": "False",
    "use_polygons": "False", "polygon_file": "(Browse for a short-circuit region file)",
    "source_file": "(Browse for a current source file)", "ground_file": "(Browse for a ground point file)",
    "ground_file_is_resistances": "True", "use_unit_currents": "False", "use_direct_grounds": "False",
    "remove_src_or_gnd": "keepall", "use_mask": "False", "mask_file": "None", "data_type": "raster",
    "scenario": "not entered", "habitat_file": "(Browse for a resistance file)",
    "habitat_map_is_resistances": "True", "solver": "cg+amg", "parallelize": "False", "precision": "double",
    "use_variable_source_strengths": "False", "variable_source_file": "None", "use_included_pairs": "False",
    "included_pairs_file": "(Browse for a file with pairs to include or exclude)",
    "point_file": "(Browse for file with locations of focal points or regions)", "use_reclass_table": "False",
    "reclass_file": "(Browse for file with reclassification data)",
    "output_file": "(Choose a base name for output files)", "write_cur_maps": "False", "write_volt_maps": "False",
    "write_cum_cur_map_only": "False", "write_max_cur_maps": "False", "set_null_currents_to_nodata": "False",
    "set_null_voltages_to_nodata": "False", "set_focal_node_currents_to_zero": "False", "compress_grids": "False",
    "log_transform_maps": "False", "write_as_tif": "False",
}


def _unset(value: str) -> bool:
    # an empty path, None, or the GUI's "(Browse for ...)" placeholder
    return value == "" or value == "None" or value.startswith("(")


def _flag(cfg: dict, key: str) -> bool:
    v = cfg[key]
    if v in TRUELIST:
        return True
    if v in FALSELIST:
        return False
    raise ValueError(f'{key} = "{v}": expected True or False')


def _resolve(name: str, base: Path) -> Path:
    # paths are read from the working directory, else next to the INI file
    p = Path(name)
    if p.is_absolute() or p.exists():
        return p
    q = base / p
    return q if q.exists() else p


def parse_config(path) -> dict:
    """An INI file as Circuitscape reads it (circuitscape.py:1032-1036): ``[section]`` lines skipped,
    every other line split at its first ``=``."""
    from .ini import parse_ini

    return parse_ini(path)


def _write_tiff(path: str, values: np.ndarray, meta: RasterMeta) -> None:
    """A float64 GeoTIFF (little-endian, one uncompressed strip) on the habitat's grid, NoData -9999."""
    ny, nx = values.shape
    data = np.ascontiguousarray(values, dtype="<f8").tobytes()
    x0, dx, _, y0, _, ndy = meta.transform
    entries = [(256, 4, (nx,)), (257, 4, (ny,)), (258, 3, (64,)), (259, 3, (1,)), (262, 3, (1,)),
               (273, 4, (0,)), (277, 3, (1,)), (278, 4, (ny,)), (279, 4, (len(data),)), (284, 3, (1,)),
               (339, 3, (3,)), (33550, 12, (dx, -ndy, 0.0)), (33922, 12, (0.0, 0.0, 0.0, x0, y0, 0.0))]
    for tag, vals in sorted((meta.crs or {}).items()):
        entries.append((tag, {34735: 3, 34736: 12, 34737: 2}[tag], vals))
    entries.append((42113, 2, "-9999"))
    entries.sort(key=lambda e: e[0])
    code = {2: "s", 3: "H", 4: "I", 12: "d"}
    size = {2: 1, 3: 2, 4: 4, 12: 8}

    def payload(typ, vals):
        if typ == 2:
            raw = vals.encode("latin-1") + b"\0"
            return raw, len(raw)
        return struct.pack("<" + f"{len(vals)}{code[typ]}", *vals), len(vals)

    ifd_at = 8
    ifd_len = 2 + 12 * len(entries) + 4
    extra = bytearray()
    blobs = []
    for tag, typ, vals in entries:
        raw, count = payload(typ, vals)
        if len(raw) > 4:
            blobs.append((tag, typ, count, ifd_at + ifd_len + len(extra), None))
            extra += raw + (b"\0" if len(raw) % 2 else b"")
        else:
            blobs.append((tag, typ, count, None, raw.ljust(4, b"\0")))
    data_at = ifd_at + ifd_len + len(extra)
    out = bytearray(b"II" + struct.pack("<HI", 42, ifd_at))
    out += struct.pack("<H", len(entries))
    for tag, typ, count, at, inline in blobs:
        if tag == 273:
            inline = struct.pack("<I", data_at)
        out += struct.pack("<HHI", tag, typ, count) + (inline if inline is not None else struct.pack("<I", at))
    out += struct.pack("<I", 0) + extra + data
    Path(path).write_bytes(bytes(out))


def _write_grid(prefix: str, values: np.ndarray, meta: RasterMeta, tif: bool) -> None:
    """One output map (circuitscape.py:1011-1029): ``prefix.asc`` with every value in its shortest
    round-trip form (and ``prefix.prj`` when the habitat had a projection), or ``prefix.tif``; NaN written as -9999."""
    out = np.where(np.isnan(values), NODATA, values)
    if tif:
        _write_tiff(prefix + ".tif", out, meta)
        return
    head = (f"ncols         {meta.ncols}\nnrows         {meta.nrows}\n"
            f"xllcorner     {meta.xllcorner!r}\nyllcorner     {meta.yllcorner!r}\n"
            f"cellsize      {meta.cellsize!r}\nNODATA_value  -9999\n")
    with open(prefix + ".asc", "wb") as f:
        f.write(head.encode())
        f.write(core().ascii_grid_rows(out))
    if meta.wkt:
        Path(prefix + ".prj").write_text(meta.wkt)


def _write_matrix(path: str, a: np.ndarray) -> None:
    with open(path, "w", encoding="utf-8") as f:
        for row in np.atleast_2d(a):
            f.write(" ".join(repr(float(v)) for v in row) + "\n")


def _prefix(cfg: dict):
    out = cfg["output_file"]
    if _unset(out):
        raise ValueError("output_file is not set")
    Path(out).parent.mkdir(parents=True, exist_ok=True)
    return out.split(".out")[0], _flag(cfg, "write_as_tif")


def _source_and_ground(source_path, ground_path, meta: RasterMeta, ground_is_resistance: bool):
    """Advanced mode's sources (amps per cell) and ground conductances (circuitscape.py:1230-1258):
    rasters on the habitat's grid or ``value x y`` lists; NoData grounds are none, a zero resistance
    a direct ground."""
    kind = file_type(ground_path)
    if kind in ("aagrid", "geotiff"):
        gnd = read_polymap(ground_path, meta, nodata_as=None)
    elif kind == "list":
        pts = _read_list(ground_path)
        gnd = np.full((meta.nrows, meta.ncols), NODATA)
        i, j = _list_cells(pts, meta)
        gnd[i - 1, j - 1] = pts[:, 0]
    else:
        raise ValueError("Cannot recognise file type.")
    kind = file_type(source_path)
    if kind in ("aagrid", "geotiff"):
        src = read_polymap(source_path, meta)
        src[src == NODATA] = 0.0
    elif kind == "list":
        pts = _read_list(source_path)
        src = np.zeros((meta.nrows, meta.ncols))
        i, j = _list_cells(pts, meta)
        src[i - 1, j - 1] = pts[:, 0]
    else:
        raise ValueError("Cannot recognize file type.")
    missing = gnd == NODATA
    if ground_is_resistance:
        with np.errstate(divide="ignore"):
            gnd = 1.0 / gnd
    gnd[missing] = 0.0
    return src, gnd


class _IniRun:
    """A Circuitscape INI file over its defaults: flags, and the input files it names (found in the
    working directory, else next to the INI file)."""

    def __init__(self, cfg: dict, base: Path):
        self.cfg, self.base = cfg, base

    def flag(self, key: str) -> bool:
        return _flag(self.cfg, key)

    def file(self, key: str) -> Path:
        value = self.cfg[key]
        if _unset(value):
            raise ValueError(f"{key} is not set")
        path = _resolve(value, self.base)
        if not path.is_file():
            raise FileNotFoundError(f'{key} = "{value}" does not exist')
        return path

    def optional_file(self, switch: str, key: str):
        return self.file(key) if self.flag(switch) else None

    def scenario(self) -> str:
        name = SCENARIOS.get(self.cfg["scenario"])
        if name is None:
            raise ValueError(f'scenario = "{self.cfg["scenario"]}" is not one of {sorted(set(SCENARIOS.values()))}')
        return name

    def policy(self) -> str:
        chosen = self.cfg["remove_src_or_gnd"]
        chosen = "keepall" if chosen == "not entered" else chosen  # Circuitscape 4's default
        if chosen not in POLICIES:
            raise ValueError(f"remove_src_or_gnd must be one of {POLICIES}")
        return chosen

    def habitat(self):
        """The conductance on the habitat's grid (reclassified, masked), the polygons and the grid."""
        habitat = self.file("habitat_file")
        reclass_file = self.optional_file("use_reclass_table", "reclass_file")
        cells, meta = read_raster(habitat)
        g = habitat_conductance(cells, is_resistance=self.flag("habitat_map_is_resistances"),
                                reclass=None if reclass_file is None else read_reclass_table(reclass_file))
        poly = None
        if self.flag("use_polygons"):
            poly = np.ascontiguousarray(read_polymap(self.file("polygon_file"), meta, integer=True), dtype=np.int64)
        if self.flag("use_mask"):
            g = g * (read_polymap(self.file("mask_file"), meta) > 0)
            if g.sum() == 0:
                raise ValueError("Mask file deleted everything!")
        return g, poly, meta


def circuitscape_ini(path, *, write_outputs: bool = True, n_threads: int = 0):
    """Run a Circuitscape INI file, as ``Circuitscape.compute(path)`` does (circuitscape.py:1039-1213).

    Rasters (``data_type = raster``): advanced, pairwise, one-to-all and
    all-to-one, with polygons, masks, reclass tables, include / exclude pairs,
    variable source strengths and focal regions, giving a
    :class:`CircuitscapeResult`. Networks (``data_type = network``): pairwise
    or advanced on an edge list, giving a ``NetworkResult``. Paths are read
    from the working directory, else next to the INI file. With
    ``write_outputs`` the maps go next to ``output_file`` under Circuitscape's
    names (``<base>_voltmap``, ``<base>_curmap_<i>_<j>``, ``<base>_cum_curmap``
    ...; ASCII grids, or GeoTIFF with ``write_as_tif``) and ``output_file``
    holds the configuration; otherwise the per-pair / per-point maps are
    returned in ``result.maps`` under the same names. ``solver``,
    ``precision`` and ``parallelize`` have no effect: every solve runs on the
    GPU to the solver settings' residual.
    """
    base = Path(path).resolve().parent
    cfg = {**DEFAULTS, **parse_config(path)}
    kind = cfg["data_type"].lower() if cfg["data_type"] in ("raster", "Raster", "network", "Network") else None
    if kind == "network":
        from .network import network_ini

        return network_ini(cfg, base, write_outputs=write_outputs, n_threads=n_threads)
    if kind != "raster":
        raise ValueError(f'data_type = "{cfg["data_type"]}": expected raster or network')
    run = _IniRun(cfg, base)
    scenario = run.scenario()
    for key in ("low_memory_mode", "preemptive_memory_release"):
        if cfg.get(key, "False") in TRUELIST:
            raise ValueError(f"{key} is not implemented in Circuitscape.jl; remove it or set it to False")
    g, poly, meta = run.habitat()
    four, avg = run.flag("connect_four_neighbors_only"), run.flag("connect_using_avg_resistances")
    policy = run.policy()
    # the per-pair and per-point maps: written as they come, or kept in result.maps
    prefix, tif = _prefix(cfg) if write_outputs and scenario != "advanced" else (None, False)
    kept: dict = {}

    def deliver(name, grid):
        if grid is not None:
            if write_outputs:
                _write_grid(f"{prefix}_{name}", grid, meta, tif)
            else:
                kept[name] = grid

    if scenario == "advanced":
        src, gnd = _source_and_ground(run.file("source_file"), run.file("ground_file"), meta,
                                      run.flag("ground_file_is_resistances"))
        result = _advanced(g, poly, src, gnd, four, avg, policy, run.flag("use_unit_currents"),
                           run.flag("use_direct_grounds"), n_threads)
    else:
        points = read_point_map(run.file("point_file"), meta)
        pairs_file = run.optional_file("use_included_pairs", "included_pairs_file")
        pairs = None if pairs_file is None else read_included_pairs(pairs_file)
        if scenario == "pairwise":
            points, excluded = _apply_included_pairs(points, pairs)
            opts = _pair_options(*(run.flag(k) for k in (
                "write_volt_maps", "write_cur_maps", "write_cum_cur_map_only", "write_max_cur_maps",
                "set_focal_node_currents_to_zero", "set_null_currents_to_nodata", "set_null_voltages_to_nodata",
                "log_transform_maps")))
            res, out = _pairwise(g, poly, points, excluded, four, avg, opts,
                                 lambda a, b, volt, cur: (deliver(f"voltmap_{a}_{b}", volt),
                                                          deliver(f"curmap_{a}_{b}", cur)), n_threads)
            result = CircuitscapeResult(resistances=res, cum_current=out.get("cum"), max_current=out.get("max"))
        else:
            strengths_file = run.optional_file("use_variable_source_strengths", "variable_source_file")
            strengths = None if strengths_file is None else read_point_strengths(strengths_file)
            # Circuitscape writes every point's current map with write_cum_cur_map_only too
            point_maps = run.flag("write_cur_maps") or run.flag("write_cum_cur_map_only")
            opts = {"volt_maps": run.flag("write_volt_maps"), "cur_maps": point_maps, "cum": point_maps,
                    "max": run.flag("write_max_cur_maps"), "focal_zero": run.flag("set_focal_node_currents_to_zero")}
            result = _one_to_all(g, poly, points, strengths, pairs, scenario == "one-to-all", four, avg, opts,
                                 lambda n, volt, cur: (deliver(f"voltmap_{n}", volt), deliver(f"curmap_{n}", cur)),
                                 n_threads)
    result.maps.update(kept)
    if write_outputs:
        _write_outputs(cfg, scenario, result, meta)
    return result


def _write_outputs(cfg: dict, scenario: str, result: CircuitscapeResult, meta: RasterMeta) -> None:
    """The configuration echo, the run's maps and, for pairwise, the resistance tables
    (circuitscape.py:1261-1286)."""
    prefix, tif = _prefix(cfg)
    Path(cfg["output_file"]).write_text("".join(f"{k} = {v}\n" for k, v in cfg.items()), encoding="utf-8")
    if scenario == "advanced":
        grids = [("_voltmap", result.voltage, _flag(cfg, "write_volt_maps")),
                 ("_curmap", result.current, _flag(cfg, "write_cur_maps") or _flag(cfg, "write_cum_cur_map_only"))]
    else:
        grids = [("_cum_curmap", result.cum_current, result.cum_current is not None),
                 ("_max_curmap", result.max_current, result.max_current is not None)]
    for suffix, grid, wanted in grids:
        if wanted:
            _write_grid(prefix + suffix, grid, meta, tif)
    if scenario == "pairwise":
        table = result.resistances
        ids = table[1:, 0]
        lo, hi = np.triu_indices(ids.size, 1)
        _write_matrix(prefix + "_resistances.out", table)
        _write_matrix(prefix + "_resistances_3columns.out",
                      np.stack([ids[lo], ids[hi], table[hi + 1, lo + 1]], axis=1).reshape(-1, 3))
