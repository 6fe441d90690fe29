"""Omniscape's INI files: ``run_omniscape(path)`` on the GPU path.

Mirrors the reference's src/landschaft/circuit/ini.py (:14-40 options and
defaults, :41-75 parsing, :142-287 the driver). Rasters are ESRI ASCII grids
(``.asc``), read as GDAL's AAIGrid driver types them (Int32, or Float32 once
any value has a decimal point or an exponent) so that the values match the
reference's rasterio reads bit for bit; other formats need rasterio, which
this path does not use.
"""

from __future__ import annotations

import os
import warnings
from pathlib import Path

import numpy as np

from .omniscape import omniscape, read_reclass_table, reclassify

TRUELIST = ("true", "True", "TRUE")
REQUIRED = ("resistance_file", "radius", "project_name")
SUPPORTED = (
    "resistance_file", "resistance_is_conductance", "source_file", "project_name", "parallelize",
    "parallel_batch_size", "block_size", "radius", "buffer", "source_threshold", "source_from_resistance",
    "r_cutoff", "precision", "connect_four_neighbors_only", "solver", "calc_flow_potential",
    "calc_normalized_current", "correct_artifacts", "write_raw_currmap", "write_as_tif", "mask_nodata",
    "suppress_cs_messages", "conditional", "n_conditions", "compare_to_future", "condition1_file",
    "condition2_file", "condition1_future_file", "condition2_future_file", "comparison1", "comparison2",
    "condition1_lower", "condition2_lower", "condition1_upper", "condition2_upper", "reclassify_resistance",
    "reclass_table", "write_reclassified_resistance", "allow_different_projections",
)
DEFAULTS = {
    "resistance_file": "", "resistance_is_conductance": "false", "source_file": "",
    "parallelize": "true", "parallel_batch_size": "10", "block_size": "1", "buffer": "0",
    "source_threshold": "0", "source_from_resistance": "false", "r_cutoff": "Inf", "precision": "double",
    "connect_four_neighbors_only": "false", "solver": "cg+amg", "calc_flow_potential": "false",
    "calc_normalized_current": "false", "correct_artifacts": "true", "write_raw_currmap": "true",
    "write_as_tif": "true", "mask_nodata": "true", "suppress_cs_messages": "true", "conditional": "false",
    "n_conditions": "1", "compare_to_future": "none", "condition1_file": "", "condition2_file": "",
    "condition1_future_file": "", "condition2_future_file": "", "comparison1": "within", "comparison2": "within",
    "condition1_lower": "0", "condition2_lower": "0", "condition1_upper": "0", "condition2_upper": "0",
    "reclassify_resistance": "false", "reclass_table": "", "write_reclassified_resistance": "false",
    "allow_different_projections": "false",
}
_HEADER = ("ncols", "nrows", "xllcorner", "yllcorner", "xllcenter", "yllcenter", "cellsize", "dx", "dy",
           "nodata_value")


def parse_ini(path) -> dict:
    """An INI file as Omniscape's ``parse_cfg`` reads it: ``[`` lines skipped, every other line split at
    its first ``=`` (key right-stripped, value stripped)."""
    cfg: dict = {}
    with open(path, encoding="utf-8") as f:
        for line in f:
            if line.startswith("["):
                continue
            key, sep, value = line.partition("=")
            if sep:
                cfg[key.rstrip()] = value.strip()
    return cfg


def read_ascii_grid(path):
    """An ESRI ASCII grid as (values float64 with NaN for NoData, header dict).

    The values go through the type GDAL's AAIGrid driver gives the band
    (Int32, or Float32 when any value is written with a decimal point or an
    exponent), then to float64, as a rasterio read does.
    """
    header: dict = {}
    tokens: list = []
    with open(path, encoding="utf-8") as f:
        for line in f:
            parts = line.split()
            if not parts:
                continue
            if parts[0].lower() in _HEADER and not tokens:
                header[parts[0].lower()] = parts[1]
                continue
            tokens.extend(parts)
    nx, ny = int(header["ncols"]), int(header["nrows"])
    if len(tokens) != nx * ny:
        raise ValueError(f"{path}: {len(tokens)} values for a {ny} x {nx} grid")
    is_float = any(("." in t) or ("e" in t.lower() and not t.lower().lstrip("+-").startswith("inf")) or
                   ("nan" in t.lower()) for t in tokens)
    band = np.float32 if is_float else np.int32
    raw = np.array([float(t) for t in tokens], dtype=np.float64).reshape(ny, nx)
    a = raw.astype(band).astype(np.float64)
    if "nodata_value" in header:
        nodata = np.array(float(header["nodata_value"])).astype(band).astype(np.float64)
        a[a == nodata] = np.nan
    return a, header


def write_ascii_grid(path, values: np.ndarray, header: dict, dtype=np.float64) -> None:
    """An ESRI ASCII grid with the header of the input, NoData -9999, values written in full precision."""
    out = np.where(np.isnan(values), -9999.0, values).astype(dtype)
    ny, nx = out.shape
    keys = [k for k in ("xllcorner", "yllcorner", "xllcenter", "yllcenter", "cellsize", "dx", "dy") if k in header]
    with open(path, "w", encoding="utf-8") as f:
        f.write(f"ncols {nx}\nnrows {ny}\n")
        for k in keys:
            f.write(f"{k} {header[k]}\n")
        f.write("nodata_value -9999\n")
        for row in out:
            f.write(" ".join(repr(float(v)) for v in row) + "\n")


def _float(value: str) -> float:
    # Julia's parse(Float64, ...) takes Inf / -Inf / NaN too
    return float(value.replace("Inf", "inf"))


def _resolve(name: str, base: Path) -> Path:
    # Omniscape reads paths from the working directory, else next to the INI file
    p = Path(name)
    if p.is_absolute() or p.exists():
        return p
    q = base / p
    return q if q.exists() else p


def _read(path: Path, dtype):
    if not path.exists():
        raise FileNotFoundError(f'the file "{path}" does not exist')
    if path.suffix.lower() != ".asc":
        raise ValueError(f"{path}: only ESRI ASCII grids (.asc) are read on this path")
    a, header = read_ascii_grid(path)
    if dtype is np.float32:
        a = a.astype(np.float32).astype(np.float64)
    return a, header


def _aligned(ref, other, name1, name2, allow_projections: bool) -> None:
    (a, ha), (b, hb) = ref, other
    geo = ("xllcorner", "yllcorner", "xllcenter", "yllcenter", "cellsize")
    if any(ha.get(k) != hb.get(k) for k in geo) and not allow_projections:
        warnings.warn(f"{name1} and {name2} have different projections. This could mean they are not properly "
                      "aligned. Attempting to proceed anyway.", stacklevel=3)
    if a.shape != b.shape:
        raise ValueError(f"{name1} and {name2} are different sizes.")


def _project_dir(name: str) -> Path:
    base = Path(name)
    if not base.is_dir():
        base.mkdir(parents=True, exist_ok=True)
        return base
    k = 1
    while Path(f"{name}_{k}").is_dir():
        k += 1
    out = Path(f"{name}_{k}")
    warnings.warn(f"Your specified project directory, {name}, already exists. Writing outputs to {out}.",
                  stacklevel=3)
    out.mkdir(parents=True)
    return out


def omniscape_ini(path, *, write_outputs: bool = True, n_threads: int = 0):
    """Run an Omniscape INI file as ``run_omniscape(path)`` does (the reference's ini.py:142-287).

    Every option of Omniscape 0.7 is read with its defaults (``radius``,
    ``resistance_file`` and ``project_name`` are required); paths are taken
    from the working directory, else next to the INI file. With
    ``write_outputs`` the outputs go to ``project_name`` (``_1``, ``_2``...
    when it exists): ``config.ini``, ``cum_currmap``, ``flow_potential``,
    ``normalized_cum_currmap``, ``classified_resistance`` as ASCII grids,
    NoData -9999. ``solver``, ``parallelize``, ``parallel_batch_size`` and
    ``suppress_cs_messages`` have no effect.
    """
    path = Path(path)
    base = path.resolve().parent
    user = parse_ini(path)
    missing = [k for k in REQUIRED if k not in user]
    if missing:
        raise ValueError("The following arguments are missing from your .ini with no defaults: " + " ".join(missing))
    unknown = [k for k in user if k not in SUPPORTED and k != ""]
    if unknown:
        warnings.warn("The following unsupported arguments were provided and will be ignored: " + " ".join(unknown),
                      stacklevel=2)
    cfg = {**DEFAULTS, **user}
    flag = {k: cfg[k] in TRUELIST for k in cfg}
    precision = "single" if cfg["precision"] in ("single", "Single", "SINGLE") else "double"
    dtype = np.float32 if precision == "single" else np.float64
    allow = flag["allow_different_projections"]

    res = _read(_resolve(cfg["resistance_file"], base), dtype)
    resistance = res[0]
    reclass = None
    if flag["reclassify_resistance"]:
        reclass = read_reclass_table(_resolve(cfg["reclass_table"], base))
    elif cfg["reclass_table"] != "":
        warnings.warn("You provided a reclass_table, but did not set reclassify_resistance to true. Resistance will "
                      "not be reclassified.", stacklevel=2)

    source = None
    if not flag["source_from_resistance"]:
        if cfg["source_file"] == "":
            raise ValueError("You did not provide a source raster file path. Set source_from_resistance to true in "
                             "your config if you want to generate source strength from the resistance layer.")
        src = _read(_resolve(cfg["source_file"], base), dtype)
        _aligned(res, src, "resistance_file", "source_file", allow)
        source = src[0]

    conds: dict = {}
    if flag["conditional"]:
        compare = cfg["compare_to_future"].lower()
        n_cond = int(round(_float(cfg["n_conditions"])))

        def cond(key):
            c = _read(_resolve(cfg[key], base), dtype)
            _aligned(res, c, "resistance_file", key, allow)
            return c[0]

        conds["condition1"] = cond("condition1_file")
        if compare in ("1", "both"):
            conds["condition1_future"] = cond("condition1_future_file")
        if n_cond == 2:
            conds["condition2"] = cond("condition2_file")
            if compare in ("2", "both"):
                conds["condition2_future"] = cond("condition2_future_file")
        elif n_cond != 1:
            raise ValueError(f"n_conditions must be one of ['1', '2']. Got {cfg['n_conditions']}.")
        conds.update(compare_to_future=compare, comparison1=cfg["comparison1"], comparison2=cfg["comparison2"],
                     condition1_lower=_float(cfg["condition1_lower"]), condition1_upper=_float(cfg["condition1_upper"]),
                     condition2_lower=_float(cfg["condition2_lower"]), condition2_upper=_float(cfg["condition2_upper"]))
    elif cfg["condition1_file"] != "" or cfg["condition2_file"] != "":
        warnings.warn("conditional is set to false in your INI but you provided file paths to condition files.",
                      stacklevel=2)

    with warnings.catch_warnings():
        warnings.simplefilter("ignore", UserWarning)  # block_size: warned below, once
        ds = omniscape(
            resistance, _float(cfg["radius"]), source=source, block_size=_float(cfg["block_size"]),
            buffer=_float(cfg["buffer"]), source_threshold=_float(cfg["source_threshold"]),
            r_cutoff=_float(cfg["r_cutoff"]), resistance_is_conductance=flag["resistance_is_conductance"],
            four_neighbors=flag["connect_four_neighbors_only"], flow_potential=flag["calc_flow_potential"],
            normalized_current=flag["calc_normalized_current"], correct_artifacts=flag["correct_artifacts"],
            mask_nodata=flag["mask_nodata"], reclass_table=reclass, precision=precision, n_threads=n_threads,
            **conds)
    if int(round(_float(cfg["block_size"]))) % 2 == 0:
        warnings.warn("block_size is even, but must be odd. Using block_size + 1.", stacklevel=2)
    if reclass is not None:
        ds["classified_resistance"] = reclassify(resistance, reclass).astype(dtype)

    if write_outputs:
        out = _project_dir(cfg["project_name"])
        with open(out / "config.ini", "a", encoding="utf-8") as f:
            for k, v in user.items():
                if k:
                    f.write(f"{k} = {v}\n")
        header = res[1]

        def put(name):
            write_ascii_grid(out / f"{name}.asc", np.asarray(ds[name]), header, dtype)

        if reclass is not None and flag["write_reclassified_resistance"]:
            put("classified_resistance")
        if flag["write_raw_currmap"]:
            put("cum_currmap")
        if flag["calc_flow_potential"]:
            put("flow_potential")
        if flag["calc_normalized_current"]:
            put("normalized_cum_currmap")
        ds.attrs["project_dir"] = os.fspath(out)
    return ds
