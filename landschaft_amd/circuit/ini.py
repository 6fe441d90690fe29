"""Omniscape's INI files (``run_omniscape(path)``) on the GPU path, and the ESRI ASCII grids they name.

Options, defaults, messages and outputs are Omniscape 0.7's, as the reference's
src/landschaft/circuit/ini.py keeps them (:14-40 options, :41-75 parsing, :142-287 the driver).
A run is an :class:`_Run`: the parsed file over the defaults, with typed accessors, the rasters it
names (checked against the resistance's grid), and the output directory it writes. Rasters are ESRI
ASCII grids (``.asc``) read as GDAL's AAIGrid driver types them (Int32, or Float32 once any value has a
decimal point or an exponent), so the values are those of the reference's rasterio reads; other formats
need rasterio, which this path does not use.
"""

from __future__ import annotations

import os
import warnings
from pathlib import Path

import numpy as np

from .omniscape import omniscape, read_reclass_table, reclassify

TRUELIST = ("true", "True", "TRUE")
REQUIRED = ("resistance_file", "radius", "project_name")
# Omniscape 0.7's optional settings and their defaults (its own names and spellings)
DEFAULTS = dict(
    resistance_file="", resistance_is_conductance="false", source_file="", parallelize="true",
    parallel_batch_size="10", block_size="1", buffer="0", source_threshold="0", source_from_resistance="false",
    r_cutoff="Inf", precision="double", connect_four_neighbors_only="false", solver="cg+amg",
    calc_flow_potential="false", calc_normalized_current="false", correct_artifacts="true",
    write_raw_currmap="true", write_as_tif="true", mask_nodata="true", suppress_cs_messages="true",
    conditional="false", n_conditions="1", compare_to_future="none", condition1_file="", condition2_file="",
    condition1_future_file="", condition2_future_file="", comparison1="within", comparison2="within",
    condition1_lower="0", condition2_lower="0", condition1_upper="0", condition2_upper="0",
    reclassify_resistance="false", reclass_table="", write_reclassified_resistance="false",
    allow_different_projections="false",
)
SUPPORTED = tuple(dict.fromkeys(("project_name", "radius", *DEFAULTS)))
_GRID_KEYS = ("ncols", "nrows", "xllcorner", "yllcorner", "xllcenter", "yllcenter", "cellsize", "dx", "dy",
              "nodata_value")
_PLACEMENT = ("xllcorner", "yllcorner", "xllcenter", "yllcenter", "cellsize")


def parse_ini(path) -> dict:
    """An INI file as Omniscape's ``parse_cfg`` reads it: ``[`` lines skipped, every other line split at
    its first ``=`` (key right-stripped, value stripped)."""
    text = Path(path).read_text(encoding="utf-8").splitlines()
    entries = (line.split("=", 1) for line in text if not line.startswith("[") and "=" in line)
    return {key.rstrip(): value.strip() for key, value in entries}


# ---- ESRI ASCII grids

def _band_type(tokens) -> type:
    """The AAIGrid band type: Float32 when a value has a decimal point, an exponent or a NaN, else Int32."""
    for t in tokens:
        low = t.lower()
        if "." in low or "nan" in low or ("e" in low and not low.lstrip("+-").startswith("inf")):
            return np.float32
    return np.int32


def read_ascii_grid(path):
    """An ESRI ASCII grid as (values float64 with NaN for NoData, header dict).

    The values go through the type GDAL's AAIGrid driver gives the band
    (Int32, or Float32 when any value is written with a decimal point or an
    exponent), then to float64, as a rasterio read does.
    """
    header: dict = {}
    body: list = []
    for line in Path(path).read_text(encoding="utf-8").splitlines():
        words = line.split()
        if not words:
            continue
        if not body and words[0].lower() in _GRID_KEYS:
            header[words[0].lower()] = words[1]
        else:
            body += words
    ny, nx = int(header["nrows"]), int(header["ncols"])
    if len(body) != nx * ny:
        raise ValueError(f"{path}: {len(body)} values for a {ny} x {nx} grid")
    band = _band_type(body)
    values = np.asarray(body, dtype=np.float64).reshape(ny, nx).astype(band).astype(np.float64)
    if "nodata_value" in header:
        marker = np.float64(np.asarray(float(header["nodata_value"])).astype(band))
        values[values == marker] = np.nan
    return values, header


def write_ascii_grid(path, values: np.ndarray, header: dict, dtype=np.float64) -> None:
    """An ESRI ASCII grid with the header of the input, NoData -9999, values written in full precision."""
    grid = np.where(np.isnan(values), -9999.0, values).astype(dtype)
    ny, nx = grid.shape
    head = [f"ncols {nx}", f"nrows {ny}"] + [f"{k} {header[k]}" for k in _GRID_KEYS[2:-1] if k in header]
    head.append("nodata_value -9999")
    rows = (" ".join(map(repr, map(float, row))) for row in grid)
    Path(path).write_text("\n".join([*head, *rows]) + "\n", encoding="utf-8")


# ---- one run of an INI file

def _number(text: str) -> float:
    """A number as Julia's parse(Float64, ...) reads it (Inf, -Inf and NaN included)."""
    return float(text.replace("Inf", "inf"))


class _Run:
    """An Omniscape INI file over the defaults: typed settings and the rasters it names."""

    def __init__(self, path):
        path = Path(path)
        self.base = path.resolve().parent
        self.user = parse_ini(path)
        absent = [k for k in REQUIRED if k not in self.user]
        if absent:
            raise ValueError("The following arguments are missing from your .ini with no defaults: " + " ".join(absent))
        ignored = [k for k in self.user if k and k not in SUPPORTED]
        if ignored:
            warnings.warn("The following unsupported arguments were provided and will be ignored: "
                          + " ".join(ignored), stacklevel=3)
        self.cfg = {**DEFAULTS, **self.user}
        self.precision = "single" if self.cfg["precision"] in ("single", "Single", "SINGLE") else "double"
        self.dtype = np.float32 if self.precision == "single" else np.float64
        self.grid = None  # (values, header) of the resistance

    def flag(self, key: str) -> bool:
        return self.cfg[key] in TRUELIST

    def number(self, key: str) -> float:
        return _number(self.cfg[key])

    def file(self, key: str) -> Path:
        """The path an option names: as given when it exists (or is absolute), else next to the INI."""
        given = Path(self.cfg[key])
        if given.is_absolute() or given.exists():
            return given
        beside = self.base / given
        return beside if beside.exists() else given

    def raster(self, key: str) -> np.ndarray:
        """The grid an option names, in the run's precision; every grid after the resistance must have its
        size (and, unless allowed, its placement)."""
        path = self.file(key)
        if not path.exists():
            raise FileNotFoundError(f'the file "{path}" does not exist')
        if path.suffix.lower() != ".asc":
            raise ValueError(f"{path}: only ESRI ASCII grids (.asc) are read on this path")
        values, header = read_ascii_grid(path)
        if self.dtype is np.float32:
            values = values.astype(np.float32).astype(np.float64)
        if self.grid is None:
            self.grid = (values, header)
            return values
        ref_values, ref_header = self.grid
        moved = any(ref_header.get(k) != header.get(k) for k in _PLACEMENT)
        if moved and not self.flag("allow_different_projections"):
            warnings.warn(f"resistance_file and {key} have different projections. This could mean they are not "
                          "properly aligned. Attempting to proceed anyway.", stacklevel=4)
        if values.shape != ref_values.shape:
            raise ValueError(f"resistance_file and {key} are different sizes.")
        return values

    def conditions(self) -> dict:
        """omniscape()'s condition arguments (empty when the run is not conditional)."""
        if not self.flag("conditional"):
            if self.cfg["condition1_file"] != "" or self.cfg["condition2_file"] != "":
                warnings.warn("conditional is set to false in your INI but you provided file paths to condition "
                              "files.", stacklevel=3)
            return {}
        compare = self.cfg["compare_to_future"].lower()
        count = round(self.number("n_conditions"))
        layers = {"condition1": self.raster("condition1_file")}
        if compare in ("1", "both"):
            layers["condition1_future"] = self.raster("condition1_future_file")
        if count == 2:
            layers["condition2"] = self.raster("condition2_file")
            if compare in ("2", "both"):
                layers["condition2_future"] = self.raster("condition2_future_file")
        elif count != 1:
            raise ValueError(f"n_conditions must be one of ['1', '2']. Got {self.cfg['n_conditions']}.")
        bounds = {k: self.number(k) for k in ("condition1_lower", "condition1_upper", "condition2_lower",
                                              "condition2_upper")}
        return {**layers, **bounds, "compare_to_future": compare, "comparison1": self.cfg["comparison1"],
                "comparison2": self.cfg["comparison2"]}

    def output_dir(self) -> Path:
        """project_name, or project_name_1, _2, ... when it exists (created)."""
        name = self.cfg["project_name"]
        if not Path(name).is_dir():
            Path(name).mkdir(parents=True, exist_ok=True)
            return Path(name)
        k = next(k for k in range(1, 1 << 30) if not Path(f"{name}_{k}").is_dir())
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
    run = _Run(path)
    resistance = run.raster("resistance_file")
    reclass = None
    if run.flag("reclassify_resistance"):
        reclass = read_reclass_table(run.file("reclass_table"))
    elif run.cfg["reclass_table"] != "":
        warnings.warn("You provided a reclass_table, but did not set reclassify_resistance to true. Resistance will "
                      "not be reclassified.", stacklevel=2)
    source = None
    if not run.flag("source_from_resistance"):
        if run.cfg["source_file"] == "":
            raise ValueError("You did not provide a source raster file path. Set source_from_resistance to true in "
                             "your config if you want to generate source strength from the resistance layer.")
        source = run.raster("source_file")
    conditions = run.conditions()
    with warnings.catch_warnings():
        warnings.simplefilter("ignore", UserWarning)  # an even block_size is reported below, once
        ds = omniscape(resistance, run.number("radius"), source=source, block_size=run.number("block_size"),
                       buffer=run.number("buffer"), source_threshold=run.number("source_threshold"),
                       r_cutoff=run.number("r_cutoff"), resistance_is_conductance=run.flag("resistance_is_conductance"),
                       four_neighbors=run.flag("connect_four_neighbors_only"),
                       flow_potential=run.flag("calc_flow_potential"),
                       normalized_current=run.flag("calc_normalized_current"),
                       correct_artifacts=run.flag("correct_artifacts"), mask_nodata=run.flag("mask_nodata"),
                       reclass_table=reclass, precision=run.precision, n_threads=n_threads, **conditions)
    if round(run.number("block_size")) % 2 == 0:
        warnings.warn("block_size is even, but must be odd. Using block_size + 1.", stacklevel=2)
    if reclass is not None:
        ds["classified_resistance"] = reclassify(resistance, reclass).astype(run.dtype)
    if not write_outputs:
        return ds
    out = run.output_dir()
    with open(out / "config.ini", "a", encoding="utf-8") as f:
        f.writelines(f"{k} = {v}\n" for k, v in run.user.items() if k)
    wanted = [("classified_resistance", reclass is not None and run.flag("write_reclassified_resistance")),
              ("cum_currmap", run.flag("write_raw_currmap")), ("flow_potential", run.flag("calc_flow_potential")),
              ("normalized_cum_currmap", run.flag("calc_normalized_current"))]
    for name, on in wanted:
        if on:
            write_ascii_grid(out / f"{name}.asc", np.asarray(ds[name]), run.grid[1], run.dtype)
    ds.attrs["project_dir"] = os.fspath(out)
    return ds
