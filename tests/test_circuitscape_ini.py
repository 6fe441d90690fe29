"""Circuitscape's INI files and file formats (the reference's tests/test_circuitscape.py).

Every INI of Circuitscape's own raster test suite (advanced, pairwise,
one-to-all, all-to-one: 50 files, with polygons, masks, GeoTIFF inputs,
include / exclude pairs, variable source strengths and focal regions) runs
through ``circuitscape_ini`` and is compared with the outputs Circuitscape
wrote for it (``circuitscape.npz``), at the reference's tolerances: on the CPU
with ``core()`` answered by the FP64 oracle, and on the GPU with the HIP
library (mixed precision refined to a relative residual of 1e-14).
"""

from __future__ import annotations

import importlib
from pathlib import Path

import numpy as np
import pytest

from oracle_core import OracleCore

cs = importlib.import_module("landschaft_amd.circuit.circuitscape")

FIX = Path(__file__).resolve().parent / "fixtures" / "circuit"
CS = FIX / "circuitscape"


def _inis(scenario):
    return sorted(p.relative_to(CS).as_posix() for p in (CS / "input" / "raster" / scenario).glob("*/*.ini"))


@pytest.fixture(scope="module")
def outputs():
    return np.load(FIX / "circuitscape.npz")


@pytest.fixture
def in_cs(monkeypatch):
    # the INI files hold paths relative to Circuitscape's test directory
    monkeypatch.chdir(CS)


@pytest.fixture(params=["oracle", pytest.param("gpu", marks=pytest.mark.gpu)])
def engine(request, monkeypatch, oracle):
    if request.param == "oracle":
        oc = OracleCore(oracle)
        monkeypatch.setattr(cs, "core", lambda: oc)
    else:
        from landschaft_amd.circuit import _backend

        monkeypatch.setattr(_backend.settings, "precision", "mixed")
        monkeypatch.setattr(_backend.settings, "tol", 1e-14)
    return request.param


# ---- readers (no solves)

def test_file_types(tmp_path):
    root = CS / "input" / "raster"
    assert cs.file_type(root / "advanced" / "1" / "cellmap5x5.asc") == "aagrid"
    assert cs.file_type(root / "pairwise" / "1" / "polygons.tif") == "geotiff"
    assert cs.file_type(root / "advanced" / "3" / "grounds10x10.tif.gz") == "geotiff"
    assert cs.file_type(root / "pairwise" / "13" / "include_matrix.txt") == "pairs_aagrid"
    f = tmp_path / "pts.txt"
    f.write_text("1 105 145\n2 115 125\n")
    assert cs.file_type(f) == "list"
    f.write_text("mode include\n1 2\n")
    assert cs.file_type(f) == "pairs"


def test_ascii_grids_are_read_in_double(tmp_path):
    f = tmp_path / "a.asc"
    f.write_text("ncols 2\nnrows 2\nxllcenter 0.5\nyllcenter 0.5\ncellsize 1\nNODATA_value -1\n"
                 "0.1 0.2\n-1 0.30000000000000004\n")
    a, meta = cs.read_raster(f)
    np.testing.assert_array_equal(a, [[0.1, 0.2], [-9999.0, 0.30000000000000004]])
    assert (meta.xllcorner, meta.yllcorner, meta.cellsize) == (0.0, 0.0, 1.0)


@pytest.mark.parametrize("stem", ["cellmap10x10", "sources10x10", "grounds10x10", "regions_grid"])
def test_geotiff_matches_its_ascii_twin(stem):
    # Circuitscape ships some inputs as both: the GeoTIFF reader must give the same cells and grid
    root = CS / "input" / "raster" / "advanced" / "3"
    tif = root / f"{stem}.tif"
    if not tif.exists():
        tif = root / f"{stem}.tif.gz"
    a, ma = cs.read_raster(root / f"{stem}.asc")
    b, mb = cs.read_raster(tif)
    np.testing.assert_array_equal(a, b)
    assert (ma.nrows, ma.ncols) == (mb.nrows, mb.ncols)
    np.testing.assert_allclose([mb.xllcorner, mb.yllcorner, mb.cellsize], [ma.xllcorner, ma.yllcorner, ma.cellsize],
                               rtol=1e-12, atol=1e-9)


def test_geotiff_written_reads_back(tmp_path):
    a, meta = cs.read_raster(CS / "input" / "raster" / "pairwise" / "1" / "polygons.tif")
    vals = np.arange(a.size, dtype=np.float64).reshape(a.shape) / 7.0
    vals[0, 0] = np.nan
    cs._write_grid(str(tmp_path / "m"), vals, meta, True)
    b, mb = cs.read_raster(tmp_path / "m.tif")
    want = vals.copy()
    want[0, 0] = -9999.0
    np.testing.assert_array_equal(b, want)
    assert (mb.xllcorner, mb.yllcorner, mb.cellsize) == (meta.xllcorner, meta.yllcorner, meta.cellsize)
    assert mb.crs == meta.crs


def test_point_lists_and_pairs_files():
    root = CS / "input" / "raster" / "pairwise"
    _, meta = cs.read_raster(root / "17" / "raster_331464e2628f.asc")
    pts = cs.read_point_map(root / "17" / "sites.txt", meta)
    assert np.all(np.diff(pts.ids) >= 0) and not pts.has_regions
    assert pts.rows.min() >= 1 and pts.rows.max() <= meta.nrows and pts.cols.max() <= meta.ncols
    pairs = cs.read_included_pairs(root / "17" / "pairs_to_include.txt")
    assert pairs.mode == "include" and pairs.matrix[0, 4] == 1 and pairs.matrix[4, 0] == 1
    m = cs.read_included_pairs(root / "13" / "include_matrix.txt")
    # min 2, max 50: values outside [2, 50] are not included
    assert m.mode == "include" and list(m.ids) == [1, 2, 3, 4, 5, 6, 8, 9]
    assert m.matrix[0, 1] == 1 and m.matrix[0, 3] == 0 and m.matrix[0, 0] == 0


def test_region_polymap():
    cells = np.array([0, 1, 5, 8])
    ids = np.array([1, 1, 2, 3])
    from landschaft_amd.circuit._backend import region_polymap

    # no polygons: the two ids' cells, and nothing else
    np.testing.assert_array_equal(region_polymap(None, 9, cells, ids, 1, 3), [1, 1, 0, 0, 0, 0, 0, 0, 3])
    poly = np.array([0, 4, 4, 0, 0, 0, 0, 7, 0])
    # region 1 (cells 0, 1) overlaps polygon 4: both become a new polygon 8; point 2 is left alone
    np.testing.assert_array_equal(region_polymap(poly, 9, cells, ids, 1, 2), [8, 8, 8, 0, 0, 0, 0, 7, 0])


def test_written_ascii_grid_reads_back(tmp_path):
    a, meta = cs.read_raster(CS / "input" / "raster" / "advanced" / "1" / "cellmap5x5.asc")
    vals = np.random.default_rng(3).standard_normal(a.shape) * 10.0 ** np.arange(a.size).reshape(a.shape)
    cs._write_grid(str(tmp_path / "m"), vals, meta, False)
    b, mb = cs.read_raster(tmp_path / "m.asc")
    np.testing.assert_array_equal(b, vals)
    assert (mb.nrows, mb.ncols, mb.xllcorner, mb.yllcorner, mb.cellsize) == \
        (meta.nrows, meta.ncols, meta.xllcorner, meta.yllcorner, meta.cellsize)


def test_ini_errors(tmp_path):
    f = tmp_path / "a.ini"
    f.write_text("scenario = advanced\n")
    with pytest.raises(ValueError, match="habitat_file is not set"):
        cs.circuitscape_ini(f, write_outputs=False)
    f.write_text("scenario = pairwis\nhabitat_file = x\n")
    with pytest.raises(ValueError, match="scenario"):
        cs.circuitscape_ini(f, write_outputs=False)
    f.write_text("scenario = pairwise\ndata_type = network\n")  # a network INI goes to the network driver
    with pytest.raises(ValueError, match="habitat_file is not set"):
        cs.circuitscape_ini(f, write_outputs=False)
    f.write_text("scenario = pairwise\ndata_type = graph\n")
    with pytest.raises(ValueError, match="data_type"):
        cs.circuitscape_ini(f, write_outputs=False)
    f.write_text("[Options]\nscenario = advanced\nuse_polygons = maybe\nhabitat_file = "
                 + (CS / "input/raster/advanced/1/cellmap5x5.asc").as_posix() + "\n")
    with pytest.raises(ValueError, match="use_polygons"):
        cs.circuitscape_ini(f, write_outputs=False)


# ---- runs against Circuitscape's outputs

@pytest.mark.parametrize("ini", _inis("advanced"))
def test_advanced_matches_circuitscape(ini, in_cs, engine, outputs):
    name = Path(ini).stem
    r = cs.circuitscape_ini(ini, write_outputs=False)
    checked = 0
    for key, ours in (("voltmap.asc", r.voltage), ("curmap.asc", r.current), ("result.txt", r.voltage)):
        k = f"{name}/{key}"
        if k not in outputs.files:
            continue
        ref = outputs[k]
        np.testing.assert_array_equal(ours == 0, ref == 0)
        np.testing.assert_allclose(ours, ref, rtol=1e-9, atol=1e-11 * np.abs(ref).max())
        checked += 1
    assert checked >= 2


@pytest.mark.parametrize("ini", _inis("pairwise"))
def test_pairwise_matches_circuitscape(ini, in_cs, engine, outputs):
    name = Path(ini).stem
    r = cs.circuitscape_ini(ini, write_outputs=False)
    # sgVerify17: resistances up to 3e3, where Circuitscape's own solve is 1e-9 off
    np.testing.assert_allclose(r.resistances, outputs[f"{name}/result.txt"], rtol=1e-8, atol=1e-12)
    for key, ours in (("cum_curmap.asc", r.cum_current), ("max_curmap.asc", r.max_current)):
        if f"{name}/{key}" in outputs.files:
            want = outputs[f"{name}/{key}"]
            np.testing.assert_allclose(ours, want, rtol=1e-9, atol=1e-11 * np.abs(want).max())
    theirs = {k.split("/")[1][:-4] for k in outputs.files if k.startswith(name + "/")
              and k.split("/")[1].startswith(("curmap_", "voltmap_"))}
    assert set(r.maps) == theirs
    for kind in theirs:
        want = outputs[f"{name}/{kind}.asc"]
        np.testing.assert_allclose(r.maps[kind], want, rtol=1e-9, atol=1e-11 * max(np.abs(want).max(), 1e-300))


@pytest.mark.parametrize("ini", _inis("one_to_all") + _inis("all_to_one"))
def test_one_to_all_matches_circuitscape(ini, in_cs, engine, outputs):
    name = Path(ini).stem
    r = cs.circuitscape_ini(ini, write_outputs=False)
    np.testing.assert_allclose(r.resistances, outputs[f"{name}/result.txt"], rtol=1e-11, atol=1e-13)
    for key, ours in (("cum_curmap.asc", r.cum_current), ("max_curmap.asc", r.max_current)):
        if f"{name}/{key}" in outputs.files:
            want = outputs[f"{name}/{key}"]
            np.testing.assert_allclose(ours, want, rtol=1e-11, atol=1e-12 * np.abs(want).max())
    theirs = {k.split("/")[1][:-4] for k in outputs.files if k.startswith(name + "/")
              and k.split("/")[1].startswith(("curmap_", "voltmap_"))}
    assert set(r.maps) == theirs
    for kind in theirs:
        want = outputs[f"{name}/{kind}.asc"]
        np.testing.assert_allclose(r.maps[kind], want, rtol=1e-11, atol=1e-12 * max(np.abs(want).max(), 1e-300))


# ---- the INI driver against the array API, and the files it writes

def test_advanced_arrays_match_the_ini(in_cs, engine):
    a = cs.circuitscape_ini("input/raster/advanced/1/mgVerify1.ini", write_outputs=False)
    root = Path("input/raster/advanced/1")
    b = cs.circuitscape_advanced(cs.read_raster(root / "cellmap5x5.asc")[0], cs.read_raster(root / "sources5x5.asc")[0],
                                 cs.read_raster(root / "grounds5x5.asc")[0],
                                 polygons=cs.read_raster(root / "polygons5x5.asc")[0], four_neighbors=True,
                                 avg_resistances=True)
    np.testing.assert_array_equal(a.voltage, b.voltage)
    np.testing.assert_array_equal(a.current, b.current)


def test_writes_outputs(tmp_path, in_cs, engine):
    ini = tmp_path / "run.ini"
    text = (CS / "input/raster/advanced/1/mgVerify1.ini").read_text()
    ini.write_text(text.replace("output/mgVerify1.out", (tmp_path / "out" / "mg1.out").as_posix()))
    r = cs.circuitscape_ini(ini)
    v, _ = cs.read_raster(tmp_path / "out" / "mg1_voltmap.asc")
    c, _ = cs.read_raster(tmp_path / "out" / "mg1_curmap.asc")
    np.testing.assert_array_equal(v, r.voltage)
    np.testing.assert_array_equal(c, r.current)
    assert "scenario = advanced" in (tmp_path / "out" / "mg1.out").read_text()


def test_pairwise_shortcut_matches_full_solves(in_cs, engine):
    cfg = cs.parse_config("input/raster/pairwise/1/sgVerify1.ini")
    cells, meta = cs.read_raster(cfg["habitat_file"])
    points = cs.read_point_map(cfg["point_file"], meta)
    poly = cs.read_raster(cfg["polygon_file"])[0]
    kw = dict(polygons=poly, four_neighbors=cfg["connect_four_neighbors_only"] == "True",
              avg_resistances=cfg["connect_using_avg_resistances"] == "True")
    short = cs.circuitscape_pairwise(cells, points, **kw)
    full = cs.circuitscape_pairwise(cells, points, cum_map=True, n_threads=1, **kw)
    np.testing.assert_allclose(full.resistances, short.resistances, rtol=1e-11, atol=1e-13)
    again = cs.circuitscape_pairwise(cells, points, cum_map=True, **kw)
    np.testing.assert_array_equal(again.resistances, full.resistances)
    np.testing.assert_array_equal(again.cum_current, full.cum_current)


def test_pairwise_writes_outputs(tmp_path, in_cs, engine):
    ini = tmp_path / "run.ini"
    text = (CS / "input/raster/pairwise/1/sgVerify1.ini").read_text()
    ini.write_text(text.replace("output/sgVerify1.out", (tmp_path / "pw.out").as_posix()))
    r = cs.circuitscape_ini(ini)
    np.testing.assert_array_equal(np.loadtxt(tmp_path / "pw_resistances.out"), r.resistances)
    three = np.loadtxt(tmp_path / "pw_resistances_3columns.out")
    n = len(r.resistances) - 1
    assert three.shape == (n * (n - 1) // 2, 3)
    np.testing.assert_array_equal(cs.read_raster(tmp_path / "pw_cum_curmap.asc")[0], r.cum_current)
    assert (tmp_path / "pw_curmap_1_2.asc").exists()


def test_pairwise_regions_api_matches_the_ini(in_cs, engine):
    # focal regions through the array API: the same per-pair graphs as the INI run
    ini = "input/raster/pairwise/9/sgVerify9.ini"
    cfg = cs.parse_config(ini)
    a = cs.circuitscape_ini(ini, write_outputs=False)
    cells, meta = cs.read_raster(cfg["habitat_file"])
    points = cs.read_point_map(cfg["point_file"], meta)
    assert points.has_regions
    b = cs.circuitscape_pairwise(cells, points, polygons=cs.read_raster(cfg["polygon_file"])[0],
                                 habitat_is_resistance=cfg["habitat_map_is_resistances"] == "True",
                                 four_neighbors=cfg["connect_four_neighbors_only"] == "True",
                                 avg_resistances=cfg["connect_using_avg_resistances"] == "True", current_maps=True)
    np.testing.assert_array_equal(a.resistances, b.resistances)
    np.testing.assert_array_equal(a.cum_current, b.cum_current)
    for (kind, i, j), m in b.maps.items():
        np.testing.assert_array_equal(m, a.maps[f"curmap_{i}_{j}"])


def test_one_to_all_api_matches_the_ini(in_cs, engine):
    ini = "input/raster/one_to_all/1/oneToAllVerify1.ini"
    cfg = cs.parse_config(ini)
    a = cs.circuitscape_ini(ini, write_outputs=False)
    cells, meta = cs.read_raster(cfg["habitat_file"])
    b = cs.circuitscape_one_to_all(cells, cs.read_point_map(cfg["point_file"], meta),
                                   polygons=cs.read_raster(cfg["polygon_file"])[0],
                                   four_neighbors=cfg["connect_four_neighbors_only"] == "True",
                                   avg_resistances=cfg["connect_using_avg_resistances"] == "True",
                                   current_maps=True, n_threads=1)
    np.testing.assert_array_equal(a.resistances, b.resistances)
    np.testing.assert_array_equal(a.cum_current, b.cum_current)
    assert {f"curmap_{n}" for kind, n in b.maps if kind == "current"} == {k for k in a.maps if k.startswith("curmap_")}
    for kind, n in b.maps:
        np.testing.assert_array_equal(b.maps[(kind, n)], a.maps[f"curmap_{n}"])


def test_one_to_all_writes_outputs(tmp_path, in_cs, engine):
    ini = tmp_path / "run.ini"
    src = CS / "input/raster/one_to_all/2/oneToAllVerify2.ini"
    out = cs.parse_config(src)["output_file"]
    ini.write_text(src.read_text().replace(out, (tmp_path / "ota.out").as_posix()))
    r = cs.circuitscape_ini(ini)
    np.testing.assert_array_equal(cs.read_raster(tmp_path / "ota_cum_curmap.asc")[0], r.cum_current)
    assert (tmp_path / "ota_curmap_1.asc").exists() and (tmp_path / "ota_voltmap_1.asc").exists()


@pytest.mark.parametrize("seed", range(12))
def test_one_to_all_problems_from_focal_cells_match_full_rasters(seed, monkeypatch, oracle):
    # the problems built from the focal cells alone against the reference's full-raster construction:
    # repeated cells, focal regions, variable strengths, include and exclude pairs, both modes
    oc = OracleCore(oracle)
    monkeypatch.setattr(cs, "core", lambda: oc)
    rng = np.random.default_rng(seed)
    ny, nx = 14, 17
    g = np.exp(rng.standard_normal((ny, nx)))
    npts = 9
    rows = rng.integers(1, ny + 1, npts)
    cols = rng.integers(1, nx + 1, npts)
    ids = np.sort(rng.integers(1, 6, npts))
    if seed % 2:
        rows[3], cols[3] = rows[2], cols[2]  # a cell listed twice
    points = cs.FocalPoints(rows, cols, ids)
    uids = np.unique(ids)
    strengths = np.column_stack([np.arange(1, npts + 1), rng.uniform(0.5, 2.0, npts)]) if seed % 3 == 0 else None
    pairs = None
    if seed >= 2:
        m = (rng.random((len(uids), len(uids))) < 0.6).astype(np.int64)
        m = np.maximum(m, m.T)
        pairs = cs.IncludedPairs("include" if seed % 2 == 0 else "exclude", uids, m)
    for all_to_one in (False, True):
        out = {}
        for dense in (True, False):
            monkeypatch.setattr(cs, "DENSE_PROBLEMS", dense)
            try:
                r = cs.circuitscape_one_to_all(g, points, all_to_one=all_to_one, habitat_is_resistance=False,
                                               strengths=strengths, included_pairs=pairs, cum_map=True,
                                               current_maps=True, voltage_maps=True)
                out[dense] = ("ok", r)
            except Exception as e:  # the two constructions must fail alike too
                out[dense] = ("error", type(e).__name__)
        assert out[True][0] == out[False][0], (seed, all_to_one, out[True], out[False])
        if out[True][0] == "error":
            assert out[True][1] == out[False][1]
            continue
        a, b = out[True][1], out[False][1]
        np.testing.assert_array_equal(a.resistances, b.resistances)
        np.testing.assert_array_equal(a.cum_current, b.cum_current)
        assert set(a.maps) == set(b.maps)
        for k in a.maps:
            np.testing.assert_array_equal(a.maps[k], b.maps[k])
