"""Omniscape INI files, conditional connectivity and single precision (the reference's tests/test_omniscape_ini.py).

The INI fixtures are Omniscape's own test configurations (1-7, 4a-4c, 32 bit,
reclass) over its test rasters; ``omniscape_ini.npz`` holds Omniscape 0.7's
outputs for them with the ``cholmod`` and ``cg+amg`` solvers. Each case runs
twice: with ``core()`` answered by the FP64 oracle on the CPU (this pins the
oracle's conditional connectivity to Omniscape's outputs), and on the GPU.
"""

from __future__ import annotations

import contextlib
import importlib
import warnings

import numpy as np
import pytest

from conftest import FIXTURES

from landschaft_amd import circuit
from landschaft_amd.circuit import ini

omni_mod = importlib.import_module("landschaft_amd.circuit.omniscape")
INI = FIXTURES / "omniscape_ini"
INPUT = FIXTURES / "omniscape_input"
OUTPUTS = ("cum_currmap", "flow_potential", "normalized_cum_currmap", "classified_resistance")


class OracleCore:
    """``core().circuit_omniscape`` answered by the oracle."""

    def __init__(self, oracle):
        self.o = oracle

    def circuit_omniscape(self, r, s, radius, buffer, block_size, four, rc, correct, fp, n_threads=0,
                          conditions=None, memory_limit=0.0):
        out = self.o.omniscape_conditions(r, s, radius, buffer, block_size, four, rc, correct, fp,
                                          conditions or {})
        out["threads"] = 1
        return out


@pytest.fixture(params=["oracle", pytest.param("gpu", marks=pytest.mark.gpu)])
def engine(request, monkeypatch, oracle):
    if request.param == "oracle":
        oc = OracleCore(oracle)
        monkeypatch.setattr(omni_mod, "core", lambda: oc)
    else:
        from landschaft_amd.circuit import _backend

        monkeypatch.setattr(_backend.settings, "precision", "mixed")
        monkeypatch.setattr(_backend.settings, "tol", 1e-14)
    return request.param


def _map(seed=1, shape=(30, 40)):
    rng = np.random.default_rng(seed)
    r = np.exp(rng.normal(0.5, 1.0, shape))
    r[rng.random(shape) < 0.05] = np.nan
    return r


def test_parse_ini(tmp_path):
    f = tmp_path / "a.ini"
    f.write_text("[Input files]\nresistance_file = r.tif \nradius=5\nweird = a = b\nno equals here\n[x]\n")
    assert ini.parse_ini(f) == {"resistance_file": "r.tif", "radius": "5", "weird": "a = b"}


def test_ascii_grid_reader(tmp_path):
    f = tmp_path / "g.asc"
    f.write_text("ncols 3\nnrows 2\nxllcorner 0\nyllcorner 0\ncellsize 1\nnodata_value -9999\n"
                 "0.1 2 -9999\n3 4 5\n")
    a, header = ini.read_ascii_grid(f)
    # a float grid is read as float32, as GDAL's AAIGrid driver types it
    np.testing.assert_array_equal(a, np.array([[np.float32(0.1), 2, np.nan], [3, 4, 5]]))
    assert header["ncols"] == "3"
    f.write_text("ncols 2\nnrows 1\nnodata_value -1\n7 -1\n")
    a, _ = ini.read_ascii_grid(f)
    np.testing.assert_array_equal(a, [[7.0, np.nan]])
    out = tmp_path / "o.asc"
    ini.write_ascii_grid(out, np.array([[1.5, np.nan]]), {"cellsize": "1"})
    b, h = ini.read_ascii_grid(out)
    np.testing.assert_array_equal(b, [[1.5, np.nan]])
    assert h["nodata_value"] == "-9999"


def test_fixture_rasters_are_float32():
    # Omniscape's test rasters: 30 x 30 (resistance1: 23 x 30), NoData -9999 (resistance1: -99)
    r, h = ini.read_ascii_grid(INPUT / "resistance1.asc")
    assert r.shape == (23, 30) and h["nodata_value"] == "-99" and np.isnan(r).any()
    np.testing.assert_array_equal(r[~np.isnan(r)], r[~np.isnan(r)].astype(np.float32))
    assert r[0, 0] == np.float64(np.float32(0.29243631))


def _cases():
    z = np.load(FIXTURES / "omniscape_ini.npz")
    return sorted({k.split("/")[0] for k in z.files})


@pytest.mark.parametrize("name", _cases())
def test_matches_omniscape_ini(engine, name, monkeypatch):
    monkeypatch.chdir(INI)
    z = np.load(FIXTURES / "omniscape_ini.npz")
    print(f"ini case {name} on {engine}", flush=True)
    expect = pytest.warns(UserWarning) if name in ("config", "config3", "config5", "config_32bit") \
        else contextlib.nullcontext()
    with expect:
        ds = ini.omniscape_ini(INI / f"{name}.ini", write_outputs=False)
    single = name == "config_32bit"
    checked = 0
    for out in OUTPUTS:
        key = f"{name}/cholmod/{out}"
        if key not in z.files:
            continue
        ours = np.asarray(ds[out]).astype(float)
        chol, cg = z[key], z[f"{name}/cgamg/{out}"]
        np.testing.assert_array_equal(np.isnan(ours), np.isnan(chol))
        np.testing.assert_array_equal(ours == 0, chol == 0)
        top = np.nanmax(np.abs(chol))
        if single:
            # Omniscape computes in float32; here the solve is double and the result is rounded
            np.testing.assert_allclose(ours, chol, rtol=1e-3, atol=1e-3 * top)
            np.testing.assert_allclose(ours, cg, rtol=0, atol=1e-2 * top)
        else:
            np.testing.assert_allclose(ours, chol, rtol=1e-9, atol=1e-13 * top)
            np.testing.assert_allclose(ours, cg, rtol=0, atol=1e-3 * top)
        checked += 1
    assert checked


def test_conditions(engine):
    r = _map(3)
    temp = np.random.default_rng(3).normal(15, 3, r.shape)
    plain = circuit.omniscape(r, 5, block_size=3)
    # a condition every source meets changes nothing
    wide = circuit.omniscape(r, 5, block_size=3, condition1=temp, condition1_lower=-1e9, condition1_upper=1e9)
    np.testing.assert_array_equal(wide["cum_currmap"], plain["cum_currmap"])
    assert wide.attrs["conditional"] == 1
    # one that no source meets: no current at all
    none = circuit.omniscape(r, 5, block_size=3, condition1=temp, condition1_lower=1e8, condition1_upper=1e9)
    assert np.nansum(none["cum_currmap"]) == 0
    # "equal" on a constant layer, and futures equal to the present
    const = np.full(r.shape, 3.0)
    eq = circuit.omniscape(r, 5, block_size=3, condition1=const, comparison1="equal", condition2=temp,
                           condition2_future=temp, compare_to_future="2", condition2_lower=-1e9,
                           condition2_upper=1e9)
    np.testing.assert_array_equal(eq["cum_currmap"], plain["cum_currmap"])
    # a tight window keeps some sources and drops others
    tight = circuit.omniscape(r, 5, block_size=3, condition1=temp, condition1_lower=-1, condition1_upper=1)
    assert 0 < np.nansum(tight["cum_currmap"]) < np.nansum(plain["cum_currmap"])
    with pytest.raises(ValueError, match="comparison1"):
        circuit.omniscape(r, 5, condition1=temp, comparison1="near")
    with pytest.raises(ValueError, match="condition1_future"):
        circuit.omniscape(r, 5, condition1=temp, compare_to_future="1")


@pytest.mark.gpu
@pytest.mark.parametrize("block_size", [1, 3, 5])
def test_conditions_gpu_matches_oracle(oracle, block_size, monkeypatch):
    """All condition forms on a random map: ``within`` against a future layer with NaN cells, ``equal`` on
    a layer of three classes (block modes with ties, NaN skipped), against the oracle's restatement."""
    from landschaft_amd.circuit import _backend

    monkeypatch.setattr(_backend.settings, "precision", "mixed")
    monkeypatch.setattr(_backend.settings, "tol", 1e-14)
    rng = np.random.default_rng(5 + block_size)
    r = np.exp(rng.normal(0.0, 0.7, (41, 53)))
    r[rng.random(r.shape) < 0.06] = np.nan
    s = np.where(np.isnan(r), 0.0, rng.uniform(0.2, 1.0, r.shape))
    s[rng.random(r.shape) < 0.3] = 0.0
    c1 = np.round(rng.normal(0.0, 1.0, r.shape), 1)
    c1f = c1 + 0.3
    c1[rng.random(r.shape) < 0.05] = np.nan  # NaN present values keep their sources
    if block_size > 1:
        c1f[rng.random(r.shape) < 0.05] = np.nan  # skipped in the block medians
    c2 = rng.integers(0, 3, r.shape).astype(float)
    conds = {"condition1": c1, "condition1_future": c1f, "comparison1": "within", "condition1_lower": -0.6,
             "condition1_upper": 0.4, "condition2": c2, "comparison2": "equal"}
    print(f"conditions block {block_size}", flush=True)
    for fp in (False, True):
        got = _backend.circuit_omniscape(r, s, 6, 1, block_size, False, False, True, fp, conditions=conds)
        ref = oracle.omniscape_conditions(r, s, 6, 1, block_size, False, False, True, fp, conds)
        assert got["n_targets"] == ref["n_targets"]
        for k in ("current", "flow_potential") if fp else ("current",):
            top = np.abs(ref[k]).max()
            assert top > 0
            np.testing.assert_array_equal(got[k] == 0, ref[k] == 0)
            np.testing.assert_allclose(got[k], ref[k], rtol=1e-9, atol=1e-12 * top)


def test_condition_all_nodata_block_raises(engine):
    r = np.ones((12, 12))
    c = np.ones((12, 12))
    c[:6, :6] = np.nan  # holds a whole 3 x 3 block
    # a C++ runtime_error in the reference (src/cpp/circuit/omniscape.cpp:178), so RuntimeError
    with pytest.raises(RuntimeError, match="no value of a condition"):
        circuit.omniscape(r, 3, block_size=3, condition1=c)


def test_ini_errors(tmp_path):
    f = tmp_path / "a.ini"
    f.write_text("radius = 3\nproject_name = x\n")
    with pytest.raises(ValueError, match="resistance_file"):
        ini.omniscape_ini(f, write_outputs=False)
    # Omniscape's test configuration 7 has a resistance of 0
    with pytest.raises(ValueError, match="> 0"):
        ini.omniscape_ini(INI / "config7.ini", write_outputs=False)


def test_write_outputs(engine, tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)
    f = tmp_path / "run.ini"
    f.write_text((INI / "config_32bit.ini").read_text().replace("../omniscape_input/", str(INPUT) + "/"))
    with pytest.warns(UserWarning, match="block_size is even"):
        ds = ini.omniscape_ini(f)
    out = tmp_path / "test1_32bit"
    assert ds.attrs["project_dir"] == "test1_32bit"
    assert sorted(p.name for p in out.iterdir()) == ["config.ini", "cum_currmap.asc", "flow_potential.asc",
                                                    "normalized_cum_currmap.asc"]
    v, h = ini.read_ascii_grid(out / "cum_currmap.asc")
    assert h["nodata_value"] == "-9999"
    np.testing.assert_array_equal(v, np.asarray(ds["cum_currmap"]).astype(float))
    with warnings.catch_warnings(record=True) as record:
        warnings.simplefilter("always")
        ds2 = ini.omniscape_ini(f, n_threads=1)
    assert any("already exists" in str(w.message) for w in record)
    assert ds2.attrs["project_dir"] == "test1_32bit_1"
    assert "resistance_file = " in (out / "config.ini").read_text()
