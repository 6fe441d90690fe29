"""The Python circuit API (the reference's circuit/advanced.py and circuitscape.py) on its docstring examples.

Each example runs twice: on the CPU with ``core()`` answered by the FP64
oracle (this pins the Python layer and the oracle's modes to the reference's
published numbers), and on the GPU with the HIP library behind ``core()``.
"""

from __future__ import annotations

import importlib

import numpy as np
import pytest

from landschaft_amd import circuit
from oracle_core import OracleCore

adv_mod = importlib.import_module("landschaft_amd.circuit.advanced")
cs_mod = importlib.import_module("landschaft_amd.circuit.circuitscape")


@pytest.fixture(params=["oracle", pytest.param("gpu", marks=pytest.mark.gpu)])
def api(request, monkeypatch, oracle):
    if request.param == "oracle":
        oc = OracleCore(oracle)
        monkeypatch.setattr(adv_mod, "core", lambda: oc)
        monkeypatch.setattr(cs_mod, "core", lambda: oc)
    else:
        from landschaft_amd.circuit import _backend

        monkeypatch.setattr(_backend.settings, "precision", "mixed")
        monkeypatch.setattr(_backend.settings, "tol", 1e-12)
    return circuit


def _column_map():
    r = np.ones((10, 10))
    r[:, 5] = 10
    pts = np.zeros((10, 10), dtype=int)
    pts[1, 1], pts[8, 1], pts[5, 8] = 1, 2, 3
    return r, pts


def test_advanced_docstring(api):
    # advanced.py:106-123
    g = np.ones((3, 5))
    s = np.zeros((3, 5))
    s[1, 0] = 1.0
    gr = np.zeros((3, 5))
    gr[1, 4] = np.inf
    r = api.advanced(g, s, gr)
    np.testing.assert_array_equal(r.voltage.round(3), [[0.716, 0.615, 0.444, 0.272, 0.172],
                                                       [0.887, 0.615, 0.444, 0.272, 0.0],
                                                       [0.716, 0.615, 0.444, 0.272, 0.172]])
    np.testing.assert_array_equal(r.current.round(3), [[0.172, 0.293, 0.293, 0.293, 0.172],
                                                       [1.0, 0.414, 0.414, 0.414, 1.0],
                                                       [0.172, 0.293, 0.293, 0.293, 0.172]])


def test_circuitscape_advanced_docstring(api):
    # circuitscape.py:470-480
    res = np.ones((3, 5))
    src = np.full((3, 5), np.nan)
    src[1, 0] = 1.0
    gnd = np.full((3, 5), np.nan)
    gnd[1, 4] = 0.0
    r = api.circuitscape_advanced(res, src, gnd)
    np.testing.assert_array_equal(r.voltage[1].round(3), [0.887, 0.615, 0.444, 0.272, 0.0])
    assert r.maps["solved"]


def test_pairwise_docstring(api):
    # circuitscape.py:621-636
    r, pts = _column_map()
    out = api.circuitscape_pairwise(r, pts, cum_map=True)
    np.testing.assert_array_equal(out.resistances.round(4), [[0.0, 1.0, 2.0, 3.0],
                                                             [1.0, 0.0, 0.8515, 0.8567],
                                                             [2.0, 0.8515, 0.0, 0.8409],
                                                             [3.0, 0.8567, 0.8409, 0.0]])
    assert out.cum_current.shape == (10, 10)
    # the cum map is the sum of the three pair maps
    maps = api.circuitscape_pairwise(r, pts, current_maps=True)
    total = sum(v for k, v in maps.maps.items() if k[0] == "current")
    np.testing.assert_allclose(out.cum_current, total, rtol=1e-12, atol=1e-15)


def test_included_pairs_docstring(api):
    # circuitscape.py:248-259
    r, pts = _column_map()
    pairs = api.IncludedPairs("include", np.array([1, 2, 3]), np.array([[1, 1, 1], [1, 1, 0], [1, 0, 1]]))
    assert sorted(pairs.excluded()) == [(2, 3), (3, 2)]
    out = api.circuitscape_pairwise(r, pts, included_pairs=pairs)
    assert float(out.resistances[2, 3]) == -1.0
    np.testing.assert_allclose(out.resistances[1, 2:].round(4), [0.8515, 0.8567])


def test_one_to_all_docstring(api):
    # circuitscape.py:919-929
    r, pts = _column_map()
    out = api.circuitscape_one_to_all(r, pts, cum_map=True)
    np.testing.assert_array_equal(out.resistances.round(4), [[1.0, 0.6438], [2.0, 0.632], [3.0, 0.6359]])
    assert out.cum_current.shape == (10, 10) and np.all(out.cum_current >= 0)


def test_all_to_one_kirchhoff(api):
    # all-to-one: 1 A at each of the two other points flows into the grounded one
    r, pts = _column_map()
    out = api.circuitscape_one_to_all(r, pts, all_to_one=True, current_maps=True, voltage_maps=True)
    np.testing.assert_array_equal(out.resistances, [[1.0, 0.0], [2.0, 0.0], [3.0, 0.0]])
    for pid, cell in ((1, (1, 1)), (2, (8, 1)), (3, (5, 8))):
        assert out.maps[("voltage", pid)][cell] == 0.0
        np.testing.assert_allclose(out.maps[("current", pid)][cell], 2.0, rtol=1e-9)


def test_one_to_all_maps_and_sums_agree(api):
    # the per-point maps path and the reduced (cum / max only) path of run_onetoall
    r, pts = _column_map()
    r = r * np.exp(np.random.default_rng(3).normal(0, 0.5, r.shape))
    for all_to_one in (False, True):
        a = api.circuitscape_one_to_all(r, pts, all_to_one=all_to_one, cum_map=True, max_map=True,
                                        set_focal_node_currents_to_zero=True)
        b = api.circuitscape_one_to_all(r, pts, all_to_one=all_to_one, current_maps=True, max_map=True,
                                        set_focal_node_currents_to_zero=True)
        np.testing.assert_array_equal(a.resistances, b.resistances)
        np.testing.assert_allclose(a.cum_current, b.cum_current, rtol=1e-13, atol=1e-15)
        np.testing.assert_allclose(a.max_current, b.max_current, rtol=1e-13, atol=1e-15)
        total = sum(v for k, v in b.maps.items() if k[0] == "current")
        np.testing.assert_allclose(b.cum_current, total, rtol=1e-13, atol=1e-15)
