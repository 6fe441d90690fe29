"""Circuitscape's raster advanced mode through the package's API: the closed forms, the polygon merge
and the large-factorisation check that the reference runs on its own front end
(tests/test_circuitscape.py:48-78, :234-264), on the oracle and on the GPU."""

from __future__ import annotations

import importlib

import numpy as np
import pytest

from oracle_core import OracleCore

cs = importlib.import_module("landschaft_amd.circuit.circuitscape")


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


def test_conflict_policies(engine):
    g = np.ones((1, 3))
    src = np.array([[1.0, 0.0, 2.0]])
    gnd = np.array([[1.0, np.nan, 1.0]])  # ground resistances of 1 at both ends
    kw = dict(habitat_is_resistance=False)
    keep = cs.circuitscape_advanced(g, src, gnd, **kw)
    # both ends are sources and grounds: 3 A leave through two grounds of 1 S, 1 A crosses the chain
    np.testing.assert_allclose(keep.voltage, [[1.5 - 0.25, 1.5, 1.5 + 0.25]], rtol=1e-13)
    rmvsrc = cs.circuitscape_advanced(g, src, gnd, remove_src_or_gnd="rmvsrc", **kw)
    assert not rmvsrc.maps["solved"] and np.all(rmvsrc.voltage == 0)
    rmvgnd = cs.circuitscape_advanced(g, src, gnd, remove_src_or_gnd="rmvgnd", **kw)
    assert not rmvgnd.maps["solved"]
    with pytest.raises(ValueError, match="remove_src_or_gnd"):
        cs.circuitscape_advanced(g, src, gnd, remove_src_or_gnd="none", **kw)


def test_polygon_cells_are_one_node(engine):
    g = np.ones((3, 5))
    g[1, 1:4] = np.nan  # a gap that the polygon bridges
    poly = np.zeros((3, 5))
    poly[0, 1] = poly[2, 3] = 7
    src = np.zeros((3, 5))
    src[0, 0] = 1.0
    gnd = np.full((3, 5), np.nan)
    gnd[2, 4] = 0.0  # resistance 0: a direct ground
    r = cs.circuitscape_advanced(g, src, gnd, polygons=poly)
    assert r.nodemap[0, 1] == r.nodemap[2, 3]
    assert r.voltage[0, 1] == r.voltage[2, 3] and r.current[0, 1] == r.current[2, 3]
    assert r.voltage[2, 4] == 0 and r.current[2, 4] == pytest.approx(1.0, rel=1e-12)


def test_large_factorisation_solves_the_laplacian(engine):
    """400 x 400 cells (an L large enough for the factorisation by levels), host threads 1 and 8 give
    the same bits, and the voltages solve the grounded Laplacian."""
    import scipy.sparse as sp

    rng = np.random.default_rng(12)
    g = rng.uniform(0.2, 5.0, (400, 400))
    src = np.full(g.shape, np.nan)
    src[3, 5], src[200, 300] = 1.0, 2.0
    gnd = np.full(g.shape, np.nan)
    gnd[399, 399] = 0.0
    kw = dict(habitat_is_resistance=False, four_neighbors=True)
    one = cs.circuitscape_advanced(g, src, gnd, n_threads=1, **kw)
    many = cs.circuitscape_advanced(g, src, gnd, n_threads=8, **kw)
    np.testing.assert_array_equal(one.voltage, many.voltage)
    np.testing.assert_array_equal(one.current, many.current)
    ny, nx = g.shape
    idx = np.arange(ny * nx).reshape(ny, nx)
    rows, cols, vals = [], [], []
    for a, b in ((idx[:, :-1], idx[:, 1:]), (idx[:-1, :], idx[1:, :])):
        w = (g.ravel()[a.ravel()] + g.ravel()[b.ravel()]) / 2
        rows += [a.ravel(), b.ravel()]
        cols += [b.ravel(), a.ravel()]
        vals += [w, w]
    A = sp.coo_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                      shape=(ny * nx,) * 2).tocsr()
    L = sp.diags(np.asarray(A.sum(axis=1)).ravel()) - A
    res = (L @ one.voltage.ravel() - np.nan_to_num(src).ravel())[:-1]
    assert np.abs(res).max() < 1e-9
