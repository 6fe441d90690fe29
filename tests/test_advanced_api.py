"""Raster circuits on the GPU (landschaft_amd.circuit.advanced): the closed forms and the dense-solve
check the reference runs on its own advanced mode (tests/test_circuit.py:16-127), restated on this
package's API, in mixed precision and in FP64."""

from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(params=["mixed", "fp64"])
def api(request):
    from landschaft_amd import circuit
    from landschaft_amd.circuit import _backend

    saved = (_backend.settings.precision, _backend.settings.tol)
    _backend.settings.precision, _backend.settings.tol = request.param, 1e-14
    yield circuit
    _backend.settings.precision, _backend.settings.tol = saved


def _laplacian(g, four):
    """The grid Laplacian of the cells with g > 0, numbered column by column, and the node list."""
    ny, nx = g.shape
    nodes = [(y, x) for x in range(nx) for y in range(ny) if g[y, x] > 0]
    index = {c: k for k, c in enumerate(nodes)}
    L = np.zeros((len(nodes), len(nodes)))
    steps = [(0, 1), (1, 0), (-1, 0), (0, -1)] + ([] if four else [(1, 1), (-1, -1), (1, -1), (-1, 1)])
    for (y, x), k in index.items():
        for dy, dx in steps:
            j = index.get((y + dy, x + dx))
            if j is not None:
                w = (g[y, x] + g[y + dy, x + dx]) / (2 * np.sqrt(2) if dy and dx else 2)
                L[k, j] -= w
                L[k, k] += w
    return nodes, L


def test_two_cells(api):
    r = api.advanced(np.array([[1.0, 1.0]]), np.array([[1.0, 0.0]]), np.array([[0.0, np.inf]]))
    np.testing.assert_allclose(r.current, [[1.0, 1.0]], rtol=1e-13)
    np.testing.assert_allclose(r.voltage, [[1.0, 0.0]], rtol=1e-13, atol=1e-15)


def test_series_chain_and_diagonal_branch(api):
    # 2 A through branches of (1 + 3) / 2 = 2 and (3 + 2) / 2 = 2.5
    r = api.advanced(np.array([[1.0, 3.0, 2.0]]), np.array([[2.0, 0, 0]]), np.array([[0, 0, np.inf]]))
    np.testing.assert_allclose(r.voltage, [[2 / 2.5 + 2 / 2, 2 / 2.5, 0]], rtol=1e-13, atol=1e-15)
    # two cells joined by a diagonal only: (1 + 1) / (2 sqrt 2)
    g = np.array([[1.0, np.nan], [np.nan, 1.0]])
    src = np.array([[1.0, 0], [0, 0]])
    gnd = np.array([[0, 0], [0, np.inf]])
    r = api.advanced(g, src, gnd)
    np.testing.assert_allclose(r.voltage[0, 0], np.sqrt(2), rtol=1e-13)
    assert np.isnan(r.voltage[0, 1]) and r.current[0, 1] == 0
    # with four neighbours the two cells are separate components: no current
    r = api.advanced(g, src, gnd, four_neighbors=True)
    np.testing.assert_array_equal(r.current, 0)


@pytest.mark.parametrize("four", [False, True])
def test_dense_solve_and_node_currents(api, four):
    rng = np.random.default_rng(3)
    g = rng.uniform(0.1, 2, (9, 11))
    s = np.where(rng.random(g.shape) < 0.3, rng.uniform(0, 1, g.shape), 0.0)
    gnd = np.zeros(g.shape)
    gnd[4, 5] = np.inf
    r = api.advanced(g, s, gnd, four_neighbors=four)
    nodes, L = _laplacian(g, four)
    ground = nodes.index((4, 5))
    keep = [k for k in range(len(nodes)) if k != ground]
    b = np.array([s[c] for c in nodes])
    v = np.zeros(len(nodes))
    v[keep] = np.linalg.solve(L[np.ix_(keep, keep)], b[keep])
    np.testing.assert_allclose([r.voltage[c] for c in nodes], v, rtol=1e-11, atol=1e-14)
    # a node's current: the larger of its inflow and outflow
    flow = -L * (v[:, None] - v[None, :])
    np.fill_diagonal(flow, 0)
    pos = np.clip(flow, 0, None)
    np.testing.assert_allclose([r.current[c] for c in nodes], np.maximum(pos.sum(axis=0), pos.sum(axis=1)),
                               rtol=1e-9, atol=1e-12)
    # Kirchhoff: the ground takes every source's current (up to the 1e-8 branch threshold)
    np.testing.assert_allclose(r.current[4, 5], s.sum() - s[4, 5], rtol=1e-7)


def test_components_and_a_source_on_a_ground(api):
    g = np.ones((3, 7))
    g[:, 3] = np.nan
    s = np.zeros(g.shape)
    gnd = np.zeros(g.shape)
    s[0, 0] = 1.0
    gnd[2, 2] = np.inf
    s[2, 2] = 5.0  # a source on a direct ground is dropped
    s[1, 5] = 1.0  # the right-hand component has no ground
    r = api.advanced(g, s, gnd)
    np.testing.assert_allclose(r.current[2, 2], 1.0, rtol=1e-12)
    np.testing.assert_array_equal(r.current[:, 4:], 0)
    np.testing.assert_array_equal(r.voltage[:, 4:], 0)
    assert np.isnan(r.voltage[:, 3]).all()
    np.testing.assert_array_equal(r.current[:, 3], 0)


def test_finite_grounds_share_the_current(api):
    r = api.advanced(np.ones((1, 3)), np.array([[0.0, 1.0, 0.0]]), np.array([[1.0, 0.0, 1.0]]))
    np.testing.assert_allclose(r.current, [[0.5, 1.0, 0.5]], rtol=1e-13)


def test_host_threads_do_not_change_the_result(api):
    rng = np.random.default_rng(5)
    g = rng.uniform(0.1, 1, (30, 40))
    g[:, 20] = np.nan
    s = rng.uniform(0, 1, g.shape)
    gnd = np.zeros(g.shape)
    gnd[5, 5] = gnd[25, 30] = np.inf
    r1 = api.advanced(g, s, gnd, n_threads=1)
    r2 = api.advanced(g, s, gnd)
    np.testing.assert_array_equal(r1.current, r2.current)
    np.testing.assert_array_equal(r1.voltage, r2.voltage)


def test_bad_inputs(api):
    with pytest.raises(ValueError, match="same shape"):
        api.advanced(np.ones((2, 2)), np.ones((2, 3)), np.ones((2, 2)))
    with pytest.raises(ValueError, match="negative"):
        api.advanced(np.ones((2, 2)), np.ones((2, 2)), -np.ones((2, 2)))
