"""The symbolic analysis on the host (lsc_circuitscape_analyze / lsc_network_analyze; no GPU).

The graph builders, component labels, nested dissections and plans run on up to 16 host threads. Their
outputs must not depend on the thread count: the digests of the CSR (node map, adjacency, weights,
components) and of every plan (order, fronts, update rows, child maps, assembled entries) are compared
between 1 thread and several, on a 512 x 512 raster with NoData holes and short-circuit regions and on
an edge list of 2^17 nodes. The figures are pinned independently: node and component counts against
scipy's connected components, adjacency entries against a count of the raster's neighbour pairs, and
the largest front of a full square grid against the size of its top separators.
"""

from __future__ import annotations


import numpy as np
import pytest
import scipy.sparse as sp
from scipy.sparse.csgraph import connected_components

from landschaft_amd import _native as N


@pytest.fixture
def threads(monkeypatch):
    """Run a callable with LSC_HOST_THREADS set (read at every call)."""

    def run(k, fn, *args, **kw):
        monkeypatch.setenv("LSC_HOST_THREADS", str(k))
        try:
            return fn(*args, **kw)
        finally:
            monkeypatch.delenv("LSC_HOST_THREADS")

    return run


def _raster(n=512, seed=3):
    rng = np.random.default_rng(seed)
    g = rng.uniform(0.2, 5.0, size=(n, n))
    g[rng.random((n, n)) < 0.35] = 0.0  # NoData: many islands and holes
    return g


def _components_scipy(g, four):
    """Node count and components of the raster graph (cells with g > 0, joined to their 4 or 8
    neighbours), by scipy."""
    ny, nx = g.shape
    idx = -np.ones(g.size, dtype=np.int64)
    on = (g > 0).ravel()
    idx[on] = np.arange(on.sum())
    rows, cols = [], []
    offs = [(0, 1), (1, 0)] if four else [(0, 1), (1, 0), (1, 1), (1, -1)]
    for dy, dx in offs:
        y0, y1 = 0, ny - dy
        x0, x1 = max(0, -dx), nx - max(0, dx)
        a = (np.arange(y0, y1)[:, None] * nx + np.arange(x0, x1)[None, :]).ravel()
        b = a + dy * nx + dx
        keep = on[a] & on[b]
        rows.append(idx[a[keep]])
        cols.append(idx[b[keep]])
    r, c = np.concatenate(rows), np.concatenate(cols)
    m = int(on.sum())
    A = sp.coo_matrix((np.ones(r.size), (r, c)), shape=(m, m))
    ncomp, _ = connected_components(A, directed=False)
    return m, ncomp, 2 * r.size


@pytest.mark.parametrize("four", [False, True])
def test_raster_analysis_counts(four):
    g = _raster(256)
    a = N.circuitscape_analyze(g, four_neighbors=four)
    m, ncomp, entries = _components_scipy(g, four)
    assert a["n_nodes"] == m
    assert a["n_components"] == ncomp
    assert a["n_entries"] == entries


def test_full_grid_plan_figures():
    """A full 256 x 256 grid (4-neighbour): one component; the root front holds the middle row's
    separator (its 256 cells minus the cut of the middle column)."""
    g = np.ones((256, 256))
    a = N.circuitscape_analyze(g, four_neighbors=True)
    assert a["n_nodes"] == 256 * 256 and a["n_components"] == 1
    assert a["n_entries"] == 4 * 256 * 255
    assert 256 <= a["largest_front"] <= 3 * 256
    assert a["factor_entries"] > 256 * 256 and a["factor_flops"] > 0


def test_raster_analysis_thread_count_identity(threads):
    g = _raster(512)
    poly = np.zeros(g.shape, dtype=np.int64)
    poly[100:104, 200:260] = 7  # short-circuit regions, one across NoData
    poly[300:340, 50:53] = 9
    poly[0:2, 0:512] = 11
    one = threads(1, N.circuitscape_analyze, g)
    many = threads(8, N.circuitscape_analyze, g)
    assert one["n_components"] > 20  # the labels meet many components
    for key in ("graph_digest", "plan_digest", "n_fronts", "factor_entries", "largest_front", "factor_flops"):
        assert one[key] == many[key], key
    p1 = threads(1, N.circuitscape_analyze, g, poly)
    p8 = threads(5, N.circuitscape_analyze, g, poly)
    assert p1["graph_digest"] == p8["graph_digest"] and p1["plan_digest"] == p8["plan_digest"]
    assert p1["n_nodes"] < one["n_nodes"]


def test_full_raster_thread_count_identity(threads):
    """One component of 2^18 nodes: the dissection's helper threads and the subtree tasks of the
    update-row pass run (the raster with holes has smaller components)."""
    g = np.random.default_rng(5).uniform(0.5, 2.0, size=(512, 512))
    one = threads(1, N.circuitscape_analyze, g)
    many = threads(8, N.circuitscape_analyze, g)
    assert one["n_components"] == 1
    assert one["graph_digest"] == many["graph_digest"]
    assert one["plan_digest"] == many["plan_digest"]


def test_digests_follow_values_and_structure():
    """A new conductance changes the graph digest but not the plan's (same structure); a new NoData
    cell changes both."""
    g = np.random.default_rng(6).uniform(0.5, 2.0, size=(128, 128))
    a = N.circuitscape_analyze(g)
    g2 = g.copy()
    g2[64, 64] *= 1.5
    b = N.circuitscape_analyze(g2)
    assert a["graph_digest"] != b["graph_digest"] and a["plan_digest"] == b["plan_digest"]
    g3 = g.copy()
    g3[64, 64] = 0.0
    c = N.circuitscape_analyze(g3)
    assert c["graph_digest"] != a["graph_digest"] and c["plan_digest"] != a["plan_digest"]


def _edge_list(nx=512, ny=256, extra=20000, seed=7):
    """A grid of 2^17 nodes as an edge list, plus random long edges, repeats and self loops."""
    rng = np.random.default_rng(seed)
    ids = np.arange(nx * ny).reshape(ny, nx)
    a = np.concatenate([ids[:, :-1].ravel(), ids[:-1, :].ravel(), rng.integers(0, nx * ny, extra)])
    b = np.concatenate([ids[:, 1:].ravel(), ids[1:, :].ravel(), rng.integers(0, nx * ny, extra)])
    w = rng.uniform(0.5, 2.0, a.size)
    a = np.concatenate([a, a[:100], [5, 6]])
    b = np.concatenate([b, b[:100], [5, 6]])
    w = np.concatenate([w, w[:100], [1.0, 1.0]])
    return a.astype(np.int32), b.astype(np.int32), w, nx * ny


def test_network_analysis_thread_count_identity(threads):
    a, b, w, m = _edge_list()
    assert m >= 1 << 17
    one = threads(1, N.network_analyze, a, b, w, m)
    many = threads(8, N.network_analyze, a, b, w, m)
    assert one["graph_digest"] == many["graph_digest"]
    assert one["plan_digest"] == many["plan_digest"]
    # counts against scipy: components of the symmetric pattern without self loops
    keep = a != b
    A = sp.coo_matrix((np.ones(keep.sum()), (a[keep], b[keep])), shape=(m, m)).tocsr()
    A = ((A + A.T) > 0).astype(np.int8)
    ncomp, _ = connected_components(A, directed=False)
    assert one["n_nodes"] == m and one["n_components"] == ncomp == 1
    assert one["n_entries"] == A.nnz


def test_network_components_with_isolated_nodes(threads):
    """Two paths, one edge and isolated nodes: a forest of 12 nodes and 5 edges has 7 components."""
    a = np.array([4, 5, 0, 1, 9], dtype=np.int32)
    b = np.array([5, 6, 1, 2, 8], dtype=np.int32)
    r = threads(3, N.network_analyze, a, b, np.ones(5), 12)
    assert r["n_nodes"] == 12 and r["n_components"] == 7 and r["n_entries"] == 10


def test_analysis_argument_errors():
    with pytest.raises(ValueError):
        N.network_analyze(np.array([0]), np.array([5]), np.array([1.0]), 3)
    with pytest.raises(ValueError):
        N.network_analyze(np.array([0]), np.array([1]), np.array([1.0]), 0)
