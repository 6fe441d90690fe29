"""Circuitscape's raster modes on the GPU graph engine against the FP64 oracle."""

from __future__ import annotations

import numpy as np
import pytest

from conftest import FIXTURES

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def backend():
    from landschaft_amd.circuit import _backend

    saved = (_backend.settings.precision, _backend.settings.tol)
    _backend.settings.tol = 1e-14
    yield _backend
    _backend.settings.precision, _backend.settings.tol = saved


def _advanced_cases():
    z = np.load(FIXTURES / "advanced_inputs.npz")
    return sorted({k.split("/")[0] for k in z.files})


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
@pytest.mark.parametrize("name", _advanced_cases())
def test_advanced_matches_cholmod(backend, name, precision):
    inp = np.load(FIXTURES / "advanced_inputs.npz")
    ref = np.load(FIXTURES / "advanced.npz")
    backend.settings.precision = precision
    print(f"advanced {name} {precision}", flush=True)
    g = np.nan_to_num(inp[f"{name}/conductance"], nan=0.0)
    s = np.nan_to_num(inp[f"{name}/source"], nan=0.0)
    gr = np.nan_to_num(inp[f"{name}/ground"], nan=0.0)
    r = backend.circuit_advanced(g, s, gr, bool(inp[f"{name}/four"]))
    chol = ref[f"{name}/current_cholmod"]
    top = chol.max()
    np.testing.assert_array_equal(r["current"] != 0, chol != 0)
    np.testing.assert_allclose(r["current"], chol, rtol=1e-9, atol=1e-12 * top)


@pytest.mark.parametrize("shape,four", [((9, 11), False), ((60, 70), True), ((230, 210), False)])
def test_advanced_random_matches_oracle(backend, oracle, shape, four):
    # 230 x 210 has fronts beyond 159 rows: the HBM-tiled large-front path
    rng = np.random.default_rng(sum(shape))
    g = np.exp(rng.normal(0, 1, shape))
    g[rng.random(shape) < 0.05] = 0.0
    s = np.where(rng.random(shape) < 0.02, rng.uniform(0, 1, shape), 0.0)
    gr = np.zeros(shape)
    gr[shape[0] // 2, shape[1] // 3] = np.inf
    gr[1, 1] = 0.5
    backend.settings.precision = "mixed"
    print(f"advanced random {shape} {four}", flush=True)
    got = backend.circuit_advanced(g, s, gr, four)
    ref = oracle.advanced(g, s, gr, four)
    top = ref["current"].max()
    np.testing.assert_allclose(got["current"], ref["current"], rtol=1e-9, atol=1e-12 * top)
    np.testing.assert_allclose(got["voltage"], ref["voltage"], rtol=1e-9, atol=1e-12 * np.nanmax(np.abs(ref["voltage"])))


@pytest.mark.parametrize("policy", ["keepall", "rmvsrc", "rmvgnd", "rmvall"])
def test_circuitscape_advanced_polygons_and_policies(backend, oracle, policy):
    rng = np.random.default_rng(5)
    shape = (40, 50)
    g = np.exp(rng.normal(0, 0.7, shape))
    g[10:14, 5:30] = 0.0
    poly = np.zeros(shape, dtype=np.int64)
    poly[20:24, 20:26] = 3
    poly[30:33, 40:45] = 7
    src = np.zeros(shape)
    src[2, 3], src[35, 44], src[21, 21] = 1.0, 2.0, 0.5
    gnd = np.zeros(shape)
    gnd[38, 2], gnd[21, 22], gnd[35, 44] = np.inf, 2.0, 1.0
    backend.settings.precision = "mixed"
    for avg in (False, True):
        got = backend.circuitscape_advanced(g, poly, src, gnd, avg, False, policy)
        ref = oracle.cs_advanced(g, poly, src, gnd, avg, False, policy)
        np.testing.assert_array_equal(got["nodemap"], ref["nodemap"])
        nm = got["nodemap"]
        on = nm >= 0
        top_v = np.abs(ref["voltage"]).max() or 1.0
        np.testing.assert_allclose(got["voltage"][nm[on]], ref["voltage"][on], rtol=1e-9, atol=1e-12 * top_v)
        top_c = np.abs(ref["current"]).max() or 1.0
        np.testing.assert_allclose(got["current"][nm[on]], ref["current"][on], rtol=1e-9, atol=1e-12 * top_c)
        assert got["solved"] == ref["solved"] and got["n_components"] == ref["n_components"]


@pytest.mark.parametrize("maps", [False, True])
def test_pairwise_matches_oracle(backend, oracle, maps):
    rng = np.random.default_rng(9)
    shape = (60, 80)
    g = np.exp(rng.normal(0, 1, shape))
    g[:, 40] = 0.0
    g[30, 40] = 1.0  # one gap joins the halves
    cells = np.array([5 * 80 + 5, 50 * 80 + 10, 20 * 80 + 70, 55 * 80 + 75, 5 * 80 + 5], dtype=np.int64)
    ex = np.zeros((5, 5), dtype=np.uint8)
    opts = {"cum": maps, "max": maps, "volt_maps": maps, "cur_maps": maps}
    backend.settings.precision = "mixed"
    got_maps, ref_maps = {}, {}
    got = backend.circuitscape_pairwise(g, None, cells, ex, False, False, opts,
                                        emit=lambda ci, cj, v, c: got_maps.__setitem__((ci, cj), (v, c)))
    ref = oracle.cs_pairwise(g, None, cells, ex, False, False, opts,
                             emit=lambda ci, cj, v, c: ref_maps.__setitem__((ci, cj), (v, c)))
    np.testing.assert_allclose(got["resistances"], ref["resistances"], rtol=1e-9, atol=1e-12)
    if maps:
        for k in ("cum", "max"):
            np.testing.assert_allclose(got[k], ref[k], rtol=1e-9, atol=1e-12 * np.abs(ref[k]).max())
        assert got_maps.keys() == ref_maps.keys()
        for key in ref_maps:
            for a, b in zip(got_maps[key], ref_maps[key]):
                np.testing.assert_allclose(a, b, rtol=1e-9, atol=1e-12 * max(np.abs(b).max(), 1e-300))


def test_pairwise_docstring_example(backend):
    # the reference's circuitscape.py:621-636 (resistances on a 10 x 10 map)
    r = np.ones((10, 10))
    r[:, 5] = 10
    cells = np.array([1 * 10 + 1, 8 * 10 + 1, 5 * 10 + 8], dtype=np.int64)
    got = backend.circuitscape_pairwise(1 / r, None, cells, np.zeros((3, 3), np.uint8), False, False, {"cum": True})
    np.testing.assert_allclose(got["resistances"][0, 1:], [0.8515, 0.8567], atol=5e-5)
    np.testing.assert_allclose(got["resistances"][1, 2], 0.8409, atol=5e-5)


def test_multi_advanced_matches_oracle(backend, oracle):
    rng = np.random.default_rng(4)
    shape = (30, 40)
    g = np.exp(rng.normal(0, 1, shape))
    pts = [3 * 40 + 3, 25 * 40 + 30, 10 * 40 + 35, 28 * 40 + 2]
    problems = []
    for i, c in enumerate(pts):  # all-to-one: point grounded, 1 A at the others
        others = [p for p in pts if p != c]
        problems.append({"src_cells": np.array(others), "src_vals": np.ones(len(others)),
                         "gnd_cells": np.array([c]), "gnd_vals": np.array([np.inf]), "check_cell": c})
    backend.settings.precision = "mixed"
    got = backend.circuitscape_multi_advanced(g, None, False, False, problems, "rmvsrc")
    ref = oracle.cs_multi_advanced(g, None, False, False, problems, "rmvsrc")
    nm = got["nodemap"]
    on = nm >= 0
    for p, r in enumerate(got["results"]):
        assert r["solved"] == ref["solved"][p]
        np.testing.assert_allclose(r["voltage"][nm[on]], ref["voltage"][p][on], rtol=1e-9, atol=1e-12)
        np.testing.assert_allclose(r["current"][nm[on]], ref["current"][p][on], rtol=1e-9, atol=1e-12)
        assert set(r["active"]) == set(np.unique(nm[ref["active"][p] & on]))


@pytest.mark.parametrize("policy", ["rmvgnd", "keepall"])
def test_multi_advanced_one_to_all_finite_grounds(backend, oracle, policy):
    # one-to-all problems with a finite ground too (the Schur path with a finite diagonal), polygons
    rng = np.random.default_rng(12)
    g = np.exp(rng.normal(0, 1, (45, 60)))
    g[rng.random(g.shape) < 0.08] = 0.0
    poly = np.zeros(g.shape, dtype=np.int64)
    poly[30:33, 10:14] = 4
    pts = [2 * 60 + 3, 40 * 60 + 55, 31 * 60 + 11, 44 * 60 + 2, 5 * 60 + 57]
    problems = []
    for c in pts:
        others = [p for p in pts if p != c]
        problems.append({"src_cells": np.array([c]), "src_vals": np.array([1.5]),
                         "gnd_cells": np.array(others + [0]), "gnd_vals": np.array([np.inf] * len(others) + [0.3]),
                         "check_cell": c})
    backend.settings.precision = "mixed"
    got = backend.circuitscape_multi_advanced(g, poly, False, False, problems, policy)
    ref = oracle.cs_multi_advanced(g, poly, False, False, problems, policy, schur=False)
    nm = got["nodemap"]
    on = nm >= 0
    for p, r in enumerate(got["results"]):
        assert r["solved"] == ref["solved"][p]
        top = np.abs(ref["voltage"][p]).max()
        np.testing.assert_allclose(r["voltage"][nm[on]], ref["voltage"][p][on], rtol=1e-9, atol=1e-12 * top)
        np.testing.assert_allclose(r["current"][nm[on]], ref["current"][p][on], rtol=1e-9,
                                   atol=1e-12 * np.abs(ref["current"][p]).max())


@pytest.mark.parametrize("focal_zero", [False, True])
def test_multi_advanced_sum_matches_per_problem_maps(backend, focal_zero):
    # the device reduction against the per-problem outputs summed on the host, in problem order
    rng = np.random.default_rng(21)
    g = np.exp(rng.normal(0, 1, (50, 64)))
    g[rng.random(g.shape) < 0.1] = 0.0
    g[:, 30] = 0.0  # two components
    pts = [3 * 64 + 3, 45 * 64 + 50, 20 * 64 + 10, 40 * 64 + 5, 10 * 64 + 60, 25 * 64 + 40]
    problems = []
    for c in pts:
        others = [p for p in pts if p != c]
        problems.append({"src_cells": np.array(others), "src_vals": np.ones(len(others)),
                         "gnd_cells": np.array([c]), "gnd_vals": np.array([np.inf]), "check_cell": c})
    backend.settings.precision = "mixed"
    full = backend.circuitscape_multi_advanced(g, None, False, False, problems, "rmvsrc")
    nm = full["nodemap"]
    on = nm >= 0
    cum0 = np.zeros(g.shape)
    mx0 = np.full(g.shape, -9999.0)
    for r in full["results"]:
        cur = np.zeros(g.shape)
        cur[on] = r["current"][nm[on]]
        if focal_zero:
            cur[np.isin(nm, r["active"])] = 0.0
        cum0 += cur
        np.maximum(mx0, cur, out=mx0)
    cum = np.zeros(g.shape)
    mx = np.full(g.shape, -9999.0)
    out = backend.circuitscape_multi_advanced_sum(g, None, False, False, problems, "rmvsrc", cum, mx, focal_zero)
    np.testing.assert_array_equal(out["nodemap"], nm)
    np.testing.assert_array_equal(out["solved"], [r["solved"] for r in full["results"]])
    np.testing.assert_array_equal(cum, cum0)
    np.testing.assert_array_equal(mx, mx0)
    for p, (c, r) in enumerate(zip(pts, full["results"])):
        assert out["check_voltage"][p] == r["voltage"][nm.flat[c]]


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_row_parallel_large_front_substitution(backend, oracle, monkeypatch, precision):
    # the large fronts' substitutions by rows (every level with a front of >= 64 rows) against the
    # one-workgroup-per-front kernels (no level) and the oracle: 20 pairs with maps on a 300 x 300
    # 8-neighbour raster (largest front about 600 rows), so 19 right-hand sides per solve
    rng = np.random.default_rng(21)
    g = np.exp(rng.standard_normal((300, 300)))
    g[rng.random((300, 300)) < 0.05] = 0.0
    on = np.flatnonzero(g.ravel() > 0)
    cells = rng.choice(on, 7, replace=False).astype(np.int64)
    ex = np.zeros((7, 7), np.uint8)
    opts = {"cum": True, "cur_maps": True}
    backend.settings.precision = precision
    out = {}
    for label, rows in (("per-front", "100000000"), ("by rows", "64")):
        monkeypatch.setenv("LSC_WIDE_ROWS", rows)
        print(f"large-front substitution {label} {precision}", flush=True)
        maps = {}
        r = backend.circuitscape_pairwise(g, None, cells, ex, False, False, opts,
                                          lambda ci, cj, v, c: maps.__setitem__((ci, cj), c))
        out[label] = (r["resistances"], r["cum"], maps)
    ref = oracle.cs_pairwise(g, None, cells, ex, False, False, opts)
    for label, (R, cum, maps) in out.items():
        np.testing.assert_allclose(R, ref["resistances"], rtol=1e-10, atol=0, err_msg=label)
        np.testing.assert_allclose(cum, ref["cum"], rtol=1e-9, atol=1e-12 * np.abs(ref["cum"]).max(), err_msg=label)
    a, b = out["per-front"], out["by rows"]
    np.testing.assert_allclose(b[0], a[0], rtol=1e-12)
    for key in a[2]:
        np.testing.assert_allclose(b[2][key], a[2][key], rtol=1e-9, atol=1e-12 * np.abs(a[2][key]).max())
    # the row-parallel path is deterministic
    monkeypatch.setenv("LSC_WIDE_ROWS", "64")
    again = backend.circuitscape_pairwise(g, None, cells, ex, False, False, opts)
    np.testing.assert_array_equal(again["resistances"], b[0])
    np.testing.assert_array_equal(again["cum"], b[1])


def test_graph_replay_matches_eager_pcg_sweeps(backend, oracle, monkeypatch):
    """Mixed precision on a badly conditioned raster (r = exp(2 z)) takes several preconditioned CG
    iterations; every sweep after the first is replayed as a hipGraph by default. The resistances and
    the cumulative map are bit-identical to launching the sweeps one by one (LSC_GRAPHS=0)."""
    rng = np.random.default_rng(23)
    g = np.exp(2.0 * rng.standard_normal((160, 160)))
    cells = rng.choice(160 * 160, 6, replace=False).astype(np.int64)
    ex = np.zeros((6, 6), np.uint8)
    backend.settings.precision = "mixed"
    out = {}
    for graphs in ("0", "1"):
        monkeypatch.setenv("LSC_GRAPHS", graphs)
        print(f"pcg graphs {graphs}", flush=True)
        out[graphs] = backend.circuitscape_pairwise(g, None, cells, ex, False, True, {"cum": True})
        assert backend.last_stats["refine_max"] >= 3
    np.testing.assert_array_equal(out["1"]["resistances"], out["0"]["resistances"])
    np.testing.assert_array_equal(out["1"]["cum"], out["0"]["cum"])
    ref = oracle.cs_pairwise(g, None, cells, ex, False, True, {"cum": True})
    np.testing.assert_allclose(out["1"]["resistances"], ref["resistances"], rtol=1e-10, atol=0)


@pytest.mark.parametrize("network", [False, True])
def test_pairwise_jobs_in_chunks(backend, oracle, network, monkeypatch):
    """40 focal points (780 pairs) with the cumulative map: the right-hand sides solved in chunks of 64
    (LSC_PAIR_CHUNK) give the same resistances and map as one solve of all of them (bit for bit: each
    right-hand side is solved on its own and the maps add in pair order), and match the oracle."""
    rng = np.random.default_rng(40)
    shape = (60, 70)
    g = np.exp(rng.normal(0, 1, shape))
    g[rng.random(shape) < 0.05] = 0.0
    data = np.flatnonzero(g > 0)
    cells = rng.choice(data, 40, replace=False).astype(np.int64)
    ex = np.zeros((40, 40), np.uint8)
    ex[3, 7] = ex[7, 3] = 1
    backend.settings.precision, backend.settings.tol = "mixed", 1e-12
    out = {}
    for chunk in ("0", "64"):
        if chunk == "0":
            monkeypatch.delenv("LSC_PAIR_CHUNK", raising=False)
        else:
            monkeypatch.setenv("LSC_PAIR_CHUNK", chunk)
        print(f"pairwise 40 points, chunk {chunk}, network {network}", flush=True)
        if network:
            idx = np.arange(g.size).reshape(shape)
            a = np.concatenate([idx[:, :-1].ravel(), idx[:-1, :].ravel()])
            b = np.concatenate([idx[:, 1:].ravel(), idx[1:, :].ravel()])
            w = (g.ravel()[a] + g.ravel()[b]) / 2
            keep = w > 0
            out[chunk] = backend.network_pairwise(a[keep].astype(np.int32), b[keep].astype(np.int32), w[keep], g.size,
                                                  cells.astype(np.int32), ex, {"cum": True})
        else:
            out[chunk] = backend.circuitscape_pairwise(g, None, cells, ex, False, True, {"cum": True})
    np.testing.assert_array_equal(out["0"]["resistances"], out["64"]["resistances"])
    np.testing.assert_array_equal(out["0"]["cum"], out["64"]["cum"])
    if not network:
        ref = oracle.cs_pairwise(g, None, cells, ex, False, True, {"cum": True})
        np.testing.assert_allclose(out["64"]["resistances"], ref["resistances"], rtol=1e-9)
        top = ref["cum"].max()
        np.testing.assert_allclose(out["64"]["cum"], ref["cum"], rtol=1e-9, atol=1e-12 * top)


def test_auto_precision_by_right_hand_sides(backend, oracle):
    """precision "auto": the pairwise factorisation of 12 points with the cumulative map (66 jobs) runs
    in FP64 (the same bits as precision "fp64"), the advanced solve of one right-hand side in mixed
    precision (the same bits as "mixed"); both match the oracle."""
    rng = np.random.default_rng(21)
    shape = (60, 70)
    g = np.exp(rng.normal(0, 1, shape))
    cells = rng.choice(shape[0] * shape[1], 12, replace=False).astype(np.int64)
    ex = np.zeros((12, 12), dtype=np.uint8)
    s = np.zeros(shape)
    s[5, 5] = 1.0
    gr = np.zeros(shape)
    gr[50, 60] = np.inf
    runs = {}
    try:
        for prec in ("auto", "fp64", "mixed"):
            backend.settings.precision = prec
            runs[prec] = (backend.circuitscape_pairwise(g, None, cells, ex, False, False, {"cum": True}),
                          backend.circuit_advanced(g, s, gr, False))
    finally:
        backend.settings.precision = "mixed"
    for k in ("resistances", "cum"):
        np.testing.assert_array_equal(runs["auto"][0][k], runs["fp64"][0][k])
    np.testing.assert_array_equal(runs["auto"][1]["current"], runs["mixed"][1]["current"])
    ref = oracle.cs_pairwise(g, None, cells, ex, False, False, {"cum": True})
    np.testing.assert_allclose(runs["auto"][0]["resistances"], ref["resistances"], rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(runs["auto"][0]["cum"], ref["cum"], rtol=1e-9, atol=1e-12 * np.abs(ref["cum"]).max())
    ref_adv = oracle.advanced(g, s, gr, False)
    top = np.abs(ref_adv["current"]).max()
    np.testing.assert_allclose(runs["auto"][1]["current"], ref_adv["current"], rtol=1e-9, atol=1e-12 * top)
