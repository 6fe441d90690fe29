"""Circuitscape's network (graph) mode (the reference's tests/test_circuitscape_network.py).

``circuitscape_network.npz`` holds Circuitscape 6.1's outputs for its own
network test INI files and for synthetic ones (``network_extra``: 10^3-10^4
nodes, include lists, advanced with finite and direct grounds). Every case
runs on the CPU with ``core()`` answered by the numpy oracle
(``oracle/network.py``) and on the GPU with the HIP library (mixed precision
refined to a relative residual of 1e-14), at the reference's tolerances.
"""

from __future__ import annotations

import importlib
from pathlib import Path

import numpy as np
import pytest

from oracle_core import OracleCore

cs = importlib.import_module("landschaft_amd.circuit.circuitscape")
net = importlib.import_module("landschaft_amd.circuit.network")

FIX = Path(__file__).resolve().parent / "fixtures" / "circuit"
CS = FIX / "circuitscape"
INIS = sorted(p.relative_to(CS).as_posix() for d in ("network", "network_extra") for p in (CS / "input" / d).glob("*.ini"))


@pytest.fixture(scope="module")
def outputs():
    return np.load(FIX / "circuitscape_network.npz")


@pytest.fixture
def in_cs(monkeypatch):
    monkeypatch.chdir(CS)


@pytest.fixture(params=["oracle", pytest.param("gpu", marks=pytest.mark.gpu)])
def engine(request, monkeypatch, oracle):
    if request.param == "oracle":
        oc = OracleCore(oracle)
        monkeypatch.setattr(net, "core", lambda: oc)
    else:
        from landschaft_amd.circuit import _backend

        monkeypatch.setattr(_backend.settings, "precision", "mixed")
        monkeypatch.setattr(_backend.settings, "tol", 1e-14)
    return request.param


def _outputs(r) -> dict:
    if r.resistances is not None:
        out = {"result.txt": r.resistances, "resistances.out": r.resistances}
        out.update({f"{k}_{p}_{q}.txt": t for (k, p, q), t in r.maps.items()})
        if r.cum_node_currents is not None:
            out["node_currents_cum.txt"] = r.cum_node_currents
            out["branch_currents_cum.txt"] = r.cum_branch_currents
        return out
    return {"result.txt": r.voltages, "voltages.txt": r.voltages, "node_currents.txt": r.node_currents,
            "branch_currents.txt": r.branch_currents}


@pytest.mark.parametrize("ini", INIS)
def test_matches_circuitscape(ini, in_cs, engine, outputs):
    name = Path(ini).stem
    refs = {k.split("/", 1)[1]: outputs[k] for k in outputs.files if k.split("/", 1)[0] == name}
    assert refs, name
    ours = _outputs(cs.circuitscape_ini(ini, write_outputs=False))
    # every output Circuitscape wrote, and no other
    assert set(refs) - {"resistances_3columns.out"} == set(ours), name
    for key, want in refs.items():
        if key not in ours:
            continue
        got = np.asarray(ours[key], dtype=float)
        want = np.asarray(want, dtype=float)
        if want.size == 0:
            assert got.size == 0, key
            continue
        got = got.reshape(want.shape) if got.size == want.size else got
        assert got.shape == want.shape, key
        np.testing.assert_allclose(got, want, rtol=1e-11, atol=1e-12 * np.abs(want).max(), err_msg=f"{name} {key}")


def test_series_and_parallel(engine):
    # 1 -2- 2 -1- 3, and the edge 1-2 twice (conductances add: 0.5 + 0.5)
    edges = [[1, 2, 2.0], [2, 3, 1.0], [2, 1, 2.0]]
    r = net.circuitscape_network_pairwise(edges, [1, 3, 2])
    np.testing.assert_allclose(r.resistances[1:, 1:], [[0, 2, 1], [2, 0, 1], [1, 1, 0]], rtol=1e-13, atol=1e-14)
    np.testing.assert_array_equal(r.resistances[0, 1:], [1, 3, 2])
    r = net.circuitscape_network_pairwise([[1, 2, 4.0]], [1, 2], graph_is_resistance=False)
    assert r.resistances[1, 2] == pytest.approx(0.25, rel=1e-13)


def test_numbering_from_zero(engine):
    a = net.circuitscape_network_pairwise([[0, 1, 1.0], [1, 2, 1.0]], [0, 2])
    b = net.circuitscape_network_pairwise([[1, 2, 1.0], [2, 3, 1.0]], [1, 3])
    np.testing.assert_array_equal(a.resistances, b.resistances)  # ids written from 1
    with pytest.raises(ValueError, match="from 2"):
        net.circuitscape_network_pairwise([[2, 3, 1.0]], [2, 3])


def test_disconnected_and_missing_points(engine):
    # a zero-sum pair of lines drops the edge 3-4; node 5 has no edges
    edges = [[1, 2, 1.0], [3, 4, 1.0], [4, 3, -1.0], [2, 5, 0.0], [1, 1, 3.0]]
    r = net.circuitscape_network_pairwise(edges, [1, 2, 3, 4, 9, 2], graph_is_resistance=False)
    res = r.resistances[1:, 1:]
    assert res[0, 1] == pytest.approx(1.0) and res[1, 5] == 0 and res[0, 5] == pytest.approx(1.0)
    assert res[2, 3] == -1
    assert np.all(res[4, [0, 1, 2, 3, 5]] == -1)  # node 9 is not in the graph
    assert np.all(np.diag(res) == 0)
    assert r.n_components == 4  # {1, 2}, {3}, {4}, {5}


def test_advanced_by_hand(engine):
    # 1 A into node 1 of 1 -1- 2 -2- 3, node 3 grounded directly: V = 3, 2, 0
    r = net.circuitscape_network_advanced([[1, 2, 1.0], [2, 3, 2.0]], [[1, 1.0]], [[3, 0.0]])
    np.testing.assert_allclose(r.voltages, [[1, 3], [2, 2], [3, 0]], rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(r.node_currents[:, 1], [1, 1, 1], rtol=1e-13)
    np.testing.assert_allclose(r.branch_currents, [[1, 2, 1], [2, 3, 1]], rtol=1e-13)
    # a ground of 1 ohm at node 3 instead: 1 V more everywhere
    r = net.circuitscape_network_advanced([[1, 2, 1.0], [2, 3, 2.0]], [[1, 1.0]], [[3, 1.0]])
    np.testing.assert_allclose(r.voltages[:, 1], [4, 3, 1], rtol=1e-13)
    with pytest.raises(ValueError, match="not in the graph"):
        net.circuitscape_network_advanced([[1, 2, 1.0]], [[5, 1.0]], [[2, 0.0]])


def _random_graph(n, chords, seed):
    rng = np.random.default_rng(seed)
    e = [[v, v + 1, rng.uniform(0.5, 3)] for v in range(1, n)]
    for _ in range(chords):
        a, b = rng.integers(1, n + 1, 2)
        e.append([a, b, rng.uniform(0.5, 3)])
    return np.array(e, dtype=float)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_effective_resistances_match_the_pseudo_inverse(seed, engine):
    # the graph nested dissection's factor against a dense solve
    edges = _random_graph(600, 900, seed)
    pts = np.random.default_rng(seed).choice(np.arange(1, 601), 7, replace=False)
    r = net.circuitscape_network_pairwise(edges, pts)
    n = 600
    L = np.zeros((n, n))
    for a, b, w in edges:
        a, b = int(a) - 1, int(b) - 1
        if a == b:
            continue
        L[a, b] -= 1 / w
        L[b, a] -= 1 / w
        L[a, a] += 1 / w
        L[b, b] += 1 / w
    P = np.linalg.inv(L + 1.0 / n)  # (L + 11'/n)^-1: the differences below are those of pinv(L)
    idx = pts - 1
    Q = P[idx][:, idx]
    want = Q.diagonal()[:, None] + Q.diagonal()[None, :] - 2 * Q
    np.testing.assert_allclose(r.resistances[1:, 1:], want, rtol=1e-9, atol=1e-12)


def test_repeated_runs_give_the_same_bits(engine):
    edges = _random_graph(3000, 2000, 7)
    pts = [5, 900, 1777, 2999]
    a = net.circuitscape_network_pairwise(edges, pts, current_maps=True, voltage_maps=True)
    b = net.circuitscape_network_pairwise(edges, pts, current_maps=True, voltage_maps=True)
    np.testing.assert_array_equal(a.resistances, b.resistances)
    np.testing.assert_array_equal(a.cum_node_currents, b.cum_node_currents)
    for k in a.maps:
        np.testing.assert_array_equal(a.maps[k], b.maps[k])
    src, gnd = [[5, 1.0], [900, 2.0]], [[2999, 0.0], [1777, 3.0]]
    x = net.circuitscape_network_advanced(edges, src, gnd)
    y = net.circuitscape_network_advanced(edges, src, gnd)
    np.testing.assert_array_equal(x.voltages, y.voltages)
    np.testing.assert_array_equal(x.branch_currents, y.branch_currents)


def test_large_network_matches_the_oracle(oracle):
    # 3,000 nodes with chords: the GPU path's maps against the oracle's, pair by pair
    pytest.importorskip("landschaft_amd._native").lib()
    edges = _random_graph(3000, 2000, 7)
    pts = [5, 900, 1777, 2999]
    gpu = net.circuitscape_network_pairwise(edges, pts, current_maps=True, voltage_maps=True)
    oc = OracleCore(oracle)
    saved = net.core
    net.core = lambda: oc
    try:
        cpu = net.circuitscape_network_pairwise(edges, pts, current_maps=True, voltage_maps=True)
    finally:
        net.core = saved
    np.testing.assert_allclose(gpu.resistances, cpu.resistances, rtol=1e-11)
    np.testing.assert_allclose(gpu.cum_node_currents, cpu.cum_node_currents, rtol=1e-10, atol=1e-12)
    assert set(gpu.maps) == set(cpu.maps)
    for k in cpu.maps:
        np.testing.assert_allclose(gpu.maps[k], cpu.maps[k], rtol=1e-10, atol=1e-11, err_msg=str(k))


test_large_network_matches_the_oracle = pytest.mark.gpu(test_large_network_matches_the_oracle)


def test_included_pairs(tmp_path, engine):
    edges = _random_graph(50, 20, 3)
    f = tmp_path / "pairs.txt"
    f.write_text("mode include\n1 20\n20 40\n")
    r = net.circuitscape_network_pairwise(edges, [1, 20, 40, 7], included_pairs=f)
    # include mode drops the focal nodes it does not name (7) and solves only the pairs it names
    np.testing.assert_array_equal(r.resistances[0, 1:], [1, 20, 40])
    res = r.resistances[1:, 1:]
    assert res[0, 1] > 0 and res[1, 2] > 0 and res[0, 2] == -1
    f.write_text("mode exclude\n1 20\n")
    r = net.circuitscape_network_pairwise(edges, [1, 20, 40], included_pairs=f)
    res = r.resistances[1:, 1:]
    assert res[0, 1] == -1 and res[0, 2] > 0 and res[1, 2] > 0


def test_writes_outputs(tmp_path, in_cs, engine):
    ini = (CS / "input" / "network_extra" / "net1.ini").read_text()
    out = tmp_path / "out" / "net.out"
    ini = ini.replace("output_file = output/net1.out", f"output_file = {out.as_posix()}")
    (tmp_path / "net.ini").write_text(ini)
    r = cs.circuitscape_ini(tmp_path / "net.ini")
    names = sorted(p.name for p in out.parent.iterdir())
    assert "net_resistances.out" in names and "net_resistances_3columns.out" in names
    assert "net_node_currents_cum.txt" in names and "net_branch_currents_cum.txt" in names
    p, q = next((p, q) for (k, p, q) in r.maps if k == "voltages")
    assert f"net_voltages_{p}_{q}.txt" in names and f"net_branch_currents_{p}_{q}.txt" in names
    np.testing.assert_allclose(np.loadtxt(out.parent / "net_resistances.out"), r.resistances, rtol=1e-15)
    np.testing.assert_allclose(np.loadtxt(out.parent / f"net_voltages_{p}_{q}.txt"), r.maps[("voltages", p, q)],
                               rtol=1e-15)
    assert "data_type = network" in out.read_text()


def test_top_level_names():
    from landschaft_amd import circuit

    assert circuit.NetworkResult is not None
    assert callable(circuit.circuitscape_network_advanced)
    assert callable(circuit.read_graph)


@pytest.mark.gpu
def test_grid_network_matches_the_raster_path():
    # a 4-neighbour grid given as an edge list (conductance (g_a + g_b) / 2) has the raster's
    # Laplacian: the general-graph ordering must give the raster path's resistances
    from landschaft_amd.circuit import _backend as B

    n, npts = 384, 6
    rng = np.random.default_rng(11)
    g = np.exp(rng.standard_normal((n, n)))
    idx = np.arange(n * n).reshape(n, n)
    a = np.concatenate([idx[:, :-1].ravel(), idx[:-1, :].ravel()])
    b = np.concatenate([idx[:, 1:].ravel(), idx[1:, :].ravel()])
    w = (g.ravel()[a] + g.ravel()[b]) / 2
    cells = rng.choice(n * n, npts, replace=False)
    ex = np.zeros((npts, npts), np.uint8)
    Rn = B.network_pairwise(a, b, w, n * n, cells, ex, {})["resistances"]
    Rr = B.circuitscape_pairwise(g, None, cells.astype(np.int64), ex, False, True, {})["resistances"]
    np.testing.assert_allclose(Rn, Rr, rtol=1e-9, atol=0)
    assert np.all(Rn[~np.eye(npts, dtype=bool)] > 0)
