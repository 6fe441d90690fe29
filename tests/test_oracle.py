"""The FP64 CPU oracle, pinned before it is trusted.

Against the reference's stored Circuitscape / Omniscape outputs (CHOLMOD runs,
tests/fixtures/circuit, copied from the reference's
tests/data/reference/circuit), at the reference's tolerances
(tests/test_circuit.py:135-148, tests/test_omniscape.py:136-163 of the
reference), and against closed forms and independent dense solves.
"""

from __future__ import annotations

import json

import numpy as np
import pytest

from conftest import FIXTURES


def _dense_advanced(g, s, gr, four=False):
    """Direct dense solve of one component with direct grounds (column-major nodes, as Julia)."""
    ny, nx = g.shape
    nodes = [(y, x) for x in range(nx) for y in range(ny) if g[y, x] > 0]
    index = {c: k for k, c in enumerate(nodes)}
    n = len(nodes)
    L = np.zeros((n, n))
    steps = [(0, 1), (1, 0), (-1, 0), (0, -1)] + ([] if four else [(1, 1), (-1, -1), (1, -1), (-1, 1)])
    for (y, x), k in index.items():
        for dy, dx in steps:
            j = index.get((y + dy, x + dx))
            if j is None:
                continue
            w = (g[y, x] + g[y + dy, x + dx]) / (2 * np.sqrt(2) if dy and dx else 2)
            L[k, j] -= w
            L[k, k] += w
    b = np.array([s[c] for c in nodes])
    grounded = [index[c] for c in nodes if gr[c] == np.inf]
    keep = [k for k in range(n) if k not in grounded]
    b[grounded] = 0
    v = np.zeros(n)
    v[keep] = np.linalg.solve(L[np.ix_(keep, keep)], b[keep])
    return nodes, L, v


def _advanced_cases():
    z = np.load(FIXTURES / "advanced_inputs.npz")
    return sorted({k.split("/")[0] for k in z.files})


@pytest.mark.parametrize("name", _advanced_cases())
def test_oracle_advanced_matches_cholmod(oracle, name):
    inp = np.load(FIXTURES / "advanced_inputs.npz")
    ref = np.load(FIXTURES / "advanced.npz")
    g = np.nan_to_num(inp[f"{name}/conductance"], nan=0.0)
    s = np.nan_to_num(inp[f"{name}/source"], nan=0.0)
    gr = np.nan_to_num(inp[f"{name}/ground"], nan=0.0)
    r = oracle.advanced(g, s, gr, bool(inp[f"{name}/four"]))
    chol = ref[f"{name}/current_cholmod"]
    top = chol.max()
    np.testing.assert_array_equal(r["current"] != 0, chol != 0)
    np.testing.assert_allclose(r["current"], chol, rtol=1e-9, atol=1e-12 * top)
    np.testing.assert_allclose(r["current"], ref[f"{name}/current_cgamg"], rtol=0, atol=1e-5 * top)


def _omniscape_inputs(z, name):
    o = json.loads(str(z[f"{name}/options"]))
    res = z[f"{name}/resistance"]
    bs = o.get("block_size", 1)
    bs = bs + 1 if bs % 2 == 0 else bs
    rc = o.get("resistance_is_conductance", False)
    if o.get("source_from_resistance"):
        with np.errstate(divide="ignore"):
            s = res if rc else 1 / res
        s = np.where(np.isnan(s) | (s < 1 / float(o.get("r_cutoff", np.inf))), 0.0, s)
    else:
        s = np.nan_to_num(z[f"{name}/source"], nan=0.0)
        s = np.where(s < o.get("source_threshold", 0.0), 0.0, s)
    fp = bool(o.get("calc_flow_potential", False) or o.get("calc_normalized_current", False))
    args = dict(radius=o["radius"], buffer=o.get("buffer", 0), block_size=bs,
                four_neighbors=bool(o.get("connect_four_neighbors_only", False)), resistance_is_conductance=rc,
                correct_artifacts=bool(o.get("correct_artifacts", True)), flow_potential=fp)
    return res, s, args, o


def _omniscape_cases():
    z = np.load(FIXTURES / "omniscape.npz")
    return sorted({k.split("/")[0] for k in z.files})


@pytest.mark.parametrize("name", _omniscape_cases())
def test_oracle_omniscape_matches_cholmod(oracle, name):
    z = np.load(FIXTURES / "omniscape.npz")
    res, s, args, o = _omniscape_inputs(z, name)
    out = oracle.omniscape(res, s, **args)
    maps = {"cum_currmap": out["current"]}
    if args["flow_potential"]:
        maps["flow_potential"] = out["flow_potential"]
    for k, v in maps.items():
        key = f"{name}/cholmod/{k}"
        if key not in z.files:
            continue
        v = v.copy()
        if o.get("mask_nodata", True):
            v[np.isnan(res)] = np.nan
        chol = z[key]
        top = np.nanmax(chol)
        np.testing.assert_array_equal(np.isnan(v), np.isnan(chol))
        np.testing.assert_allclose(v, chol, rtol=1e-9, atol=1e-13 * top)


def test_two_cells(oracle):
    r = oracle.advanced(np.array([[1.0, 1.0]]), np.array([[1.0, 0.0]]), np.array([[0.0, np.inf]]))
    np.testing.assert_allclose(r["current"], [[1.0, 1.0]], rtol=1e-14)
    np.testing.assert_allclose(r["voltage"], [[1.0, 0.0]], rtol=1e-14)


def test_series_and_diagonal(oracle):
    r = oracle.advanced(np.array([[1.0, 3.0, 2.0]]), np.array([[2.0, 0, 0]]), np.array([[0, 0, np.inf]]))
    np.testing.assert_allclose(r["voltage"], [[2 / 2.5 + 2 / 2, 2 / 2.5, 0]], rtol=1e-13)
    g = np.array([[1.0, 0.0], [0.0, 1.0]])
    r = oracle.advanced(g, np.array([[1.0, 0], [0, 0]]), np.array([[0, 0], [0, np.inf]]))
    np.testing.assert_allclose(r["voltage"][0, 0], np.sqrt(2), rtol=1e-13)


@pytest.mark.parametrize("four", [False, True])
def test_oracle_matches_dense_solve(oracle, four):
    rng = np.random.default_rng(3)
    g = rng.uniform(0.1, 2, (9, 11))
    s = np.where(rng.random(g.shape) < 0.3, rng.uniform(0, 1, g.shape), 0.0)
    gr = np.zeros(g.shape)
    gr[4, 5] = np.inf
    r = oracle.advanced(g, s, gr, four)
    nodes, L, v = _dense_advanced(g, s, gr, four)
    np.testing.assert_allclose([r["voltage"][c] for c in nodes], v, rtol=1e-12, atol=1e-14)
    B = -L * (v[:, None] - v[None, :])
    np.fill_diagonal(B, 0)
    node_current = np.maximum(np.clip(B, 0, None).sum(axis=0), np.clip(B, 0, None).sum(axis=1))
    np.testing.assert_allclose([r["current"][c] for c in nodes], node_current, rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(r["current"][4, 5], s.sum() - s[4, 5], rtol=1e-7)


def test_finite_grounds_split(oracle):
    r = oracle.advanced(np.ones((1, 3)), np.array([[0.0, 1.0, 0.0]]), np.array([[1.0, 0.0, 1.0]]))
    np.testing.assert_allclose(r["current"], [[0.5, 1.0, 0.5]], rtol=1e-13)


def test_omniscape_single_target_is_one_advanced_solve(oracle):
    # reference tests/test_omniscape.py:28-47: block 1, two source cells
    rng = np.random.default_rng(2)
    r = np.exp(rng.normal(0.5, 1.0, (15, 15)))
    s = np.zeros(r.shape)
    s[7, 7], s[5, 9] = 1.0, 2.0
    out = oracle.omniscape(r, s, 4, 0, 1, False, False, True, False)
    assert out["n_targets"] == 2
    total = np.zeros(r.shape)
    for (ty, tx), amps, other in [((7, 7), 1.0, (5, 9)), ((5, 9), 2.0, (7, 7))]:
        yy, xx = np.mgrid[:15, :15]
        g = np.where((yy - ty) ** 2 + (xx - tx) ** 2 <= 16, 1 / r, 0.0)
        src = np.zeros(r.shape)
        src[other] = amps
        gr = np.zeros(r.shape)
        gr[ty, tx] = np.inf
        total += oracle.advanced(g, src, gr)["current"]
    np.testing.assert_allclose(out["current"], total, rtol=1e-12, atol=1e-15)


def test_correction_factors(oracle):
    np.testing.assert_array_equal(oracle.omniscape_correction(4, 1, 1, False, 50, 50), 1.0)
    c = oracle.omniscape_correction(6, 2, 5, False, 50, 50)
    assert c.shape == (17, 17)
    np.testing.assert_allclose(c, c.T, rtol=1e-9)
    np.testing.assert_allclose(c, c[::-1, :], rtol=1e-9)
    assert np.all(c > 0)


def test_window_statistics(oracle):
    # The radius-50 window: 7,844 unknowns, 903 fronts, largest front 150 rows.
    st = oracle.window_stats(50)
    assert st["unknowns"] == 7844 and st["fronts"] == 903 and st["max_front"] == 150


def test_pairwise_docstring_and_pseudo_inverse(oracle):
    # circuitscape.py:621-636: a 10 x 10 map with a column of resistance 10
    r = np.ones((10, 10))
    r[:, 5] = 10
    cells = np.array([11, 81, 58], dtype=np.int64)
    got = oracle.cs_pairwise(1 / r, None, cells, np.zeros((3, 3), np.uint8), False, False, {"cum": True})
    R = got["resistances"]
    np.testing.assert_allclose([R[0, 1], R[0, 2], R[1, 2]], [0.8515, 0.8567, 0.8409], atol=5e-5)
    np.testing.assert_allclose(R, R.T, rtol=1e-12)
    # the full matrix against R_ij = P_ii + P_jj - 2 P_ij, P = (L + 1/n)^-1
    g = 1 / r
    n = g.size
    L = np.zeros((n, n))
    for y in range(10):
        for x in range(10):
            for dy, dx in ((0, 1), (1, 0), (1, 1), (1, -1)):
                yy, xx = y + dy, x + dx
                if 0 <= yy < 10 and 0 <= xx < 10:
                    w = (g[y, x] + g[yy, xx]) / (2 * np.sqrt(2) if dy and dx else 2)
                    a, b = y * 10 + x, yy * 10 + xx
                    L[a, b] -= w
                    L[b, a] -= w
                    L[a, a] += w
                    L[b, b] += w
    P = np.linalg.inv(L + 1.0 / n)
    for i in range(3):
        for j in range(3):
            s, t = cells[i], cells[j]
            np.testing.assert_allclose(R[i, j], P[s, s] + P[t, t] - 2 * P[s, t], rtol=1e-10, atol=1e-14)
    # the shortcut (no maps) and the explicit solves agree
    full = oracle.cs_pairwise(1 / r, None, cells, np.zeros((3, 3), np.uint8), False, False, {"cur_maps": True})
    np.testing.assert_allclose(full["resistances"], R, rtol=1e-11)


def test_multi_advanced_schur_matches_per_problem_factors(oracle):
    # the Schur complement on the special nodes against one factorisation per problem
    rng = np.random.default_rng(12)
    g = np.exp(rng.normal(0, 1, (25, 30)))
    g[rng.random(g.shape) < 0.08] = 0.0
    pts = [2 * 30 + 3, 20 * 30 + 25, 12 * 30 + 14, 23 * 30 + 2, 5 * 30 + 27]
    problems = []
    for c in pts:
        others = [p for p in pts if p != c]
        problems.append({"src_cells": np.array([c]), "src_vals": np.array([1.5]),
                         "gnd_cells": np.array(others + [0]), "gnd_vals": np.array([np.inf] * len(others) + [0.3]),
                         "check_cell": c})
    for policy in ("rmvgnd", "keepall"):
        a = oracle.cs_multi_advanced(g, None, False, False, problems, policy, schur=True)
        b = oracle.cs_multi_advanced(g, None, False, False, problems, policy, schur=False)
        np.testing.assert_array_equal(a["solved"], b["solved"])
        np.testing.assert_allclose(a["voltage"], b["voltage"], rtol=1e-9, atol=1e-12)
        np.testing.assert_allclose(a["current"], b["current"], rtol=1e-9, atol=1e-12)
