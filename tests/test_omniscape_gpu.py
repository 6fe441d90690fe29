"""Omniscape on the GPU against the oracle and the reference's stored outputs."""

from __future__ import annotations

import numpy as np
import pytest

from conftest import FIXTURES
from test_oracle import _omniscape_cases, _omniscape_inputs

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def backend():
    from landschaft_amd.circuit import _backend

    saved = (_backend.settings.precision, _backend.settings.tol)
    _backend.settings.tol = 1e-14
    yield _backend
    _backend.settings.precision, _backend.settings.tol = saved


@pytest.mark.parametrize("precision", [0, 1])
def test_mfma_operand_layout(precision):
    from landschaft_amd import _native as N

    got = N.probe_mfma(precision)
    i = np.arange(16)[:, None]
    j = np.arange(16)[None, :]
    expect = sum((i + 16 * k) * (k + 4 * j) for k in range(4)).astype(float)
    np.testing.assert_array_equal(got, expect)


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
@pytest.mark.parametrize("name", _omniscape_cases())
def test_omniscape_matches_cholmod(backend, oracle, name, precision):
    z = np.load(FIXTURES / "omniscape.npz")
    res, s, args, o = _omniscape_inputs(z, name)
    backend.settings.precision = precision
    print(f"case {name} {precision}", flush=True)
    out = backend.circuit_omniscape(res, s, args["radius"], args["buffer"], args["block_size"],
                                    args["four_neighbors"], args["resistance_is_conductance"],
                                    args["correct_artifacts"], args["flow_potential"])
    ref = oracle.omniscape(res, s, **args)
    assert out["n_targets"] == ref["n_targets"]
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
        np.testing.assert_array_equal(v == 0, chol == 0)
        np.testing.assert_allclose(v, chol, rtol=1e-9, atol=1e-13 * top)


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_omniscape_random_map_matches_oracle(backend, oracle, precision):
    rng = np.random.default_rng(11)
    r = np.exp(rng.normal(0.0, 1.0, (90, 120)))
    r[rng.random(r.shape) < 0.08] = np.nan
    s = np.where(np.isnan(r), 0.0, 1 / np.nan_to_num(r, nan=1.0))
    backend.settings.precision = precision
    for radius, block, four in [(12, 5, False), (9, 3, True), (20, 1, False)]:
        print(f"random r={radius} b={block} four={four} {precision}", flush=True)
        out = backend.circuit_omniscape(r, s, radius, 0, block, four, False, True, False)
        ref = oracle.omniscape(r, s, radius, 0, block, four, False, True, False)
        assert out["n_targets"] == ref["n_targets"]
        top = ref["current"].max()
        np.testing.assert_allclose(out["current"], ref["current"], rtol=1e-9, atol=1e-12 * top)


def test_correction_matches_oracle(backend, oracle):
    backend.settings.precision = "mixed"
    for args in [(6, 2, 5, False, 50, 50), (10, 0, 3, True, 30, 40), (4, 1, 1, False, 50, 50)]:
        got = backend.circuit_omniscape_correction(*args)
        ref = oracle.omniscape_correction(*args)
        np.testing.assert_allclose(got, ref, rtol=1e-9)


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_batches_in_flight_give_identical_maps(backend, oracle, precision, monkeypatch):
    """The same small batches on one stream, on three (the default) and on four streams (the
    accumulation is chained in batch order): the cumulative and flow-potential maps are bit-identical,
    and match the oracle. The streams share the one-batch budget, so `batch` 384 over three streams and
    512 over four are batches of 128."""
    rng = np.random.default_rng(21)
    r = np.exp(rng.normal(0.0, 1.0, (70, 90)))
    r[rng.random(r.shape) < 0.08] = np.nan
    s = np.where(np.isnan(r), 0.0, 1 / np.nan_to_num(r, nan=1.0))
    backend.settings.precision = precision
    maps = {}
    for streams, batch in (("1", 128), ("3", 384), ("4", 512)):
        monkeypatch.setattr(backend.settings, "batch", batch)
        monkeypatch.setenv("LSC_BATCH_STREAMS", streams)
        print(f"streams {streams} {precision}", flush=True)
        maps[streams] = backend.circuit_omniscape(r, s, 7, 1, 1, False, False, True, True)
        assert backend.last_stats["n_batches"] > 8
    for k in ("current", "flow_potential"):
        np.testing.assert_array_equal(maps["1"][k], maps["3"][k])
        np.testing.assert_array_equal(maps["1"][k], maps["4"][k])
    ref = oracle.omniscape(r, s, 7, 1, 1, False, False, True, True)
    top = ref["current"].max()
    np.testing.assert_allclose(maps["4"]["current"], ref["current"], rtol=1e-9, atol=1e-12 * top)


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_graph_replay_matches_eager_launches(backend, oracle, precision, monkeypatch):
    """Each batch's launch sequences (flood, factorisation, substitution sweeps, residual) replayed as
    hipGraphs, the default, and launched one by one (LSC_GRAPHS=0): bit-identical maps on one and on
    four streams. The first batch stream is the caller's null stream, which launches eagerly; the
    other three replay graphs. 100-window batches over about 5,800 targets replay each graph many
    times, and the smaller last batch captures again."""
    rng = np.random.default_rng(22)
    r = np.exp(rng.normal(0.0, 1.0, (70, 90)))
    r[rng.random(r.shape) < 0.08] = np.nan
    s = np.where(np.isnan(r), 0.0, 1 / np.nan_to_num(r, nan=1.0))
    backend.settings.precision = precision
    monkeypatch.setattr(backend.settings, "batch", 100)
    for streams in ("1", "4"):
        monkeypatch.setenv("LSC_BATCH_STREAMS", streams)
        maps = {}
        for graphs in ("0", "1"):
            monkeypatch.setenv("LSC_GRAPHS", graphs)
            print(f"graphs {graphs} streams {streams} {precision}", flush=True)
            maps[graphs] = backend.circuit_omniscape(r, s, 7, 1, 1, False, False, True, True)
            assert backend.last_stats["n_batches"] > 10
            if precision == "mixed":
                assert backend.last_stats["refine_max"] >= 2  # the residual and sweep graphs replay
        for k in ("current", "flow_potential"):
            np.testing.assert_array_equal(maps["1"][k], maps["0"][k])
    ref = oracle.omniscape(r, s, 7, 1, 1, False, False, True, True)
    top = ref["current"].max()
    np.testing.assert_allclose(maps["1"]["current"], ref["current"], rtol=1e-9, atol=1e-12 * top)


@pytest.mark.parametrize("four", [False, True])
@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_lane_substitutions_match_per_front_kernels(backend, oracle, precision, four, monkeypatch):
    """The small levels' substitutions four lanes per window (the default) against the
    one-front-per-wavefront kernels (LSC_LANE_LEVELS=0), radius 12 windows with 8 and 4 neighbours:
    the maps agree to rounding (the two sum in different orders) and both match the oracle."""
    rng = np.random.default_rng(31)
    r = np.exp(rng.normal(0.0, 1.0, (60, 80)))
    r[rng.random(r.shape) < 0.1] = np.nan
    s = np.where(np.isnan(r), 0.0, 1 / np.nan_to_num(r, nan=1.0))
    backend.settings.precision = precision
    maps = {}
    for lanes in ("0", "1"):
        monkeypatch.setenv("LSC_LANE_LEVELS", lanes)
        print(f"lane levels {lanes} {precision} four {four}", flush=True)
        maps[lanes] = backend.circuit_omniscape(r, s, 12, 0, 3, four, False, True, True)
    ref = oracle.omniscape(r, s, 12, 0, 3, four, False, True, True)
    for k in ("current", "flow_potential"):
        top = np.nanmax(np.abs(ref[k]))
        np.testing.assert_allclose(maps["1"][k], maps["0"][k], rtol=1e-11, atol=1e-13 * top)
        np.testing.assert_allclose(maps["1"][k], ref[k], rtol=1e-9, atol=1e-12 * top)


@pytest.mark.parametrize("contrast", [1e6, 1e10])
def test_fp64_fallback_of_high_contrast_windows(backend, oracle, contrast):
    """Corridors of conductance `contrast` times the background's: where the FP32 factor's refinement
    stalls (windows that reach the corridors: about half of them at 1e6, all 216 at 1e10) the window is
    solved again in FP64 (n_fallback); the map matches the FP64 path and the oracle."""
    rng = np.random.default_rng(3)
    r = np.full((60, 80), np.sqrt(contrast))
    r[:, 38:42] = 1 / np.sqrt(contrast)
    r[28:32, :] = 1 / np.sqrt(contrast)
    r *= np.exp(rng.normal(0, 0.3, r.shape))
    s = 1 / r
    ref = oracle.omniscape(r, s, 10, 0, 3, False, False, True, False)
    out = {}
    for precision in ("mixed", "fp64"):
        backend.settings.precision, backend.settings.tol = precision, 1e-10
        print(f"corridor contrast {contrast:g} {precision}", flush=True)
        out[precision] = backend.circuit_omniscape(r, s, 10, 0, 3, False, False, True, False)["current"]
        st = backend.last_stats
        if precision == "mixed":
            assert 0 < st["n_fallback"] < st["n_solved"]
        else:
            assert st["n_fallback"] == 0
    top = ref["current"].max()
    np.testing.assert_allclose(out["mixed"], out["fp64"], rtol=1e-9, atol=1e-12 * top)
    np.testing.assert_allclose(out["mixed"], ref["current"], rtol=1e-9, atol=1e-10 * top)


@pytest.mark.timeout(300)
@pytest.mark.parametrize("radius,block,four,shape,flow", [(60, 7, False, (100, 120), False),
                                                          (70, 9, True, (110, 130), True),
                                                          (100, 11, False, (150, 170), True)])
def test_large_radius_windows_match_oracle(backend, oracle, radius, block, four, shape, flow):
    """Windows with fronts of more than 159 rows (radius >= 54: the large-front tiles in HBM, a level at a
    time for every window) and, at radius 100 (201 cells a side), the wide flood fill and voltages in HBM
    for the node currents, mixed and FP64, against the oracle."""
    from landschaft_amd import _native as N

    assert N.window_plan_stats(radius)["max_front"] + 1 > 160
    rng = np.random.default_rng(radius)
    r = np.exp(rng.normal(0.0, 0.8, shape))
    r[rng.random(shape) < 0.06] = np.nan
    s = np.where(np.isnan(r), 0.0, 1 / np.nan_to_num(r, nan=1.0))
    ref = oracle.omniscape(r, s, radius, 0, block, four, False, True, flow)
    top = ref["current"].max()
    for precision in ("mixed", "fp64"):
        backend.settings.precision, backend.settings.tol = precision, 1e-12
        print(f"radius {radius} block {block} four {four} {shape} {precision}", flush=True)
        out = backend.circuit_omniscape(r, s, radius, 0, block, four, False, True, flow)
        assert out["n_targets"] == ref["n_targets"] > 0
        np.testing.assert_allclose(out["current"], ref["current"], rtol=1e-9, atol=1e-12 * top)
        if flow:
            np.testing.assert_allclose(out["flow_potential"], ref["flow_potential"], rtol=1e-9,
                                       atol=1e-12 * ref["flow_potential"].max())


def test_large_radius_correction_matches_oracle(backend, oracle):
    """Correction factors of windows with large fronts (and, at radius 100, the wide flood fill). A factor
    is a quotient of two currents: where the block circuit's current is small (down to 1e-6 of its
    maximum, the floor) the quotient carries a larger relative error than either solve, hence 1e-8."""
    backend.settings.precision, backend.settings.tol = "mixed", 1e-12
    for args in [(100, 0, 11, False, 300, 300), (64, 4, 7, True, 200, 150)]:
        print(f"correction {args}", flush=True)
        np.testing.assert_allclose(backend.circuit_omniscape_correction(*args), oracle.omniscape_correction(*args),
                                   rtol=1e-8)
