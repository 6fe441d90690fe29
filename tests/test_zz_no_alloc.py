"""The allocation-free Omniscape entry point (lsc_circuit_omniscape_device_static) against the allocating
one (lsc_circuit_omniscape_device), through the C-ABI on device arrays, on two calls in a row, the second
on a smaller map (the scratch keeps the first call's contents): the same target counts and the same maps
up to rounding. Each window is solved alike; the cumulative map adds a batch's windows at once, and the
scratch holds smaller batches, so the map's sums are grouped differently (differences of a few ulp)."""

from __future__ import annotations

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _run(entry, r, s, radius, block, flow, precision):
    import torch

    from landschaft_amd import _native as N

    ny, nx = r.shape
    dev = torch.device("cuda:0")
    tr = torch.from_numpy(r).to(dev)
    ts = torch.from_numpy(s).to(dev)
    cur = torch.full((ny, nx), np.nan, dtype=torch.float64, device=dev)
    fp = torch.full((ny, nx), np.nan, dtype=torch.float64, device=dev)
    o = N.options(precision, 1e-10)
    st = N.SolveStats()
    nt = ctypes.c_int64(0)
    fn = getattr(N.lib(), entry)
    N.check(fn(ctypes.c_void_p(tr.data_ptr()), ctypes.c_void_p(ts.data_ptr()), nx, ny, radius, 0, block, 0, 0, 1,
               int(flow), None, ctypes.byref(o), ctypes.c_void_p(cur.data_ptr()),
               ctypes.c_void_p(fp.data_ptr()) if flow else None, ctypes.byref(nt), ctypes.byref(st)))
    torch.cuda.synchronize()
    return cur.cpu().numpy(), (fp.cpu().numpy() if flow else None), nt.value, st.as_dict()


def _map(shape, seed):
    rng = np.random.default_rng(seed)
    r = np.exp(rng.normal(0.0, 1.0, shape))
    r[rng.random(shape) < 0.08] = np.nan
    s = np.where(np.isnan(r), 0.0, 1 / np.nan_to_num(r, nan=1.0))
    return r, s


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_static_entry_matches_allocating_entry(precision):
    # first a larger map and radius (many batches of the scratch's size), then a smaller one
    for shape, radius, block, flow, seed in (((200, 260), 30, 3, True, 41), ((70, 90), 8, 5, False, 42)):
        r, s = _map(shape, seed)
        print(f"static vs allocating: {shape} radius {radius} block {block} flow {flow} {precision}", flush=True)
        a_cur, a_fp, a_nt, a_st = _run("lsc_circuit_omniscape_device", r, s, radius, block, flow, precision)
        b_cur, b_fp, b_nt, b_st = _run("lsc_circuit_omniscape_device_static", r, s, radius, block, flow, precision)
        assert b_nt == a_nt > 0
        assert b_st["n_streams"] == 1 and b_st["n_solved"] == a_st["n_solved"]
        np.testing.assert_allclose(b_cur, a_cur, rtol=1e-13, atol=1e-14 * np.max(a_cur))
        np.testing.assert_array_equal(b_cur == 0, a_cur == 0)
        if flow:
            np.testing.assert_allclose(b_fp, a_fp, rtol=1e-13, atol=1e-14 * np.max(a_fp))
        if radius == 30:
            assert b_st["n_batches"] > a_st["n_batches"]  # the scratch held fewer windows per batch
