"""The Omniscape front end (landschaft_amd.circuit.omniscape) on the GPU: the properties the reference
checks of its own front end (tests/test_omniscape.py:28-135), restated on this package's API."""

from __future__ import annotations

import warnings

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

OUTPUTS = ("cum_currmap", "flow_potential", "normalized_cum_currmap")


@pytest.fixture(scope="module")
def api():
    from landschaft_amd import circuit
    from landschaft_amd.circuit import _backend

    saved = (_backend.settings.precision, _backend.settings.tol)
    _backend.settings.precision, _backend.settings.tol = "mixed", 1e-13
    yield circuit
    _backend.settings.precision, _backend.settings.tol = saved


def _map(seed=1, shape=(40, 50)):
    """A lognormal resistance map with 5 % NoData cells."""
    rng = np.random.default_rng(seed)
    r = np.exp(rng.normal(0.5, 1.0, shape))
    r[rng.random(shape) < 0.05] = np.nan
    return r


def test_two_targets_are_two_advanced_solves(api):
    """With one-cell blocks and two source cells, each source cell is a target; its window's one other
    source carries the target's own strength to it (a target's amps are shared among its window's
    sources): the cumulative map is the sum of two advanced-mode solves on the disks."""
    r = _map(2, (15, 15))
    r[np.isnan(r)] = 1.0
    s = np.zeros(r.shape)
    s[7, 7], s[5, 9] = 1.0, 2.0
    ds = api.omniscape(r, 4, source=s)
    assert ds.attrs["n_targets"] == 2
    yy, xx = np.mgrid[:15, :15]
    total = np.zeros(r.shape)
    for (ty, tx), other in (((7, 7), (5, 9)), ((5, 9), (7, 7))):
        g = np.where((yy - ty) ** 2 + (xx - tx) ** 2 <= 16, 1 / r, np.nan)
        src = np.zeros(r.shape)
        src[other] = s[ty, tx]
        ground = np.zeros(r.shape)
        ground[ty, tx] = np.inf
        total += api.advanced(g, src, ground).current
    np.testing.assert_allclose(ds.cum_currmap.values, total, rtol=1e-10, atol=1e-13 * total.max())


def test_host_threads_do_not_change_the_maps(api):
    r = _map(3)
    s = np.where(np.isnan(r), 0, 1 / np.nan_to_num(r, nan=1))
    kw = dict(source=s, block_size=3, buffer=2, flow_potential=True, normalized_current=True)
    a = api.omniscape(r, 6, n_threads=1, **kw)
    b = api.omniscape(r, 6, **kw)
    for k in OUTPUTS:
        np.testing.assert_array_equal(getattr(a, k).values, getattr(b, k).values)


def test_block_size_nodata_masks_and_outputs(api):
    r = _map(4)
    with pytest.warns(UserWarning, match="block_size is even"):
        ds = api.omniscape(r, 5, block_size=2, flow_potential=True, normalized_current=True)
    assert ds.attrs["block_size"] == 3
    nodata = np.isnan(r)
    for k in OUTPUTS:
        v = getattr(ds, k).values
        assert np.isnan(v[nodata]).all() and not np.isnan(v[~nodata]).any()
    ds2 = api.omniscape(r, 5, block_size=3, flow_potential=True, normalized_current=True, mask_nodata=False)
    assert not np.isnan(ds2.cum_currmap.values).any()
    np.testing.assert_array_equal(ds2.cum_currmap.values[~nodata], ds.cum_currmap.values[~nodata])
    # normalised current = current / flow potential, 0 where both are 0
    cur, fp = ds2.cum_currmap.values, ds2.flow_potential.values
    with np.errstate(invalid="ignore", divide="ignore"):
        expect = np.where((cur == 0) & (fp == 0), 0, cur / fp)
    np.testing.assert_array_equal(ds2.normalized_cum_currmap.values, expect)
    with warnings.catch_warnings():
        warnings.simplefilter("error")
        assert list(api.omniscape(r, 5).data_vars) == ["cum_currmap"]
        assert list(api.omniscape(r, 5, normalized_current=True).data_vars) == ["cum_currmap",
                                                                                 "normalized_cum_currmap"]


def test_sources_from_resistance(api):
    r = _map(5)
    s = api.source_from_resistance(r, r_cutoff=2.0)
    expect = np.where(np.isnan(r) | (1 / r < 0.5), 0, 1 / r)
    np.testing.assert_array_equal(s, expect)
    a = api.omniscape(r, 4, r_cutoff=2.0)
    b = api.omniscape(r, 4, source=expect)
    np.testing.assert_array_equal(a.cum_currmap.values, b.cum_currmap.values)
    # the same circuit given as conductances
    c = api.omniscape(1 / r, 4, source=expect, resistance_is_conductance=True)
    np.testing.assert_allclose(c.cum_currmap.values, a.cum_currmap.values, rtol=1e-10)


def test_source_threshold_and_nan_sources(api):
    r = _map(6)
    s = np.random.default_rng(6).uniform(0, 1, r.shape)
    s[::7, ::5] = np.nan
    a = api.omniscape(r, 4, source=s, source_threshold=0.4)
    b = api.omniscape(r, 4, source=np.where(np.isnan(s) | (s < 0.4), 0, s))
    np.testing.assert_array_equal(a.cum_currmap.values, b.cum_currmap.values)


def test_correction_factor_symmetries(api):
    from landschaft_amd.circuit import _backend

    np.testing.assert_array_equal(_backend.circuit_omniscape_correction(4, 1, 1, False, 50, 50), 1.0)
    c = _backend.circuit_omniscape_correction(6, 2, 5, False, 50, 50)
    assert c.shape == (17, 17)
    np.testing.assert_allclose(c, c.T, rtol=1e-9)
    np.testing.assert_allclose(c, c[::-1, :], rtol=1e-9)
    assert np.all(c > 0)


def test_bad_inputs(api):
    r = _map(8)
    with pytest.raises(ValueError, match="> 0"):
        api.omniscape(np.where(np.isnan(r), 0, r), 3)
    with pytest.raises(ValueError, match="shape"):
        api.omniscape(r, 3, source=np.ones((3, 3)))
