"""The C-ABI's contract beyond the solvers: the ASCII grid writer (host code, no GPU), argument errors
(status 1, ValueError) and a circuit that is not positive definite (status 2, RuntimeError with the
reference's message, src/cpp/circuit/grid.cpp:208)."""

from __future__ import annotations

import ctypes
from pathlib import Path

import numpy as np
import pytest

GOLDEN = Path(__file__).resolve().parent / "golden" / "to_chars_vectors.txt"


def _vectors():
    bits, text = [], []
    for line in GOLDEN.read_text().splitlines():
        b, t = line.split()
        bits.append(int(b, 16))
        text.append(t)
    return np.array(bits, dtype=np.uint64).view(np.float64), text


def test_ascii_grid_rows_writes_to_chars_text():
    """3,000 doubles from 1e-30 to 1e30, random bit patterns, short decimals, integers and powers of two,
    against C++17 std::to_chars (tests/golden/make_to_chars_vectors.py), as 60 rows of 50."""
    from landschaft_amd.circuit import _backend

    vals, text = _vectors()
    assert vals.size == 3000
    body = _backend.ascii_grid_rows(vals.reshape(60, 50)).decode()
    rows = body.split("\n")
    assert rows[-1] == "" and len(rows) == 61
    assert [t for r in rows[:-1] for t in r.split(" ")] == text
    back = np.array([float(t) for r in rows[:-1] for t in r.split(" ")])
    np.testing.assert_array_equal(back, vals)


def test_ascii_grid_rows_nodata_and_shapes():
    from landschaft_amd.circuit import _backend

    a = np.array([[np.nan, 0.0, -0.0], [1e16, 1e-5, 123456789012345678.0]])
    assert _backend.ascii_grid_rows(a) == b"-9999 0 -0\n1e+16 1e-05 123456789012345680\n"
    assert _backend.ascii_grid_rows(np.zeros((0, 4))) == b""
    assert _backend.ascii_grid_rows(np.zeros((3, 0))) == b"\n\n\n"
    # many rows: the rows are formatted on several threads and joined in order
    big = np.arange(20000.0).reshape(4000, 5) / 8
    lines = _backend.ascii_grid_rows(big).decode().splitlines()
    assert lines[0] == "0 0.125 0.25 0.375 0.5" and lines[-1] == "2499.375 2499.5 2499.625 2499.75 2499.875"
    with pytest.raises(ValueError, match="2-D"):
        _backend.ascii_grid_rows(np.zeros(4))


def test_ascii_grid_rows_status_codes():
    from landschaft_amd import _native as N

    lib = N.lib()
    a = np.ones((2, 3))
    size = ctypes.c_int64(0)
    assert lib.lsc_ascii_grid_rows(N.ptr(a), 3, 2, None, 0, ctypes.byref(size)) == 0
    assert size.value == len("1 1 1\n1 1 1\n")
    small = ctypes.create_string_buffer(4)
    assert lib.lsc_ascii_grid_rows(N.ptr(a), 3, 2, small, 4, ctypes.byref(size)) == 1
    assert "too small" in lib.lsc_last_error().decode()
    assert lib.lsc_ascii_grid_rows(N.ptr(a), -1, 2, None, 0, ctypes.byref(size)) == 1
    with pytest.raises(ValueError, match="negative grid size"):
        N.check(1)


def test_grid_writer_goes_through_the_library(tmp_path):
    """circuitscape's output grids are written by core().ascii_grid_rows, as the reference's
    circuitscape.py:1027 writes them."""
    from landschaft_amd.circuit import circuitscape as cs

    meta = cs.RasterMeta(nrows=2, ncols=3, xllcorner=0.0, yllcorner=0.0, cellsize=1.0)
    cs._write_grid(str(tmp_path / "m"), np.array([[0.1 + 0.2, np.nan, 2.0], [1e-7, 5e-324, -3.5]]), meta, False)
    text = (tmp_path / "m.asc").read_text().splitlines()
    assert text[6:] == ["0.30000000000000004 -9999 2", "1e-07 5e-324 -3.5"]


def test_mirror_rejects_bad_shapes_before_the_device():
    """ValueError for mismatched rasters (the binding's python.cpp:67-70; reference tests/test_circuit.py:123-127),
    raised on the host: no GPU is needed to see it."""
    from landschaft_amd.circuit import _backend
    from landschaft_amd.circuit.advanced import advanced

    with pytest.raises(ValueError, match="same shape"):
        _backend.circuit_advanced(np.ones((2, 2)), np.ones((2, 3)), np.ones((2, 2)))
    with pytest.raises(ValueError, match="same shape"):
        advanced(np.ones((2, 2)), np.ones((2, 3)), np.ones((2, 2)))
    with pytest.raises(ValueError, match="negative"):
        advanced(np.ones((2, 2)), np.ones((2, 2)), -np.ones((2, 2)))


@pytest.mark.gpu
def test_negative_ground_is_not_positive_definite():
    """A negative conductance to ground passed straight to the C-ABI (the front end rejects it first):
    status 2 and the reference's message; the mirror raises RuntimeError with it."""
    from landschaft_amd import _native as N
    from landschaft_amd.circuit import _backend

    g = np.ones((3, 4))
    s = np.zeros((3, 4))
    s[0, 0] = 1.0
    gr = np.zeros((3, 4))
    gr[2, 3] = -5.0
    cur, volt = np.empty((3, 4)), np.empty((3, 4))
    st = N.SolveStats()
    o = N.options()
    rc = N.lib().lsc_circuit_advanced(N.ptr(g), N.ptr(s), N.ptr(gr), 4, 3, 0, ctypes.byref(o), N.ptr(cur),
                                      N.ptr(volt), ctypes.byref(st))
    assert rc == 2
    assert N.lib().lsc_last_error().decode() == "a circuit component is not positive definite"
    with pytest.raises(RuntimeError, match="a circuit component is not positive definite"):
        _backend.circuit_advanced(g, s, gr)
    # the raster-graph path of Circuitscape's advanced mode, FP64 as well
    for precision in ("mixed", "fp64"):
        _backend.settings.precision = precision
        try:
            with pytest.raises(RuntimeError, match="not positive definite"):
                _backend.circuitscape_advanced(g, None, s, gr, False, False, "keepall")
        finally:
            _backend.settings.precision = "mixed"


@pytest.mark.gpu
def test_c_abi_argument_errors():
    """Status 1 (ValueError) for invalid options before any work: an even block size, a negative radius,
    a pair cell off the raster."""
    from landschaft_amd.circuit import _backend

    r = np.ones((20, 20))
    with pytest.raises(ValueError, match="odd"):
        _backend.circuit_omniscape(r, r, 3, 0, 2, False, False, True, False)
    with pytest.raises(ValueError, match=">= 0"):
        _backend.circuit_omniscape(r, r, -1, 0, 1, False, False, True, False)
    with pytest.raises(ValueError, match="out of range"):
        _backend.effective_resistance_grad(r, [0], [400])


def test_auto_precision_option():
    """"auto" reaches the C-ABI as LSC_PRECISION_AUTO and is the settings' default."""
    from landschaft_amd import _native as N
    from landschaft_amd.circuit import _backend

    assert N.options("auto").precision == N.PRECISION_AUTO == 2
    assert N.options("mixed").precision == N.PRECISION_MIXED
    assert N.options("fp64").precision == N.PRECISION_FP64
    assert _backend.SolverSettings().precision == "auto"
    header = (Path(__file__).resolve().parents[1] / "include" / "landschaft_circuit.h").read_text()
    assert "LSC_PRECISION_AUTO = 2" in header
