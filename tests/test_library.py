"""The C-ABI library: it loads and exports every symbol include/*.h declares (no GPU calls)."""

from __future__ import annotations

import ctypes
import re

from conftest import ROOT


def _declared():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        names |= set(re.findall(r"\b(lsc_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_header_declares_entry_points():
    names = _declared()
    assert "lsc_circuit_omniscape" in names and "lsc_last_error" in names


def test_library_exports_every_declared_symbol():
    from landschaft_amd import _native as N

    lib = N.lib()
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, missing
    assert isinstance(lib.lsc_last_error(), bytes)  # the calling thread's last message (earlier tests may set one)


def test_python_binding_lists_the_header():
    from landschaft_amd import _native as N

    assert set(_declared()) <= set(N.EXPORTS)
