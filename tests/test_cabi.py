"""The C-ABI library loads and exports every symbol include/pty_hip.h
declares (no compute calls: this runs without a GPU)."""

import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent


def _declared():
    text = (ROOT / "include" / "pty_hip.h").read_text()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\*?\s+\*?(pty_\w+)\(", text, re.M)))


def test_header_declares_expected_entry_points():
    names = _declared()
    assert "pty_plan_create" in names and "pty_plan_run" in names and "pty_propagate" in names


def test_library_exports_every_declared_symbol():
    from ptylab_amd import _native

    if not _native.LIB_PATH.exists():
        pytest.skip("libptyhip.so not built")
    lib = ctypes.CDLL(str(_native.LIB_PATH))
    for name in _declared():
        assert hasattr(lib, name), name
    assert set(_native.EXPORTED) == set(_declared())
    assert _native.load().pty_abi_version() == _native.PTY_ABI_VERSION


def test_struct_layout_matches_header():
    from ptylab_amd import _native

    # 10 int32 + 8 float + 14 pointers + 2 int32, naturally aligned
    assert ctypes.sizeof(_native.PtyInstance) == 10 * 4 + 8 * 4 + 14 * 8 + 2 * 4
