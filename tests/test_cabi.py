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

    # 10 int32 + 8 float + 14 pointers + 2 int32 + pointer + 2 int32 + pointer,
    # naturally aligned (no padding); the library asserts the same size
    assert ctypes.sizeof(_native.PtyInstance) == 10 * 4 + 8 * 4 + 14 * 8 + 2 * 4 + 8 + 2 * 4 + 8 == 216
    assert _native.PtyInstance.detector_error.offset == 208 and _native.PtyInstance.pc_radius.offset == 204


def _declared_net():
    text = (ROOT / "include" / "pty_net.h").read_text()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\*?\s+\*?(pty_net_\w+)\(", text, re.M)))


def test_network_library_exports_every_declared_symbol():
    from ptylab_amd import unet

    names = _declared_net()
    assert "pty_net_conv_fwd" in names and "pty_net_conv_wgrad" in names
    if not unet.LIB_PATH.exists():
        pytest.skip("libptynet.so not built")
    lib = ctypes.CDLL(str(unet.LIB_PATH))
    for name in names:
        assert hasattr(lib, name), name
    assert set(unet.EXPORTED) == set(names)
    assert unet.load().pty_net_kpad(3, 8) == unet.kpad(3, 8) == 96


def _param_counts(header):
    text = (ROOT / "include" / header).read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    out = {}
    for m in re.finditer(r"^\s*(?:const\s+)?\w+\*?\s+\*?(pty_\w+)\(([^)]*)\);", text, re.M | re.S):
        args = m.group(2).strip()
        out[m.group(1)] = 0 if args in ("", "void") else args.count(",") + 1
    return out


@pytest.mark.parametrize("header,module", [("pty_hip.h", "_native"), ("pty_net.h", "unet")])
def test_ctypes_signatures_match_header_parameter_counts(header, module):
    import importlib

    mod = importlib.import_module(f"ptylab_amd.{module}")
    if not mod.LIB_PATH.exists():
        pytest.skip("library not built")
    lib = mod.load()
    for name, n in _param_counts(header).items():
        at = getattr(lib, name).argtypes
        if at is None:
            assert n == 0, name
        else:
            assert len(at) == n, (name, len(at), n)


def test_header_constants_match_the_binding():
    """Every numeric #define PTY_* of include/pty_hip.h has the same value in
    ptylab_amd._native (engine, propagator, flag and error codes)."""
    import re
    from pathlib import Path

    from ptylab_amd import _native

    text = (Path(__file__).resolve().parents[1] / "include" / "pty_hip.h").read_text()
    found = 0
    for name, value in re.findall(r"#define (PTY_[A-Z0-9_]+) \(?(-?\d+)\)?", text):
        if hasattr(_native, name):
            assert getattr(_native, name) == int(value), name
            found += 1
    for name in ("PTY_SCALED_ASP", "PTY_SCALED_POLYCHROME_ASP", "PTY_QNEWTON", "PTY_FLAG_BATCHED"):
        assert hasattr(_native, name), name
    assert found >= 12
