"""C-ABI library checks that need no GPU: it loads, exports every function
that include/atrip_hip.h declares, and refuses work cleanly without a device."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    return sorted(set(re.findall(r"^\w[\w \*]*?\b(atrip_hip_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    from atrip_amd import hip

    L = hip.load()
    names = _declared("atrip_hip.h")
    assert len(names) >= 14
    for name in names:
        assert hasattr(L, name), name
    assert sorted(hip.SYMBOLS) == names


def test_no_device_is_an_error_not_a_crash():
    from atrip_amd import hip

    if hip.device_count() > 0:
        pytest.skip("a device is visible")
    with pytest.raises(hip.HipError):
        hip.Context(0)


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    from atrip_amd import hip

    monkeypatch.setattr(hip, "LIB_PATH", str(tmp_path / "absent.so"))
    monkeypatch.setattr(hip, "_lib", None)
    with pytest.raises(hip.HipError, match="no CPU fallback"):
        hip.load()
