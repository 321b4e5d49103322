"""CPU: the C-ABI library loads and exports every entry point include/*.h declares (no
compute calls: there may be no GPU here)."""
import ctypes
import os
import re

import pytest

from gem_amd import _lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    inc = os.path.join(ROOT, "include")
    for fn in os.listdir(inc):
        if fn.endswith(".h"):
            txt = open(os.path.join(inc, fn)).read()
            txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
            names.update(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(adz_\w+)\s*\(", txt, flags=re.M))
    return names


def test_header_declares_entry_points():
    names = _declared()
    assert {"adz_vsearch_setup_plus", "adz_tricublin_zyx1_m_n", "adz_tricublin_mono_zyx_m_n",
            "adz_tricublin_zyx3_n"} <= names


def test_library_exports_every_declared_symbol():
    if not os.path.exists(L.LIB_PATH):
        pytest.skip("libadz_hip.so not built")
    so = ctypes.CDLL(L.LIB_PATH)
    missing = [n for n in sorted(_declared()) if not hasattr(so, n)]
    assert not missing, missing


def test_python_mirror_covers_declared_symbols():
    # every declared entry point has a ctypes signature in the host mirror
    import gem_amd.tricublin  # noqa: F401  (registers its signatures)
    try:
        import gem_amd.adz  # noqa: F401
    except ImportError:
        pass
    missing = [n for n in sorted(_declared()) if n not in L.exported_symbols()]
    assert not missing, missing


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(L, "LIB_PATH", str(tmp_path / "nope.so"))
    monkeypatch.setattr(L, "_lib", None)
    with pytest.raises(L.AdzError):
        L.lib()
