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


@pytest.mark.parametrize("case", ["sweep0", "sweep5", "outside", "null", "rows"])
def test_ilmc_create_refuses_bad_arguments(case):
    """adz_ilmc_create validates before any device call (safe without a GPU)."""
    if not os.path.exists(L.LIB_PATH):
        pytest.skip("libadz_hip.so not built")
    from gem_amd import adz as A

    dom = dict(minx=-1, maxx=12, miny=-1, maxy=10, nk=8, i0=1, in_=10, j0=1, jn=8, k0=2, sweep_max=2, min_max=1,
               autobar=0)
    tr = (A.AdzIlmcTracer * 1)()
    tr[0].f, tr[0].mn, tr[0].mx = 4096, 8192, 12288  # never dereferenced: validation fails first
    if case == "sweep0":
        dom["sweep_max"] = 0
    elif case == "sweep5":
        dom["sweep_max"] = 5
    elif case == "outside":
        dom["in_"] = 13
    elif case == "null":
        tr[0].mx = None
    else:
        dom.update(maxy=70000, jn=65535, nk=300, k0=1)
    d = A.AdzIlmcDomain(**dom)
    h = ctypes.c_void_p()
    assert L.lib().adz_ilmc_create(tr, 1, ctypes.byref(d), ctypes.byref(h)) == 1
    assert "adz_ilmc_create" in L.lib().adz_last_error().decode()
    assert not h.value


@pytest.mark.parametrize("case", ["stage9", "stage8_ref", "xcd2"])
def test_plan_create_refuses_bad_tiles(case):
    """adz_plan_create checks the tile parameters before it reads an item or touches the
    device: staging modes run 0..8, staging mode 8 (16-wave blocks) exists for FP32 fields in
    the FP32 arithmetic only, XCD chunking is 0 or 1."""
    if not os.path.exists(L.LIB_PATH):
        pytest.skip("libadz_hip.so not built")
    from gem_amd import adz as A

    items = (A.AdzItem * 1)()
    items[0].pdtype = L.ADZ_F32  # the rest is never read: the tile check fails first
    t = A.AdzTune(fields_per_pass=4, points_outer=0, blocks_per_cu=1, xcd_chunks=1, stage=8, level_inner=1)
    math, want = L.ADZ_MATH_F32, "stage must be 0..8"
    if case == "stage9":
        t.stage = 9
    elif case == "stage8_ref":
        math, want = L.ADZ_MATH_REF, "stage 8 needs the FP32 arithmetic"
    else:
        t.xcd_chunks, want = 2, "xcd_chunks must be 0 or 1"
    h = ctypes.c_void_p()
    assert L.lib().adz_plan_create(items, 1, L.ADZ_F32, math, 0, ctypes.byref(t), ctypes.byref(h)) == 1
    assert want in L.lib().adz_last_error().decode()
    assert not h.value
