"""Loader for libadz_hip.so (the HIP/gfx950 library).  Fails loudly when the library is
missing: there is no CPU fallback anywhere in gem_amd."""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libadz_hip.so")

ADZ_OK, ADZ_EINVAL, ADZ_EHIP, ADZ_ENOMEM, ADZ_ESTATE = 0, 1, 2, 3, 4
ADZ_F64, ADZ_F32, ADZ_BF16 = 0, 1, 2
ADZ_MATH_REF, ADZ_MATH_F32 = 0, 1

_lib = None

# argument signatures of every exported entry point (vp = pointer, i = int, l = long,
# d = double); kept beside include/adz_hip.h
_SIGS = {
    "adz_last_error": ("s", []),
    "adz_version": ("i", []),
    "adz_vsearch_setup_plus": ("i", ["vp", "i", "i", "i", "i", "i", "vp"]),
    "adz_vsearch_set_bounds": ("i", ["vp", "i", "i", "i", "i"]),
    "adz_vsearch_free": ("i", ["vp"]),
    "adz_vcoef": ("i", ["vp", "vp", "vp", "vp", "i", "vp", "i", "vp"]),
    "adz_tricublin_zyx1_m_n": ("i", ["vp", "vp", "vp", "vp", "i", "i", "i", "i", "i", "vp"]),
    "adz_tricublin_mono_zyx_m_n": ("i", ["vp"] * 7 + ["i", "i", "i", "i", "i", "vp"]),
    "adz_tricublin_zyx3_n": ("i", ["vp", "vp", "vp", "vp", "i", "i", "i", "i", "vp"]),
}

_CT = {"vp": C.c_void_p, "i": C.c_int, "l": C.c_long, "d": C.c_double, "s": C.c_char_p}


class AdzError(RuntimeError):
    pass


def extra_sigs(sigs):
    """Register signatures of further entry points (used by the other host modules)."""
    _SIGS.update(sigs)
    if _lib is not None:
        _apply(_lib, sigs)


def _apply(L, sigs):
    for name, (res, args) in sigs.items():
        fn = getattr(L, name)
        fn.restype = _CT[res]
        fn.argtypes = [_CT[a] for a in args]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise AdzError(f"{LIB_PATH} is not built: run `python -m gem_amd.build` (or __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        _apply(L, _SIGS)
        _lib = L
    return _lib


def check(status, what=""):
    if status != ADZ_OK:
        msg = lib().adz_last_error().decode()
        raise AdzError(f"{what} failed (status {status}): {msg}")


def exported_symbols():
    return sorted(_SIGS)
