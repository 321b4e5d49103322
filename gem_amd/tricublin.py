"""Host mirror of the reference's interpolation interface (turbo_tricublin.c /
tricublin_f90.inc) on device-resident torch tensors.

Names, argument order and meaning follow the reference: `Vsearch_setup_plus(levels, nk,
ni, nj, offseti, offsetj)` returns the vertical descriptor; `Tricublin_zyx1_m_n(d, fs,
pxyz, lv, n, m)` interpolates m source fields (fs = list of tensors, each addressed at
f(1,1,1) by an element offset) at n departure points and writes d[m][n].  Where the
reference silently reads out of bounds, these raise or produce NaN (documented per call).
"""
import ctypes as C

import torch

from . import _lib as L

_DT = {torch.float64: L.ADZ_F64, torch.float32: L.ADZ_F32, torch.bfloat16: L.ADZ_BF16}


def _dtype_code(t):
    try:
        return _DT[t.dtype]
    except KeyError:
        raise L.AdzError(f"unsupported dtype {t.dtype}")


def _stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _dev(t):
    if not (isinstance(t, torch.Tensor) and t.is_cuda):
        raise L.AdzError("device (cuda/hip) tensor expected")
    return t


class Ztab:
    """Device vertical descriptor (Vsearch_setup_plus, turbo_tricublin.c:368-380)."""

    def __init__(self, levels, ni, nj, offseti=0, offsetj=0):
        import numpy as np

        lev = np.ascontiguousarray(np.asarray(levels, dtype=np.float64))
        self.nk, self.ni, self.nj, self.offi, self.offj = len(lev), int(ni), int(nj), int(offseti), int(offsetj)
        h = C.c_void_p()
        L.check(L.lib().adz_vsearch_setup_plus(lev.ctypes.data_as(C.c_void_p), self.nk, self.ni, self.nj,
                                               self.offi, self.offj, C.byref(h)), "Vsearch_setup_plus")
        self.h = h

    def set_bounds(self, ilo, ihi, jlo, jhi):
        L.check(L.lib().adz_vsearch_set_bounds(self.h, ilo, ihi, jlo, jhi), "vsearch_set_bounds")
        return self

    def free(self):
        if getattr(self, "h", None) is not None and self.h.value:
            L.lib().adz_vsearch_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def Vsearch_setup_plus(levels, nk, ni, nj, offseti, offsetj):
    assert nk == len(levels)
    return Ztab(levels, ni, nj, offseti, offsetj)


def Vsearch_setup(levels, nk, ni, nj):
    return Ztab(levels, ni, nj, 0, 0)


def _fptrs(fs, base_off):
    arr = (C.c_void_p * len(fs))()
    for k, f in enumerate(fs):
        _dev(f)
        arr[k] = f.data_ptr() + int(base_off) * f.element_size()
    return arr


def vcoef(pxyz, lv, n=None):
    """Index and weights of each target (Vcoef_pxyz4_inline): (ijk, zlinear, coef[n,20])."""
    _dev(pxyz)
    n = pxyz.numel() // 3 if n is None else n
    ijk = torch.empty(n, dtype=torch.int32, device=pxyz.device)
    zl = torch.empty(n, dtype=torch.int32, device=pxyz.device)
    coef = torch.empty(n, 20, dtype=torch.float64, device=pxyz.device)
    L.check(L.lib().adz_vcoef(C.c_void_p(ijk.data_ptr()), C.c_void_p(zl.data_ptr()), C.c_void_p(coef.data_ptr()),
                              C.c_void_p(pxyz.data_ptr()), _dtype_code(pxyz), lv.h, n, _stream()), "vcoef")
    return ijk, zl, coef


def Tricublin_zyx1_m_n(d, fs, pxyz, lv, n, m, base_off=0, math=L.ADZ_MATH_REF):
    """d (m*n,) <- m fields fs at n positions pxyz (3n,).  base_off = element offset of
    f(1,1,1) inside each tensor of fs."""
    _dev(d), _dev(pxyz)
    assert len(fs) == m and d.numel() >= m * n and pxyz.numel() >= 3 * n
    L.check(L.lib().adz_tricublin_zyx1_m_n(C.c_void_p(d.data_ptr()), _fptrs(fs, base_off), C.c_void_p(pxyz.data_ptr()),
                                           lv.h, n, m, _dtype_code(d), _dtype_code(pxyz), math, _stream()),
            "Tricublin_zyx1_m_n")


def Tricublin_mono_zyx_m_n(d, l, mi, ma, fs, pxyz, lv, n, m, base_off=0, math=L.ADZ_MATH_REF):
    for t in (d, l, mi, ma, pxyz):
        _dev(t)
    assert len(fs) == m
    L.check(L.lib().adz_tricublin_mono_zyx_m_n(C.c_void_p(d.data_ptr()), C.c_void_p(l.data_ptr()),
                                               C.c_void_p(mi.data_ptr()), C.c_void_p(ma.data_ptr()),
                                               _fptrs(fs, base_off), C.c_void_p(pxyz.data_ptr()), lv.h, n, m,
                                               _dtype_code(d), _dtype_code(pxyz), math, _stream()),
            "Tricublin_mono_zyx_m_n")


def Tricublin_zyx3_n(d, f3, pxyz, lv, n, base_off=0, math=L.ADZ_MATH_REF):
    """d (3n,) <- interleaved f3(3,...) at n positions; base_off = element offset of
    f3(1,1,1,1)."""
    for t in (d, f3, pxyz):
        _dev(t)
    L.check(L.lib().adz_tricublin_zyx3_n(C.c_void_p(d.data_ptr()), C.c_void_p(f3.data_ptr() + int(base_off) * f3.element_size()),
                                         C.c_void_p(pxyz.data_ptr()), lv.h, n, _dtype_code(d), _dtype_code(pxyz),
                                         math, _stream()), "Tricublin_zyx3_n")
