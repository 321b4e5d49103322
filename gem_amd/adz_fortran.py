"""The Fortran drop-in layer (gem_amd/fortran/adz_gpu.F90, module adz_gpu) driven from
Python through its C-callable entry points (adz_gpu_capi.F90, libadz_gpu.so).

`FortranStep` sets the module state the way GEM's modules hold it -- winds, the right-hand
sides of adz_main, tracers_P / tracers_M, the tracer groups of tr3d, the metas and Adz_post
arrays of the groups with post-processing, the fixer's and set_post_tr_hlt's inputs -- as
device addresses of torch tensors, and runs the reference's step routines on them:
adz_main() (adz_main.F90:16-71) and adz_tracers(F_before_psadj_L) (adz_tracers.F90:16-184).
The tests compare this layer with the C ABI and the oracle, and bench.py times it.  A
missing library raises: there is no fallback.
"""
import ctypes as C
import os

import torch

from . import _lib as L
from . import adz as A

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fortran", "libadz_gpu.so")
_flib = None
_current = [0]  # generation of the FortranStep that holds the module's state (one at a time)

_SIGS = {
    "adz_f_set": (None, ["vp", "i", "vp", "i", "i", "i"]),
    "adz_f_finalize": (None, []),
    "adz_f_set_stream": (None, ["vp"]),
    "adz_f_set_time": (None, ["d", "d", "d", "i"]),
    "adz_f_set_winds": (None, ["vp"] * 6),
    "adz_f_set_rhs": (None, ["vp", "vp", "i", "i"]),
    "adz_f_set_tracers": (None, ["i", "vp", "vp", "vp"]),
    "adz_f_set_groups": (None, ["vp", "vp"]),
    "adz_f_set_wp": (None, ["i", "i", "vp", "vp", "vp", "vp", "vp", "vp"]),
    "adz_f_set_bc_fields": (None, ["vp"] * 4),
    "adz_f_set_airmass": (None, ["vp"] * 5 + ["i", "i"]),
    "adz_f_set_options": (None, ["i"] * 4),
    "adz_f_set_area_mask": (None, ["vp"]),
    "adz_f_set_tune": (None, ["i", "vp"]),
    "adz_f_inittraj": (None, []),
    "adz_f_main": (None, []),
    "adz_f_tracers": (None, ["i"]),
    "adz_f_step": (None, []),
    "adz_f_deficits": (None, ["i", "i", "vp"]),
    "adz_f_set_deficits": (None, ["i", "i", "vp"]),
    "adz_f_info": (None, ["vp"]),
    "adz_f_ctx": ("vp", []),
}
_CT = {"vp": C.c_void_p, "i": C.c_int, "d": C.c_double}


def flib():
    global _flib
    if _flib is None:
        L.lib()  # libadz_hip.so first: libadz_gpu.so resolves its symbols there
        if not os.path.exists(LIB_PATH):
            raise L.AdzError(f"{LIB_PATH} is not built: run `python gem_amd/fortran/build.py` "
                             "(or __graft_entry__.build())")
        F = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(F, name)
            fn.restype = None if res is None else _CT[res]
            fn.argtypes = [_CT[a] for a in args]
        _flib = F
    return _flib


def _p(t, off=0):
    return None if t is None else t.data_ptr() + int(off) * t.element_size()


def _parr(ptrs):
    return (C.c_void_p * max(1, len(ptrs)))(*ptrs)


class FortranStep:
    """Module adz_gpu on one tile (the module holds one tile: a new FortranStep replaces the
    state of the previous one).  grid: the dict of synth.grid_desc; layout: (l_minx,
    l_maxx, l_miny, l_maxy) of the model arrays; offsets are element offsets of (1,1,1)
    inside the flat tensors (extended layout for winds and right-hand-side sources, model
    layout for the rest)."""

    def __init__(self, grid, layout, itraj=3, pos_dtype=L.ADZ_F32, field_dtype=L.ADZ_F32, math=L.ADZ_MATH_REF):
        F = flib()
        self._g = A.AdzGrid.from_dict(grid)
        lay = (C.c_int * 4)(*layout)
        F.adz_f_set(C.byref(self._g), int(itraj), lay, int(pos_dtype), int(field_dtype), int(math))
        _current[0] += 1
        self._gen = _current[0]  # a newer FortranStep replaces this one's module state
        self.layout = tuple(layout)
        self.field_dtype = field_dtype
        self._keep = []
        self.ctx = A.AdzContext.view(F.adz_f_ctx(), grid, pos_dtype)  # the module's context, not owned here
        self.set_stream(torch.cuda.current_stream())

    def set_stream(self, stream):
        flib().adz_f_set_stream(C.c_void_p(stream.cuda_stream))

    def set_time(self, dt, bA_m=0.6, bA=0.6, advec=1):
        flib().adz_f_set_time(float(dt), float(bA_m), float(bA), int(advec))

    def set_winds(self, w, off=0):
        flib().adz_f_set_winds(*[_p(w[k], off) for k in ("uu_ext", "vv_ext", "ww_ext", "ut0", "vt0", "zdt0")])
        self._keep.append(w)

    def set_rhs(self, src, dst, src_off, dst_off, hydro=True, hauteur=False):
        """src / dst: lists of 6 tensors (u, v, c, t, f, w; w may be None)"""
        flib().adz_f_set_rhs(_parr([_p(s, src_off) for s in src]), _parr([_p(d, dst_off) for d in dst]),
                             1 if hydro else 0, 1 if hauteur else 0)
        self._keep.append((src, dst))

    def set_tracers(self, P, M, B=None, off=0):
        n = len(P)
        B = B if B is not None else [None] * n
        flib().adz_f_set_tracers(n, _parr([_p(t, off) for t in P]), _parr([_p(t, off) for t in M]),
                                 _parr([_p(t, off) for t in B]))
        self._keep.append((P, M, B))

    def set_groups(self, tricub_nt=(1, 0), bichqv_nt=(1, 0), tricub_wp=(1, 0), bichqv_wp=(1, 0)):
        """(deb, ntr) of each group, deb 1-based into the tracer lists"""
        g = [tricub_nt, bichqv_nt, tricub_wp, bichqv_wp]
        flib().adz_f_set_groups((C.c_int * 4)(*[a for a, _ in g]), (C.c_int * 4)(*[b for _, b in g]))

    def set_wp(self, which, mass, mono, post, off=0, deficit=None):
        """which 1: Tr_3CWP / Adz_post_3CWP, 2: Tr_BQWP / Adz_post_BQWP; post: list of dicts
        with lin, min, max tensors"""
        n = len(mass)
        dfc = deficit if deficit is not None else [0.0] * n
        flib().adz_f_set_wp(int(which), n, (C.c_int * n)(*mass), (C.c_int * n)(*mono), (C.c_double * n)(*dfc),
                            _parr([_p(q["lin"], off) for q in post]), _parr([_p(q["min"], off) for q in post]),
                            _parr([_p(q["max"], off) for q in post]))
        self._keep.append(post)

    def set_bc_fields(self, airm1, airm0, pkps, mask, off, mask_off):
        flib().adz_f_set_bc_fields(_p(airm1, off), _p(airm0, off), _p(pkps, off), _p(mask, mask_off))
        self._keep.append((airm1, airm0, pkps, mask))

    def set_airmass(self, pm_p, pm_m, pt_m, p0_m, area, off, area_off, hu=0, dry=False):
        flib().adz_f_set_airmass(_p(pm_p, off), _p(pm_m, off), _p(pt_m, off), _p(p0_m, area_off),
                                 _p(area, area_off), int(hu), 1 if dry else 0)
        self._keep.append((pm_p, pm_m, pt_m, p0_m, area))

    def set_area_mask(self, area_mask, off):
        """geomh_area_mask_8 (FP64, 2-D), the weight of the masses under Schm_autobar_L"""
        flib().adz_f_set_area_mask(_p(area_mask, off))
        self._keep.append(area_mask)

    def set_options(self, bc_min_max=True, ilmc_sweep_max=2, ilmc_min_max=True, autobar=False):
        flib().adz_f_set_options(int(bc_min_max), int(ilmc_sweep_max), int(ilmc_min_max), int(autobar))

    def set_tune(self, which, tune):
        t = A.AdzTune(*[int(tune.get(k, 0)) for k, _ in A.AdzTune._fields_])
        flib().adz_f_set_tune(int(which), C.byref(t))

    def adz_inittraj(self):
        flib().adz_f_inittraj()

    def adz_main(self):
        flib().adz_f_main()

    def adz_tracers(self, before_psadj):
        flib().adz_f_tracers(1 if before_psadj else 0)

    def step(self):
        """adz_main, adz_tracers(.true.), adz_tracers(.false.)"""
        flib().adz_f_step()

    def deficits(self, which, n):
        out = (C.c_double * n)()
        flib().adz_f_deficits(int(which), n, out)
        return list(out)

    def set_deficits(self, which, values):
        """new carried deficits of a group (synchronous; not inside a capture)"""
        n = len(values)
        flib().adz_f_set_deficits(int(which), n, (C.c_double * n)(*values))

    def array(self, which):
        return self.ctx.array(which)

    def ztab(self, thermo):
        return self.ctx.ztab(thermo)

    def free(self):
        if getattr(self, "_keep", None) is not None:
            self.ctx.h = None
            if self._gen == _current[0]:
                flib().adz_f_finalize()
            self._keep = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
