"""Host mirror of GEM's adz_* layer (gemdyn/src/base/adz_*.F90) over libadz_hip.

`AdzContext(grid)` is adz_set (adz_set.F90:16-311): it owns the SL state that the
Fortran keeps in module adz_mem (positions, derived winds, tables) in device memory.
Methods keep the Fortran names: adz_inittraj(), adz_traject(dt, ...).  Arguments that
are arrays are torch device tensors.
"""
import ctypes as C

import numpy as np
import torch

from . import _lib as L

L.extra_sigs({
    "adz_set": ("i", ["vp", "i", "i", "vp"]),
    "adz_ctx_free": ("i", ["vp"]),
    "adz_ctx_info": ("i", ["vp", "vp", "vp"]),
    "adz_ctx_table": ("i", ["vp", "i", "vp"]),
    "adz_ctx_array": ("i", ["vp", "i", "vp", "vp"]),
    "adz_ctx_ztab": ("vp", ["vp", "i"]),
    "adz_inittraj": ("i", ["vp", "vp"]),
    "adz_traject": ("i", ["vp", "d", "i", "d", "d"] + ["vp"] * 6 + ["vp"]),
    "adz_ctx_state_bytes": ("i", ["vp", "vp"]),
    "adz_ctx_save_state": ("i", ["vp", "vp", "l", "vp"]),
    "adz_ctx_load_state": ("i", ["vp", "vp", "l", "vp"]),
})

ADZ_A_PXYZM, ADZ_A_WPXYZ, ADZ_A_UVW_DEP, ADZ_A_PM, ADZ_A_PMU, ADZ_A_PMV, ADZ_A_PT, ADZ_A_PB, ADZ_A_UVW_D, \
    ADZ_A_WARR = range(10)

INFO_INTS = ["lminx", "lmaxx", "lminy", "lmaxy", "nit", "njt", "nij", "i0", "in", "j0", "jn", "i0u", "inu", "j0v",
             "jnv", "i0b", "inb", "j0b", "jnb", "k0", "k0t", "k0m", "pnz_m", "pnz_t"]
INFO_DBLS = ["iminposx", "imaxposx", "iminposy", "imaxposy", "sig", "ovdzm", "ovdzt", "vw5", "vw6"]


class AdzGrid(C.Structure):
    """struct adz_grid (include/adz_hip.h)."""
    _fields_ = [(n, C.c_int) for n in ("ni", "nj", "nk", "i0g", "j0g", "halox", "haloy", "west", "east", "south",
                                       "north", "pil_w", "pil_e", "pil_s", "pil_n", "gbpil_t", "bc_lam_flux", "niu",
                                       "njv")] + \
              [("inv_hx", C.c_double), ("inv_hy", C.c_double), ("inv_rayt", C.c_double)] + \
              [(n, C.c_void_p) for n in ("zm", "zt", "zx", "lat")] + \
              [(n, C.c_double) for n in ("zmin", "zmax", "wmstar_nk", "idz_t_nkm1")]

    @classmethod
    def from_dict(cls, g):
        s = cls()
        keep = {}
        for name, ctype in cls._fields_:
            if ctype is C.c_void_p:
                a = np.ascontiguousarray(g[name], dtype=np.float64)
                keep[name] = a
                setattr(s, name, a.ctypes.data)
            else:
                setattr(s, name, g[name])
        s._keep = keep  # host arrays must outlive the struct
        return s


def _stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


class AdzContext:
    """adz_set + the device state of module adz_mem for one tile."""

    def __init__(self, grid, pos_dtype=L.ADZ_F32, itraj=3):
        self.grid = grid
        self.pos_dtype = pos_dtype
        self._g = AdzGrid.from_dict(grid)
        h = C.c_void_p()
        L.check(L.lib().adz_set(C.byref(self._g), pos_dtype, itraj, C.byref(h)), "adz_set")
        self.h = h
        iv = np.zeros(24, np.int32)
        dv = np.zeros(9)
        L.check(L.lib().adz_ctx_info(self.h, iv.ctypes.data, dv.ctypes.data), "adz_ctx_info")
        self.info = dict(zip(INFO_INTS, iv.tolist()))
        self.info.update(zip(INFO_DBLS, dv.tolist()))

    def table(self, which):
        nk = self.grid["nk"]
        if which in (0, 1):
            out = np.zeros(self.info["pnz_m" if which == 0 else "pnz_t"], np.int32)
        elif which == 10:
            out = np.zeros(self.info["njt"])
        else:
            out = np.zeros(nk + 1)
        L.check(L.lib().adz_ctx_table(self.h, which, out.ctypes.data), "adz_ctx_table")
        return out

    def array(self, which):
        """Device state array as a torch tensor view (no copy)."""
        p = C.c_void_p()
        n = C.c_long()
        L.check(L.lib().adz_ctx_array(self.h, which, C.byref(p), C.byref(n)), "adz_ctx_array")
        dt = torch.float64 if (which == ADZ_A_WPXYZ or self.pos_dtype == L.ADZ_F64) else torch.float32
        return _wrap(p.value, n.value, dt)

    def ztab(self, thermo):
        return L.lib().adz_ctx_ztab(self.h, 1 if thermo else 0)

    def adz_inittraj(self):
        L.check(L.lib().adz_inittraj(self.h, _stream()), "adz_inittraj")

    def adz_traject(self, dt, uu_ext, vv_ext, ww_ext, ut0, vt0, zdt0, schm_advec=1, bA_m=0.6, bA=0.6):
        args = [uu_ext, vv_ext, ww_ext, ut0, vt0, zdt0]
        for a in args:
            assert a.is_cuda and a.dtype == torch.float32 and a.is_contiguous()
        L.check(L.lib().adz_traject(self.h, float(dt), int(schm_advec), float(bA_m), float(bA),
                                    *[C.c_void_p(a.data_ptr()) for a in args], _stream()), "adz_traject")

    def save_state(self):
        """The SL state a restart needs (Adz_pxyzm, Adz_wpxyz, the next iteration count) as
        a host byte array (adz_set.F90:343-360)."""
        n = C.c_size_t()
        L.check(L.lib().adz_ctx_state_bytes(self.h, C.byref(n)), "adz_ctx_state_bytes")
        buf = np.zeros(n.value, np.uint8)
        L.check(L.lib().adz_ctx_save_state(self.h, buf.ctypes.data, n.value, _stream()), "adz_ctx_save_state")
        return buf

    def load_state(self, buf):
        buf = np.ascontiguousarray(buf, dtype=np.uint8)
        L.check(L.lib().adz_ctx_load_state(self.h, buf.ctypes.data, buf.size, _stream()), "adz_ctx_load_state")

    @classmethod
    def view(cls, h, grid, pos_dtype):
        """A context another owner created (e.g. the Fortran layer's): arrays and tables can be
        read through it; it is never freed from here."""
        v = cls.__new__(cls)
        v.grid, v.pos_dtype, v.h, v._owner = grid, pos_dtype, C.c_void_p(h), False
        iv = np.zeros(24, np.int32)
        dv = np.zeros(9)
        L.check(L.lib().adz_ctx_info(v.h, iv.ctypes.data, dv.ctypes.data), "adz_ctx_info")
        v.info = dict(zip(INFO_INTS, iv.tolist()))
        v.info.update(zip(INFO_DBLS, dv.tolist()))
        return v

    def free(self):
        if getattr(self, "_owner", True) and getattr(self, "h", None) is not None and self.h.value:
            L.lib().adz_ctx_free(self.h)
        self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class _DevBuf:
    """Minimal __cuda_array_interface__ provider to view library-owned device memory."""

    def __init__(self, ptr, n, typestr):
        self.__cuda_array_interface__ = {"shape": (n,), "typestr": typestr, "data": (ptr, False), "version": 2,
                                         "strides": None}


def _wrap(ptr, n, dt):
    typestr = "<f8" if dt == torch.float64 else "<f4"
    return torch.as_tensor(_DevBuf(ptr, n, typestr), device="cuda")


# ---------------------------------------------------------------------- adz_tricub_hlt
L.extra_sigs({
    "adz_tricub_hlt": ("i", ["vp", "i", "vp", "i", "vp"] + ["i"] * 10 + ["i", "i", "i", "i", "vp"]),
    "adz_plan_create": ("i", ["vp", "i", "i", "i", "i", "vp", "vp"]),
    "adz_plan_execute": ("i", ["vp", "vp"]),
    "adz_plan_tune": ("i", ["vp", "vp"]),
    "adz_plan_status": ("i", ["vp", "vp"]),
    "adz_plan_set_spin_limit": ("i", ["vp", "i"]),
    "adz_plan_free": ("i", ["vp"]),
    "adz_plan_autotune": ("i", ["vp", "i", "i", "i", "i", "i", "vp", "vp", "vp"]),
})

ADZ_POST_STORE, ADZ_POST_CLIP = 1, 2
_DT = {torch.float64: L.ADZ_F64, torch.float32: L.ADZ_F32, torch.bfloat16: L.ADZ_BF16}


class AdzField(C.Structure):
    _fields_ = [("src", C.c_void_p), ("dst", C.c_void_p), ("lin", C.c_void_p), ("mn", C.c_void_p),
                ("mx", C.c_void_p)]


class AdzItem(C.Structure):
    _fields_ = [("geom", C.c_void_p), ("xyz", C.c_void_p), ("pdtype", C.c_int)] + \
               [(n, C.c_int) for n in ("i0", "in_", "j0", "jn", "k0", "nk", "dminx", "dmaxx", "dminy", "dmaxy",
                                       "ext")] + [("fld", C.c_void_p), ("nptr", C.c_int)]


class AdzTune(C.Structure):
    _fields_ = [("fields_per_pass", C.c_int), ("points_outer", C.c_int), ("blocks_per_cu", C.c_int),
                ("xcd_chunks", C.c_int), ("stage", C.c_int), ("level_inner", C.c_int)]


def _ptr_at(t, off):
    return None if t is None else t.data_ptr() + int(off) * t.element_size()


def field_list(stack, src_off, dst_off):
    """stack: list of dicts with tensors 'src', 'dst' and optional 'lin','min','max'
    (the Adz_pntr_stack / Adz_post entries); offsets address element (1,1,1)."""
    arr = (AdzField * len(stack))()
    for n, e in enumerate(stack):
        arr[n].src = _ptr_at(e["src"], src_off)
        arr[n].dst = _ptr_at(e["dst"], dst_off)
        for k, a in (("lin", "lin"), ("mn", "min"), ("mx", "max")):
            setattr(arr[n], k, _ptr_at(e.get(a), dst_off))
    return arr


def adz_tricub_hlt(stack, xyz, geom, i0, in_, j0, jn, k0, nk, dlayout, ext=True, post=0, math=L.ADZ_MATH_REF,
                   src_off=None):
    """adz_tricub_hlt(F_stk, F_nptr, F_xyz, F_geom, F_num, F_i0, F_in, F_j0, F_jn, F_k0
    [,F_ext_L, F_post]).  dlayout = (dminx, dmaxx, dminy, dmaxy) of the destination (model)
    arrays; geom = an adz_ztab handle (e.g. AdzContext.ztab(thermo)); src_off = element
    offset of src(1,1,1) (default: the extended layout of the context)."""
    dminx, dmaxx, dminy, dmaxy = dlayout
    dst_off = (1 - dminx) + (dmaxx - dminx + 1) * (1 - dminy)
    if src_off is None:
        src_off = dst_off
    fl = field_list(stack, src_off, dst_off)
    fdt = _DT[stack[0]["dst"].dtype]
    pdt = _DT[xyz.dtype]
    L.check(L.lib().adz_tricub_hlt(fl, len(stack), C.c_void_p(xyz.data_ptr()), pdt, geom, i0, in_, j0, jn, k0, nk,
                                   dminx, dmaxx, dminy, dmaxy, 1 if ext else 0, post, fdt, math, _stream()),
            "adz_tricub_hlt")


L.extra_sigs({"adz_tricub_hlt_qv": ("i", ["vp", "vp", "i", "vp"] + ["i"] * 11 + ["vp"])})


def adz_tricub_hlt_qv(ctx, stack, xyz, i0, in_, j0, jn, k0, dlayout, ext=True, post=0, src_off=None):
    """adz_tricub_hlt(..., F_QV_L=.true.): bicubic horizontal, quintic vertical
    (bicubHQV_lin_mono.F90:20-172) at the positions xyz (storage of the context) on the
    thermodynamic levels of ctx; FP32 fields."""
    dminx, dmaxx, dminy, dmaxy = dlayout
    dst_off = (1 - dminx) + (dmaxx - dminx + 1) * (1 - dminy)
    if src_off is None:
        i = ctx.info
        src_off = (1 - i["lminx"]) + i["nit"] * (1 - i["lminy"]) if ext else dst_off
    for e in stack:
        _check_types(torch.float32, **{k: e.get(k) for k in ("src", "dst", "lin", "min", "max")})
    _check_types(torch.float32 if ctx.pos_dtype == L.ADZ_F32 else torch.float64, xyz=xyz)
    fl = field_list(stack, src_off, dst_off)
    L.check(L.lib().adz_tricub_hlt_qv(ctx.h, fl, len(stack), C.c_void_p(xyz.data_ptr()), i0, in_, j0, jn, k0,
                                      dminx, dmaxx, dminy, dmaxy, 1 if ext else 0, post, _stream()),
            "adz_tricub_hlt_qv")


def _item_array(items):
    """adz_item array for a list of dicts {geom, xyz, box=(i0,in,j0,jn,k0,nk),
    dlayout=(dminx,dmaxx,dminy,dmaxy), ext, fields=field_list(...)}; returns the array
    and the field lists it points into (keep them alive while the array is used)."""
    keep = []
    arr = (AdzItem * len(items))()
    for q, it in enumerate(items):
        fl = it["fields"]
        keep.append(fl)
        a = arr[q]
        a.geom = it["geom"]
        a.xyz = it["xyz"].data_ptr()
        a.pdtype = _DT[it["xyz"].dtype]
        a.i0, a.in_, a.j0, a.jn, a.k0, a.nk = it["box"]
        a.dminx, a.dmaxx, a.dminy, a.dmaxy = it["dlayout"]
        a.ext = 1 if it.get("ext", True) else 0
        a.fld = C.cast(fl, C.c_void_p)
        a.nptr = len(fl)
    return arr, keep


class Plan:
    """A persistent batched launch over several subdomains (adz_plan_*)."""

    def __init__(self, items, fdtype, math=L.ADZ_MATH_F32, post=0, tune=None):
        arr, self._keep = _item_array(items)
        t = None
        if tune is not None:
            t = AdzTune(*[int(tune.get(k, 0)) for k, _ in AdzTune._fields_])
        h = C.c_void_p()
        L.check(L.lib().adz_plan_create(arr, len(items), fdtype, math, post, C.byref(t) if t else None,
                                        C.byref(h)), "adz_plan_create")
        self.h = h

    @staticmethod
    def autotune(items, fdtype, math=L.ADZ_MATH_F32, post=0, reps=3):
        """adz_plan_autotune: the fastest tile parameters for these items (the items'
        destinations are written).  Returns (tune dict, ms per execution)."""
        arr, keep = _item_array(items)
        t, ms = AdzTune(), C.c_float()
        L.check(L.lib().adz_plan_autotune(arr, len(items), fdtype, math, post, reps, _stream(), C.byref(t),
                                          C.byref(ms)), "adz_plan_autotune")
        del keep
        return {k: getattr(t, k) for k, _ in AdzTune._fields_}, ms.value

    def execute(self):
        L.check(L.lib().adz_plan_execute(self.h, _stream()), "adz_plan_execute")

    def tune(self):
        t = AdzTune()
        L.check(L.lib().adz_plan_tune(self.h, C.byref(t)), "adz_plan_tune")
        return {k: getattr(t, k) for k, _ in AdzTune._fields_}

    def status(self):
        """Staging waits that gave up since the last call (adz_plan_status; synchronises)."""
        n = C.c_int()
        L.check(L.lib().adz_plan_status(self.h, C.byref(n)), "adz_plan_status")
        return n.value

    def set_spin_limit(self, polls):
        L.check(L.lib().adz_plan_set_spin_limit(self.h, int(polls)), "adz_plan_set_spin_limit")

    def free(self):
        if getattr(self, "h", None) is not None and self.h.value:
            L.lib().adz_plan_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


# ---------------------------------------------------------------------- Bermejo-Conde
L.extra_sigs({
    "adz_bc_create": ("i", ["vp", "i", "vp", "i", "vp"]),
    "adz_bermejo_conde": ("i", ["vp"] * 7 + ["vp"]),
    "adz_bermejo_conde_lam": ("i", ["vp"] * 8 + ["vp"]),
    "adz_bc_free": ("i", ["vp"]),
    "adz_bermejo_conde_static": ("i", ["vp", "i", "vp", "i"] + ["vp"] * 7 + ["vp"]),
    "adz_flux_create": ("i", ["vp", "i", "vp", "vp", "i", "i", "vp"]),
    "adz_bc_lam_flux": ("i", ["vp"] * 5 + ["vp"]),
    "adz_flux_active": ("i", ["vp", "vp"]),
    "adz_flux_free": ("i", ["vp"]),
})


class AdzBcTracer(C.Structure):
    _fields_ = [("p", C.c_void_p), ("m", C.c_void_p), ("lin", C.c_void_p), ("mn", C.c_void_p), ("mx", C.c_void_p),
                ("legacy", C.c_int), ("weight", C.c_int), ("pexp_n", C.c_int)]


class AdzBcDomain(C.Structure):
    _fields_ = [(n, C.c_int) for n in ("minx", "maxx", "miny", "maxy", "nk", "i0", "in_", "j0", "jn", "k0",
                                       "min_max")] + [("pexp_list", C.c_float * 9), ("autobar", C.c_int)]


def _check_types(dtype, **arrays):
    """The library takes plain pointers: refuse arrays of another storage type."""
    for name, a in arrays.items():
        if a is not None and a.dtype != dtype:
            raise TypeError(f"{name} is {a.dtype}, expected {dtype}")


def bc_settings(mass):
    """Tracer `mass=` code to (legacy, weight, pexp_n), adz_post_tr.F90:190-210."""
    if mass == 1:
        return 1, 1, 0
    weight = (mass - 100) // 10
    return 0, weight, mass - (100 + weight * 10)


class BermejoConde:
    """Bermejo_Conde(F_bc, F_ntr_bc, F_i0, F_in, F_j0, F_jn, F_k0) on device arrays.

    tracers: list of dicts with tensors p, m, lin, min, max (model layout, element (1,1,1)
    at offset `off`) and the tracer's `mass` code (1 or 100+10*weight+pexp_n).  autobar:
    Schm_autobar_L (the masses are level-1 sums with geomh_area_mask_8 given as `mask`)."""

    def __init__(self, tracers, layout, scope, off, min_max=True, pexp_list=(1.0,) * 9, autobar=False):
        minx, maxx, miny, maxy, nk = layout
        i0, in_, j0, jn, k0 = scope
        arr = (AdzBcTracer * len(tracers))()
        tdt = None
        for n, t in enumerate(tracers):
            legacy, weight, pexp_n = bc_settings(t["mass"])
            a = arr[n]
            a.p, a.m = _ptr_at(t["p"], off), _ptr_at(t["m"], off)
            a.lin, a.mn, a.mx = _ptr_at(t["lin"], off), _ptr_at(t["min"], off), _ptr_at(t["max"], off)
            a.legacy, a.weight, a.pexp_n = legacy, weight, pexp_n
            tdt = _DT[t["m"].dtype]
        dom = AdzBcDomain(minx, maxx, miny, maxy, nk, i0, in_, j0, jn, k0, 1 if min_max else 0,
                          (C.c_float * 9)(*pexp_list), 1 if autobar else 0)
        h = C.c_void_p()
        L.check(L.lib().adz_bc_create(arr, len(tracers), C.byref(dom), tdt, C.byref(h)), "adz_bc_create")
        self.h, self.off, self.ntr = h, off, len(tracers)
        self.dtype = tracers[0]["m"].dtype
        self.mask_off = (1 - minx) + (maxx - minx + 1) * (1 - miny)

    def __call__(self, airm1, airm0, mask, deficit, pkps=None, diag=None, flux=None):
        """flux: FP64 tensor [2*ntr] of (mass_fo, mass_fi) from BcLamFlux (LAM Aranami flux)."""
        _check_types(self.dtype, airm1=airm1, airm0=airm0, pkps=pkps)
        _check_types(torch.float64, mask=mask, deficit=deficit, diag=diag, flux=flux)
        if flux is not None and flux.numel() < 2 * self.ntr:
            raise ValueError("flux needs 2 values per tracer")
        args = (self.h, _ptr_at(airm1, self.off), _ptr_at(airm0, self.off), _ptr_at(pkps, self.off),
                _ptr_at(mask, self.mask_off))
        tail = (deficit.data_ptr(), None if diag is None else diag.data_ptr(), _stream())
        if flux is None:
            L.check(L.lib().adz_bermejo_conde(*args, *tail), "adz_bermejo_conde")
        else:
            L.check(L.lib().adz_bermejo_conde_lam(*args, flux.data_ptr(), *tail), "adz_bermejo_conde_lam")

    def free(self):
        if getattr(self, "h", None) is not None and self.h.value:
            L.lib().adz_bc_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


# ---------------------------------------------------------------------- ILMC (mono = 2)
L.extra_sigs({
    "adz_ilmc_create": ("i", ["vp", "i", "vp", "vp"]),
    "adz_ilmc_lam": ("i", ["vp", "vp", "vp", "vp"]),
    "adz_ilmc_free": ("i", ["vp"]),
})


class AdzIlmcTracer(C.Structure):
    _fields_ = [("f", C.c_void_p), ("mn", C.c_void_p), ("mx", C.c_void_p)]


class AdzIlmcDomain(C.Structure):
    _fields_ = [(n, C.c_int) for n in ("minx", "maxx", "miny", "maxy", "nk", "i0", "in_", "j0", "jn", "k0",
                                       "sweep_max", "min_max", "autobar")]


class IlmcLam:
    """ILMC_LAM (ILMC_LAM.F90:18-273, ilmc_lam_loop.inc:16-139): ILMC shape preservation of
    the tracers whose mono code is 2 (adz_post_tr.F90:104), on device arrays.

    tracers: list of dicts with FP32 tensors `f` (the interpolated tracer, corrected in
    place), `min`, `max` (Adz_post min/max), model layout with element (1,1,1) at offset
    `off`.  box = (i0, in, j0, jn, k0): Adz_i0..Adz_in x Adz_j0..Adz_jn x Adz_k0t..nk.
    sweep_max, min_max: Adz_ILMC_sweep_max, Adz_ILMC_min_max_L; autobar: Schm_autobar_L."""

    def __init__(self, tracers, layout, box, off, sweep_max=2, min_max=True, autobar=False):
        minx, maxx, miny, maxy, nk = layout
        i0, in_, j0, jn, k0 = box
        arr = (AdzIlmcTracer * len(tracers))()
        for n, t in enumerate(tracers):
            _check_types(torch.float32, f=t["f"], min=t["min"], max=t["max"])
            arr[n].f, arr[n].mn, arr[n].mx = _ptr_at(t["f"], off), _ptr_at(t["min"], off), _ptr_at(t["max"], off)
        dom = AdzIlmcDomain(minx, maxx, miny, maxy, nk, i0, in_, j0, jn, k0, sweep_max, 1 if min_max else 0,
                            1 if autobar else 0)
        h = C.c_void_p()
        L.check(L.lib().adz_ilmc_create(arr, len(tracers), C.byref(dom), C.byref(h)), "adz_ilmc_create")
        self.h, self.off, self.ntr = h, off, len(tracers)

    def __call__(self, air_m, counts=None):
        """air_m: FP32 air mass at time M (airm0).  counts: optional int64 tensor [3*ntr]
        receiving per tracer the sweeps that moved mass, min resets and max resets."""
        _check_types(torch.float32, air_m=air_m)
        _check_types(torch.int64, counts=counts)
        if counts is not None and counts.numel() < 3 * self.ntr:
            raise ValueError("counts needs 3 values per tracer")
        L.check(L.lib().adz_ilmc_lam(self.h, _ptr_at(air_m, self.off), None if counts is None else counts.data_ptr(),
                                     _stream()), "adz_ilmc_lam")

    def free(self):
        if getattr(self, "h", None) is not None and self.h.value:
            L.lib().adz_ilmc_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


# ---------------------------------------------------------------------- LAM boundary flux
class AdzFluxTracer(C.Structure):
    _fields_ = [("src", C.c_void_p), ("fo", C.c_void_p), ("fi", C.c_void_p)]


class AdzFluxDomain(C.Structure):
    _fields_ = [(n, C.c_int) for n in ("minx", "maxx", "miny", "maxy", "nk", "ni", "nj", "i0b", "inb", "j0b", "jnb",
                                       "ic0", "ic1", "jc0", "jc1", "kc0", "autobar")]


class BcLamFlux:
    """adz_BC_LAM_Aranami_hlt (adz_BC_LAM_Aranami_hlt.F90:18-236) for the fixer tracers,
    with the flux masses of mass_tr_PM (mass_tr_PM.F90:101-127).

    tracers: list of dicts with `src` (the tracer interpolated) and optionally `fo`, `fi`
    (Adz_flux%fo / %fi, model layout, element (1,1,1) at offset `off`).  layout =
    (minx, maxx, miny, maxy, nk); tile = (ni, nj); bbox = (i0b, inb, j0b, jnb) of the
    Adz_pb targets; core = (i0, in, j0, jn, k0) of the fixer's scope.  geom: the thermo
    descriptor (AdzContext.ztab(True)); pos_dtype: storage of the positions; autobar:
    Schm_autobar_L (level-1 masses with geomh_area_mask_8 given as `mask`)."""

    def __init__(self, tracers, layout, tile, bbox, core, off, geom, pos_dtype=L.ADZ_F32, autobar=False):
        minx, maxx, miny, maxy, nk = layout
        arr = (AdzFluxTracer * len(tracers))()
        tdt = None
        for n, t in enumerate(tracers):
            arr[n].src = _ptr_at(t["src"], off)
            arr[n].fo = _ptr_at(t.get("fo"), off)
            arr[n].fi = _ptr_at(t.get("fi"), off)
            tdt = _DT[t["src"].dtype]
        dom = AdzFluxDomain(minx, maxx, miny, maxy, nk, *tile, *bbox, *core, 1 if autobar else 0)
        h = C.c_void_p()
        L.check(L.lib().adz_flux_create(arr, len(tracers), C.byref(dom), geom, tdt, pos_dtype, C.byref(h)),
                "adz_flux_create")
        self.h, self.off, self.ntr = h, off, len(tracers)
        self.mask_off = (1 - minx) + (maxx - minx + 1) * (1 - miny)
        self.dtype = tracers[0]["src"].dtype
        self.pos_dtype = torch.float32 if pos_dtype == L.ADZ_F32 else torch.float64

    def __call__(self, xyz, airm0, mask, mass):
        """mass: FP64 device tensor [2*ntr] receiving (mass_fo, mass_fi) per tracer."""
        _check_types(self.dtype, airm0=airm0)
        _check_types(self.pos_dtype, xyz=xyz)
        _check_types(torch.float64, mask=mask, mass=mass)
        if mass.numel() < 2 * self.ntr:
            raise ValueError("mass needs 2 values per tracer")
        L.check(L.lib().adz_bc_lam_flux(self.h, xyz.data_ptr(), _ptr_at(airm0, self.off), _ptr_at(mask, self.mask_off),
                                        mass.data_ptr(), _stream()), "adz_bc_lam_flux")

    def active(self):
        n = C.c_int()
        L.check(L.lib().adz_flux_active(self.h, C.byref(n)), "adz_flux_active")
        return n.value

    def free(self):
        if getattr(self, "h", None) is not None and self.h.value:
            L.lib().adz_flux_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


# ---------------------------------------------------------------------- ZLF
L.extra_sigs({"adz_bc_lam_zlf": ("i", ["i", "vp", "vp", "i", "vp", "i", "vp"])})


class AdzZlfDomain(C.Structure):
    _fields_ = [(n, C.c_int) for n in ("minx", "maxx", "miny", "maxy", "nk", "ni", "nj", "pil_w", "pil_e", "pil_s",
                                       "pil_n", "ext", "k0")]


def adz_bc_lam_zlf(setup, f1, f2, layout, tile, pil, ext, k0, off):
    """adz_BC_LAM_zlf_hlt(F1, F2, ..., F_setup) (adz_BC_LAM_zlf_hlt.F90:18-319) on lists of
    device tensors (model layout, element (1,1,1) at offset `off`); f2 may be None for
    setups 1 and 3.  pil = (pil_w, pil_e, pil_s, pil_n); ext = Grd_maxcfl + 1; k0 = Adz_k0t."""
    nf = len(f1)
    a1 = (C.c_void_p * nf)(*[_ptr_at(t, off) for t in f1])
    a2 = (C.c_void_p * nf)(*[_ptr_at(t, off) for t in f2]) if f2 is not None else None
    dom = AdzZlfDomain(*layout, *tile, *pil, ext, k0)
    L.check(L.lib().adz_bc_lam_zlf(setup, a1, a2, nf, C.byref(dom), _DT[f1[0].dtype], _stream()), "adz_bc_lam_zlf")


# ---------------------------------------------------------------------- air mass (set_post_tr_hlt)
L.extra_sigs({
    "adz_set_post_tr": ("i", ["vp", "vp", "d", "vp", "vp", "vp", "i", "vp"]),
})

GRAV = 9.80616  # tdpack grav_8


class AdzAirmassFields(C.Structure):
    _fields_ = [(n, C.c_int) for n in ("minx", "maxx", "miny", "maxy")] + \
               [(n, C.c_void_p) for n in ("pm_p", "pm_m", "hu_p", "hu_m", "hyd_p", "hyd_m")] + \
               [("nhyd", C.c_int)] + [(n, C.c_void_p) for n in ("pt_m", "p0_m", "area")] + \
               [("dry_mixing_ratio", C.c_int), ("wload", C.c_int)]


def adz_set_post_tr(ctx, layout, pm_p, pm_m, pt_m, p0_m, area, airm1, airm0, pkps=None, hu_p=None, hu_m=None,
                    hyd_p=(), hyd_m=(), dry=False, wload=False, grav=GRAV):
    """set_post_tr_hlt (set_post_tr_hlt.F90:44-70): airm1 / airm0 by get_air_mass_hlt at
    times P and M and pkps = pr_t/p0 at M, on device tensors of the model layout
    (minx:maxx, miny:maxy, .) flattened i fastest (pm_* with nk+1 levels, p0_m and the FP64
    area 2-D).  hyd_p / hyd_m: the water-loading tracers in tracer order."""
    minx, maxx, miny, maxy = layout[:4]
    off = (1 - minx) + (maxx - minx + 1) * (1 - miny)
    dt = airm1.dtype
    _check_types(dt, pm_p=pm_p, pm_m=pm_m, pt_m=pt_m, p0_m=p0_m, airm0=airm0, pkps=pkps, hu_p=hu_p, hu_m=hu_m)
    _check_types(torch.float64, area=area)
    for h in list(hyd_p) + list(hyd_m):
        _check_types(dt, hyd=h)
    hp = (C.c_void_p * max(1, len(hyd_p)))(*[_ptr_at(h, off) for h in hyd_p])
    hm = (C.c_void_p * max(1, len(hyd_m)))(*[_ptr_at(h, off) for h in hyd_m])
    f = AdzAirmassFields(minx, maxx, miny, maxy, _ptr_at(pm_p, off), _ptr_at(pm_m, off), _ptr_at(hu_p, off),
                         _ptr_at(hu_m, off), C.cast(hp, C.c_void_p), C.cast(hm, C.c_void_p), len(hyd_p),
                         _ptr_at(pt_m, off), _ptr_at(p0_m, off), _ptr_at(area, off), 1 if dry else 0,
                         1 if wload else 0)
    L.check(L.lib().adz_set_post_tr(ctx.h, C.byref(f), float(grav), _ptr_at(airm1, off), _ptr_at(airm0, off),
                                    _ptr_at(pkps, off), _DT[dt], _stream()), "adz_set_post_tr")
