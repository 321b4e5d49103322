"""ctypes wrapper over liboracle.so -- CPU ORACLE, test infrastructure only.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import this
module, and only as the checker (or the timed CPU baseline).  The product package
(gem_amd) never imports it.

Also loads the reference's own turbo_tricublin.c when `make ref` has compiled it into
oracle/_ref/ (only where the reference tree exists), to pin the restatement.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_AVX2 = os.path.join(HERE, "_ref", "libtricublin_ref_avx2.so")
REF_SCALAR = os.path.join(HERE, "_ref", "libtricublin_ref_scalar.so")
REF_QVZYX = os.path.join(HERE, "_ref", "libqvzyx_ref.so")
REF_ILMC = os.path.join(HERE, "_ref", "libilmc_ref.so")

ORD_AVX2, ORD_SCALAR = 0, 1
_lib = None


def build():
    subprocess.check_call(["make", "-s", "-C", HERE, "liboracle.so"])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        vp, i, l, d = C.c_void_p, C.c_int, C.c_long, C.c_double
        L.ora_ztab_new.restype = vp
        L.ora_ztab_new.argtypes = [vp, i, i, i, i, i]
        L.ora_ztab_free.argtypes = [vp]
        L.ora_ztab_get.argtypes = [vp, vp, vp]
        L.ora_vcoef.argtypes = [vp, vp, i, i, i, vp, vp, vp]
        L.ora_zyx1_m_n.argtypes = [vp, vp, i, vp, i, vp, l, i, i, i]
        L.ora_mono_m_n.argtypes = [vp, vp, vp, vp, vp, i, vp, i, vp, l, i, i, i]
        L.ora_zyx3_n.argtypes = [vp, vp, i, vp, i, vp, l, i, i]
        L.ora_grid_new.restype = vp
        L.ora_grid_new.argtypes = [vp]
        L.ora_grid_free.argtypes = [vp]
        L.ora_grid_info.argtypes = [vp, vp, vp]
        L.ora_grid_table.argtypes = [vp, i, vp]
        L.ora_level_search.argtypes = [vp, i, l, vp, vp]
        L.ora_zindx.argtypes = [vp, l, vp, vp]
        L.ora_traject.argtypes = [vp, i, i, d, i, d, d, i, i] + [vp] * 6 + [vp] * 10 + [i, i]
        L.ora_almost_zero.restype = i
        L.ora_almost_zero.argtypes = [d]
        L.ora_fsum.restype = d
        L.ora_fsum.argtypes = [vp, l]
        L.ora_bermejo_conde.argtypes = [i, vp, i, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp]
        L.ora_mass.restype = d
        L.ora_mass.argtypes = [i, vp, vp, vp, vp]
        L.ora_qv_new.restype = vp
        L.ora_qv_new.argtypes = [vp, i]
        L.ora_qv_free.argtypes = [vp]
        L.ora_qv_m_n.argtypes = [vp, vp, i, vp, i, l, vp, vp, vp, vp, vp]
        L.ora_qv_coef.restype = l
        L.ora_qv_coef.argtypes = [vp, vp, vp, vp, vp]
        L.ora_qv_zyx.argtypes = [vp, vp, vp, l, l, i, i]
        L.ora_ilmc.argtypes = [vp, vp, vp, vp, vp, vp]
        L.ora_ilmc_shell.restype = l
        L.ora_ilmc_shell.argtypes = [vp, i, i, i, i, vp]
        L.ora_set_threads.argtypes = [i]
        _lib = L
    return _lib


def set_threads(n):
    """threads of the oracle's OpenMP loops"""
    lib().ora_set_threads(int(n))


def _p(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def _ptrs(arrs, offsets):
    """Array of pointers to element offsets[k] of arrs[k] (the f(1,1,1) address)."""
    out = (C.c_void_p * len(arrs))()
    for k, (a, o) in enumerate(zip(arrs, offsets)):
        out[k] = a.ctypes.data + int(o) * a.itemsize
    return out


class Ztab:
    """Oracle vertical descriptor (Vsearch_setup_plus restatement)."""

    def __init__(self, levels, ni, nj, offi=0, offj=0):
        self.levels = np.ascontiguousarray(levels, dtype=np.float64)
        self.nk, self.ni, self.nj, self.offi, self.offj = len(levels), ni, nj, offi, offj
        self.h = lib().ora_ztab_new(_p(self.levels), self.nk, ni, nj, offi, offj)

    def tables(self):
        z = np.zeros(self.nk)
        ocz = np.zeros(4 * max(self.nk, 4))
        lib().ora_ztab_get(self.h, _p(z), _p(ocz))
        return z, ocz

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.ora_ztab_free(self.h)
            self.h = None


def _pdt(pos):
    return 1 if pos.dtype == np.float32 else 0


# The reference's production build (gcc -O2 -mavx2 -mfma) contracts the absolute-z
# expression of Vcoef_pxyz4_inline (turbo_tricublin.c:438) into fma(dz, z1, (1-dz)*z0);
# the portable build has no FMA.  zabs_fma=None picks the one matching `order`.
def _zf(order, zabs_fma):
    if zabs_fma is None:
        return 1 if order == ORD_AVX2 else 0
    return zabs_fma


def vcoef(zt, pos, zabs_fma=1):
    pos = np.ascontiguousarray(pos)
    n = pos.size // 3
    ijk = np.zeros(n, np.int32)
    zl = np.zeros(n, np.int32)
    coef = np.zeros((n, 20))
    lib().ora_vcoef(zt.h, _p(pos), _pdt(pos), n, zabs_fma, _p(ijk), _p(zl), _p(coef))
    return ijk, zl, coef


def zyx1_m_n(zt, fields, base_off, pos, order=ORD_AVX2, zabs_fma=None):
    """fields: list of flat arrays (same dtype); base_off: element offset of f(1,1,1)."""
    pos = np.ascontiguousarray(pos)
    n = pos.size // 3
    dt = fields[0].dtype
    d = np.zeros(len(fields) * n, dt)
    fs = _ptrs(fields, [base_off] * len(fields))
    lib().ora_zyx1_m_n(_p(d), fs, 1 if dt == np.float32 else 0, _p(pos), _pdt(pos), zt.h, n, len(fields),
                       order, _zf(order, zabs_fma))
    return d.reshape(len(fields), n)


def mono_m_n(zt, fields, base_off, pos, order=ORD_AVX2, zabs_fma=None):
    pos = np.ascontiguousarray(pos)
    n = pos.size // 3
    dt = fields[0].dtype
    outs = [np.zeros(len(fields) * n, dt) for _ in range(4)]
    fs = _ptrs(fields, [base_off] * len(fields))
    lib().ora_mono_m_n(*[_p(o) for o in outs], fs, 1 if dt == np.float32 else 0, _p(pos), _pdt(pos), zt.h, n,
                       len(fields), order, _zf(order, zabs_fma))
    return [o.reshape(len(fields), n) for o in outs]


def zyx3_n(zt, f3, base_off, pos, order=ORD_AVX2, zabs_fma=None):
    pos = np.ascontiguousarray(pos)
    n = pos.size // 3
    d = np.zeros(3 * n, f3.dtype)
    ptr = f3.ctypes.data + int(base_off) * f3.itemsize
    lib().ora_zyx3_n(_p(d), C.c_void_p(ptr), 1 if f3.dtype == np.float32 else 0, _p(pos), _pdt(pos), zt.h, n,
                     order, _zf(order, zabs_fma))
    return d.reshape(n, 3)


def almost_zero(a):
    return bool(lib().ora_almost_zero(float(a)))


def fsum(v):
    v = np.ascontiguousarray(v, dtype=np.float64)
    return lib().ora_fsum(_p(v), v.size)


# ------------------------------------------------------------------ reference (_ref)
class RefKernel:
    """The reference's turbo_tricublin.c compiled from its own source (oracle/_ref)."""

    def __init__(self, which="avx2"):
        path = REF_AVX2 if which == "avx2" else REF_SCALAR
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        L = C.CDLL(path)
        vp, i = C.c_void_p, C.c_int
        L.Vsearch_setup_plus.restype = vp
        L.Vsearch_setup_plus.argtypes = [vp, i, i, i, i, i]
        for name in ("Tricublin_zyx1_m_n", "Tricublin_zyx1_n_m"):
            getattr(L, name).argtypes = [vp, vp, vp, vp, i, i]
        L.Tricublin_zyx1_n.argtypes = [vp, vp, vp, vp, i]
        L.Tricublin_mono_zyx_m_n.argtypes = [vp, vp, vp, vp, vp, vp, vp, i, i]
        L.Tricublin_zyx3_n.argtypes = [vp, vp, vp, vp, i]
        self.L = L

    def setup(self, levels, ni, nj, offi=0, offj=0):
        self._lev = np.ascontiguousarray(levels, dtype=np.float64)
        return self.L.Vsearch_setup_plus(_p(self._lev), len(levels), ni, nj, offi, offj)

    def zyx1_m_n(self, lv, fields, base_off, pos):
        pos = np.ascontiguousarray(pos, dtype=np.float32)
        n = pos.size // 3
        d = np.zeros(len(fields) * n, np.float32)
        self.L.Tricublin_zyx1_m_n(_p(d), _ptrs(fields, [base_off] * len(fields)), _p(pos), lv, n, len(fields))
        return d.reshape(len(fields), n)

    def mono_m_n(self, lv, fields, base_off, pos):
        pos = np.ascontiguousarray(pos, dtype=np.float32)
        n = pos.size // 3
        outs = [np.zeros(len(fields) * n, np.float32) for _ in range(4)]
        self.L.Tricublin_mono_zyx_m_n(*[_p(o) for o in outs], _ptrs(fields, [base_off] * len(fields)), _p(pos), lv,
                                      n, len(fields))
        return [o.reshape(len(fields), n) for o in outs]

    def zyx3_n(self, lv, f3, base_off, pos):
        pos = np.ascontiguousarray(pos, dtype=np.float32)
        n = pos.size // 3
        d = np.zeros(3 * n, np.float32)
        self.L.Tricublin_zyx3_n(_p(d), C.c_void_p(f3.ctypes.data + int(base_off) * 4), _p(pos), lv, n)
        return d.reshape(n, 3)


# ------------------------------------------------------------------ adz_* layer
class GridDesc(C.Structure):
    """struct GridDesc of oracle_adz.h (same layout as the library's adz_grid)."""
    _fields_ = [(n, C.c_int) for n in ("ni", "nj", "nk", "i0g", "j0g", "halox", "haloy", "west", "east", "south",
                                       "north", "pil_w", "pil_e", "pil_s", "pil_n", "gbpil_t", "bc_lam_flux", "niu",
                                       "njv")] + \
              [("inv_hx", C.c_double), ("inv_hy", C.c_double), ("inv_rayt", C.c_double)] + \
              [(n, C.c_void_p) for n in ("zm", "zt", "zx", "lat")] + \
              [(n, C.c_double) for n in ("zmin", "zmax", "wmstar_nk", "idz_t_nkm1")]


INFO_INTS = ["lminx", "lmaxx", "lminy", "lmaxy", "nit", "njt", "nij", "i0", "in", "j0", "jn", "i0u", "inu", "j0v",
             "jnv", "i0b", "inb", "j0b", "jnb", "k0", "k0t", "k0m", "pnz_m", "pnz_t"]
INFO_DBLS = ["iminposx", "imaxposx", "iminposy", "imaxposy", "sig", "ovdzm", "ovdzt", "vw5", "vw6"]


class Grid:
    """Oracle adz_set restatement for one tile."""

    def __init__(self, g):
        self.g = g
        d = GridDesc()
        self._keep = {}
        for name, ctype in GridDesc._fields_:
            if ctype is C.c_void_p:
                a = np.ascontiguousarray(g[name], dtype=np.float64)
                self._keep[name] = a
                setattr(d, name, a.ctypes.data)
            else:
                setattr(d, name, g[name])
        self._d = d
        self.h = lib().ora_grid_new(C.byref(d))
        iv = np.zeros(24, np.int32)
        dv = np.zeros(9)
        lib().ora_grid_info(self.h, _p(iv), _p(dv))
        self.info = dict(zip(INFO_INTS, iv.tolist()))
        self.info.update(zip(INFO_DBLS, dv.tolist()))

    def table(self, which):
        nk = self.g["nk"]
        if which in (0, 1):
            out = np.zeros(self.info["pnz_m" if which == 0 else "pnz_t"], np.int32)
        elif which == 10:
            out = np.zeros(self.info["njt"])
        else:
            out = np.zeros(nk + 1)
        lib().ora_grid_table(self.h, which, _p(out))
        return out

    def level_search(self, which, pos):
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        kk = np.zeros(pos.size, np.int32)
        lib().ora_level_search(self.h, which, pos.size, _p(pos), _p(kk))
        return kk

    def zindx(self, pos):
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        out = np.zeros(pos.size)
        lib().ora_zindx(self.h, pos.size, _p(pos), _p(out))
        return out

    def sizes(self):
        i, g = self.info, self.g
        ni, nj, nk = g["ni"], g["nj"], g["nk"]
        nxy = ni * nj
        return dict(p3=3 * nxy * (nk - i["k0m"] + 1), wp=(ni + 4) * (nj + 4) * nk * 3,
                    pm=3 * (i["in"] - i["i0"] + 1) * (i["jn"] - i["j0"] + 1) * (nk - i["k0"] + 1),
                    pmu=3 * (i["inu"] - i["i0u"] + 1) * (i["jn"] - i["j0"] + 1) * (nk - i["k0"] + 1),
                    pmv=3 * (i["in"] - i["i0"] + 1) * (i["jnv"] - i["j0v"] + 1) * (nk - i["k0"] + 1),
                    pt=3 * (i["in"] - i["i0"] + 1) * (i["jn"] - i["j0"] + 1) * (nk - i["k0t"] + 1),
                    pb=3 * (i["inb"] - i["i0b"] + 1) * (i["jnb"] - i["j0b"] + 1) * nk,
                    uvw_d=3 * i["nij"] * nk, arr=3 * nxy * nk)

    def traject(self, winds, dt, storage="f32", state=None, niter=3, advec=1, bA_m=0.6, bA=0.6, order=ORD_AVX2,
                band=None, want=None):
        """adz_traject restatement.  storage 'f32' = reference `real` storage, 'f64' = FP64.
        state: dict with 'pxyzm' and 'wpxyz' carried from a previous step (None: inittraj,
        and then max(6, niter) iterations as adz_inittraj sets).  band = (klo, khi): only
        the position sets of those levels are produced (exactly: the departure points of
        levels klo-1..khi+2 are iterated); want: names of the outputs to return (default
        all), to spare host memory at large sizes."""
        dt_ = np.float32 if storage == "f32" else np.float64
        sz = self.sizes()
        names = (("uvw_dep", "p3"), ("pm", "pm"), ("pmu", "pmu"), ("pmv", "pmv"), ("pt", "pt"), ("pb", "pb"),
                 ("uvw_d", "uvw_d"), ("warr", "arr"))
        out = {k: (np.zeros(sz[s], dt_) if want is None or k in want else None) for k, s in names}
        init = state is None
        pxyzm = np.zeros(sz["p3"], dt_) if init else np.ascontiguousarray(state["pxyzm"], dtype=dt_).copy()
        wpxyz = np.zeros(sz["wp"]) if init else np.ascontiguousarray(state["wpxyz"], dtype=np.float64).copy()
        nit = max(6, niter) if init else niter
        w = {k: np.ascontiguousarray(np.asarray(winds[k]), dtype=np.float32)
             for k in ("uu_ext", "vv_ext", "ww_ext", "ut0", "vt0", "zdt0")}
        klo, khi = band if band is not None else (1, self.g["nk"])
        lib().ora_traject(self.h, 1 if storage == "f32" else 0, 1 if init else 0, dt, advec, bA_m, bA, nit, order,
                          *[_p(w[k]) for k in ("uu_ext", "vv_ext", "ww_ext", "ut0", "vt0", "zdt0")],
                          _p(pxyzm), _p(wpxyz), *[None if out[k] is None else _p(out[k])
                                                  for k in ("uvw_dep", "pm", "pmu", "pmv", "pt", "pb", "uvw_d",
                                                            "warr")], int(klo), int(khi))
        out["pxyzm"], out["wpxyz"] = pxyzm, wpxyz
        return {k: v for k, v in out.items() if v is not None}

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.ora_grid_free(self.h)
            self.h = None


def bermejo_conde(tracers, layout, scope, airm1, airm0, mask, deficit, pkps=None, min_max=True,
                  pexp_list=(1.0,) * 9, sum_mode=1, flux=None, autobar=False):
    """Bermejo_Conde restatement on host arrays (array-start layout).  tracers: list of
    dicts with p, m (modified in place), lin, min, max and `mass` code.  sum_mode 1 =
    exactly rounded sums, 0 = the reference's sequential order.  flux: (mass_fo, mass_fi)
    per tracer of the LAM Aranami flux (None: zero fluxes).  autobar: Schm_autobar_L (the
    masses are level-1 sums and `mask` is geomh_area_mask_8).  Returns (deficit, diag)."""
    minx, maxx, miny, maxy, nk = layout
    i0, in_, j0, jn, k0 = scope
    ntr = len(tracers)
    dt = tracers[0]["m"].dtype
    tdt = 1 if dt == np.float32 else 0

    def ptrs(key):
        arr = (C.c_void_p * ntr)()
        for n, t in enumerate(tracers):
            arr[n] = t[key].ctypes.data
        return arr

    def codes(mass):
        if mass == 1:
            return 1, 1, 0
        w = (mass - 100) // 10
        return 0, w, mass - (100 + w * 10)

    cs = [codes(t["mass"]) for t in tracers]
    legacy = np.array([c[0] for c in cs], np.int32)
    weight = np.array([c[1] for c in cs], np.int32)
    pexp_n = np.array([c[2] for c in cs], np.int32)
    dom = np.array([minx, maxx, miny, maxy, nk, i0, in_, j0, jn, k0, 1 if min_max else 0, sum_mode,
                    1 if autobar else 0], np.int32)
    de = np.ascontiguousarray(deficit, dtype=np.float64).copy()
    diag = np.zeros(6 * ntr)
    pk = pkps if pkps is not None else np.zeros_like(airm0)
    pl = np.asarray(pexp_list, np.float32)
    lib().ora_bermejo_conde(tdt, _p(dom), ntr, ptrs("p"), ptrs("m"), ptrs("lin"), ptrs("min"), ptrs("max"),
                            _p(legacy), _p(weight), _p(pexp_n), _p(de), _p(airm1), _p(airm0), _p(pk), _p(mask),
                            _p(pl), _p(diag), None if flux is None else _p(np.ascontiguousarray(flux, np.float64)))
    return de, diag.reshape(ntr, 6)


def mass(x, air, mask, layout, scope, sum_mode=1, autobar=False):
    """mass_tr: sum of x * air * mask over the scope (autobar: level 1 of x * mask, with
    mask = geomh_area_mask_8)"""
    minx, maxx, miny, maxy, nk = layout
    i0, in_, j0, jn, k0 = scope
    dom = np.array([minx, maxx, miny, maxy, nk, i0, in_, j0, jn, k0, 0, sum_mode, 1 if autobar else 0], np.int32)
    return lib().ora_mass(1 if x.dtype == np.float32 else 0, _p(dom), _p(x), _p(air), _p(mask))


# ------------------------------------------------------------------ LAM boundary flux
def bc_lam_masks(layout, tile, core):
    """adz_BC_LAM_setup.F90:18-92 for a tile whose four sides are LAM edges.

    mask_o is 1 on the core (:50-62) and mask_i is 1 on the tile 1..ni x 1..nj x 1..nk
    outside the core (:64-76).  Both are 0 in the halo (:80; one tile: nothing is exchanged
    into it).  Returns two float32 arrays (nk, ny, nx) on `layout` = (minx, maxx, miny,
    maxy, nk); core = (ic0, ic1, jc0, jc1, kc0)."""
    minx, maxx, miny, maxy, nk = layout
    ni, nj = tile
    ic0, ic1, jc0, jc1, kc0 = core
    shape = (nk, maxy - miny + 1, maxx - minx + 1)
    one = np.zeros(shape, np.float32)
    one[:, 1 - miny:nj + 1 - miny, 1 - minx:ni + 1 - minx] = 1.0           # fld_ONE(1:ni, 1:nj, 1:nk)
    c = (slice(kc0 - 1, nk), slice(jc0 - miny, jc1 + 1 - miny), slice(ic0 - minx, ic1 + 1 - minx))
    mo = np.zeros(shape, np.float32)
    mo[c] = one[c]
    mi = one.copy()
    mi[c] = 0.0
    return mo, mi


def bc_lam_aranami(tracers, layout, tile, bbox, core, zt, pb, order=ORD_AVX2, zabs_fma=None, kernel=None):
    """adz_BC_LAM_Aranami_hlt.F90:18-236 for a tile whose four sides are LAM edges.

    tracers: flat arrays on `layout` (array start).  zt: Ztab over the layout (ni = row
    pitch, no offsets); pb: Adz_pb (3, i0b:inb, j0b:jnb, 1:nk); bbox = (i0b, inb, j0b, jnb).
    kernel: None (this oracle's Tricublin_zyx1_m_n) or (RefKernel, its ztab handle) to
    interpolate with the reference's own compiled code.  Returns (fo, fi): lists of flat
    arrays on `layout` holding the b-range values and zero elsewhere."""
    minx, maxx, miny, maxy, nk = layout
    ni, nj = tile
    i0b, inb, j0b, jnb = bbox
    ic0, ic1, jc0, jc1, kc0 = core
    pil_w, pil_e, pil_s, pil_n = ic0 - 1, ni - ic1, jc0 - 1, nj - jc1
    nx, ny = maxx - minx + 1, maxy - miny + 1
    base = (1 - minx) + nx * (1 - miny)
    mo, mi = bc_lam_masks(layout, tile, core)
    tile_sl = (slice(None), slice(1 - miny, nj + 1 - miny), slice(1 - minx, ni + 1 - minx))  # [i0:in] x [j0:jn]
    fields = []
    for f in tracers:  # :58-82: adz_o / adz_i = mask * tracer, a product in the tracer's type
        f3 = f.reshape(nk, ny, nx)
        ao, ai = np.zeros_like(f3), np.zeros_like(f3)
        ao[tile_sl] = mo[tile_sl].astype(f.dtype) * f3[tile_sl]
        ai[tile_sl] = mi[tile_sl].astype(f.dtype) * f3[tile_sl]
        fields += [ao.reshape(-1), ai.reshape(-1)]
    # :110-155: one m_n interpolation of (adz_o, adz_i) per tracer at every Adz_pb target
    if kernel is None:
        d = zyx1_m_n(zt, fields, base, pb, order, zabs_fma)
    else:
        ref, lv = kernel
        d = ref.zyx1_m_n(lv, fields, base, pb)
    nbx, nby = inb - i0b + 1, jnb - j0b + 1
    b_sl = (slice(None), slice(j0b - miny, jnb + 1 - miny), slice(i0b - minx, inb + 1 - minx))
    J = lambda a, b: slice(a - miny, b + 1 - miny)  # noqa: E731  (j = a..b)
    I = lambda a, b: slice(a - minx, b + 1 - minx)  # noqa: E731
    fo, fi = [], []
    for n, f in enumerate(tracers):
        o = np.zeros((nk, ny, nx), f.dtype)
        i_ = np.zeros((nk, ny, nx), f.dtype)
        o[b_sl] = d[2 * n].reshape(nk, nby, nbx)
        i_[b_sl] = d[2 * n + 1].reshape(nk, nby, nbx)
        o[kc0 - 1:, J(1 + pil_s, nj - pil_n), I(1 + pil_w, ni - pil_e)] = 0     # :161-169
        i_[kc0 - 1:, J(nj - pil_n + 1, nj), I(1, ni)] = 0                        # :171-181 (l_north)
        i_[kc0 - 1:, J(1, nj), I(ni - pil_e + 1, ni)] = 0                        # :183-193 (l_east)
        i_[kc0 - 1:, J(1, pil_s), I(1, ni)] = 0                                  # :195-205 (l_south)
        i_[kc0 - 1:, J(1, nj), I(1, pil_w)] = 0                                  # :207-217 (l_west)
        i_[:kc0 - 1, J(j0b, jnb), I(i0b, inb)] = 0                               # :219-227
        fo.append(o.reshape(-1))
        fi.append(i_.reshape(-1))
    return fo, fi


def flux_masses(fo, fi, airm0, mask, layout, bbox, sum_mode=1, autobar=False):
    """mass_tr_PM.F90:101-127: masses of FLUX_out and FLUX_in over Adz_i0b:Adz_inb x
    Adz_j0b:Adz_jnb x 1:nk with the air mass at time M (autobar: level 1 with mask =
    geomh_area_mask_8, :105-112).  Returns [mass_fo, mass_fi] per tracer, flattened."""
    i0b, inb, j0b, jnb = bbox
    scope = (i0b, inb, j0b, jnb, 1)
    out = []
    for o, i_ in zip(fo, fi):
        out += [mass(o, airm0, mask, layout, scope, sum_mode, autobar),
                mass(i_, airm0, mask, layout, scope, sum_mode, autobar)]
    return np.array(out)


# ------------------------------------------------------------------ bicubic H / quintic V
class QvTab:
    """Tables of the BICHQV interpolation (adz_set.F90:181-190,286-307); zt = Ver_z_8%t(0:nk+1)."""

    def __init__(self, zt, nk):
        self.zt = np.ascontiguousarray(zt, dtype=np.float64)
        self.nk = nk
        self.h = lib().ora_qv_new(_p(self.zt), nk)

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.ora_qv_free(self.h)
            self.h = None


def qv_m_n(tab, fields, layout, pos, ioff, joff):
    """bicubHQV_lin_min_max_ntr (bicubHQV_lin_mono.F90:20-172).  fields: float32 flat arrays
    starting at (minx, miny, 1) of layout = (minx, maxx, miny, maxy); ioff, joff = Adz_ioff,
    Adz_joff (l_i0 - 2, l_j0 - 2).  Returns (qut, lin, min, max), each (nfields, n)."""
    pos = np.ascontiguousarray(pos)
    n = pos.size // 3
    outs = [np.zeros(len(fields) * n, np.float32) for _ in range(4)]
    fs = _ptrs(fields, [0] * len(fields))
    dom = np.array(list(layout) + [ioff, joff], np.int32)
    lib().ora_qv_m_n(tab.h, fs, len(fields), _p(pos), _pdt(pos), n, _p(dom), *[_p(o) for o in outs])
    return [o.reshape(len(fields), n) for o in outs]


def qv_coef(tab, pos3, layout, ioff, joff):
    """Weights of one target: (w[20] = cx cy qz lx ly lz, quintic, cubic, element offset)."""
    p = np.ascontiguousarray(pos3, dtype=np.float32)
    w = np.zeros(20)
    fl = np.zeros(2, np.int32)
    dom = np.array(list(layout) + [ioff, joff], np.int32)
    off = lib().ora_qv_coef(tab.h, _p(p), _p(dom), _p(w), _p(fl))
    return w, bool(fl[0]), bool(fl[1]), off


def qv_zyx(src, off, w, ni, ninj, zq, zc):
    """bicubHQV_lin_min_max_zyx restated: (qut, lin, min, max) of the stencil at src[off]."""
    out = np.zeros(4, np.float32)
    lib().ora_qv_zyx(_p(out), C.c_void_p(src.ctypes.data + 4 * int(off)), _p(np.ascontiguousarray(w)), ni, ninj,
                     int(zq), int(zc))
    return out


class RefQvZyx:
    """The reference's bicubHQV_lin_min_max_zyx compiled from its own source (oracle/_ref)."""

    def __init__(self):
        if not os.path.exists(REF_QVZYX):
            raise FileNotFoundError(REF_QVZYX)
        self.L = C.CDLL(REF_QVZYX)
        self.f = self.L.bicubhqv_lin_min_max_zyx_
        self.f.argtypes = [C.c_void_p] * 15

    def __call__(self, src, off, w, ni, ninj, zq, zc):
        out = np.zeros(4, np.float32)
        w = np.ascontiguousarray(w, dtype=np.float64)
        ints = np.array([ni, ninj, 1 if zq else 0, 1 if zc else 0], np.int32)
        a = out.ctypes.data
        wd = w.ctypes.data
        self.f(C.c_void_p(a), C.c_void_p(a + 4), C.c_void_p(a + 8), C.c_void_p(a + 12),
               C.c_void_p(src.ctypes.data + 4 * int(off)), C.c_void_p(wd), C.c_void_p(wd + 32),
               C.c_void_p(wd + 64), C.c_void_p(wd + 112), C.c_void_p(wd + 128), C.c_void_p(wd + 144),
               C.c_void_p(ints.ctypes.data), C.c_void_p(ints.ctypes.data + 4), C.c_void_p(ints.ctypes.data + 8),
               C.c_void_p(ints.ctypes.data + 12))
        return out


def bc_lam_zlf(setup, f1, f2, layout, tile, pil, ext, k0):
    """adz_BC_LAM_zlf_hlt.F90:18-319 for a tile whose four sides are LAM edges, in place on
    flat arrays of `layout` = (minx, maxx, miny, maxy, nk) (array start); pil = (pil_w,
    pil_e, pil_s, pil_n); ext = Grd_maxcfl + 1; k0 = Adz_k0t."""
    minx, maxx, miny, maxy, nk = layout
    ni, nj = tile
    pil_w, pil_e, pil_s, pil_n = pil
    shape = (nk, maxy - miny + 1, maxx - minx + 1)
    F1 = f1.reshape(shape)
    F2 = f2.reshape(shape) if f2 is not None else None

    def box(ia, ib, ja, jb, ka, kb):  # i = ia..ib, j = ja..jb, k = ka..kb (1-based, inclusive; empty if b < a)
        if ib < ia or jb < ja or kb < ka:
            return (slice(0, 0), slice(0, 0), slice(0, 0))
        return (slice(ka - 1, kb), slice(ja - miny, jb + 1 - miny), slice(ia - minx, ib + 1 - minx))

    if setup == 1:                                            # :44-95
        i0e, ine = 1 + pil_w - ext, ni - pil_e + ext
        j0e, jne = 1 + pil_s - ext, nj - pil_n + ext
        k0d2 = int(np.floor(k0 / 2.0 + 0.5))                  # nint(k0/2.) for k0 >= 0
        F1[box(1, i0e - 1, 1, nj, 1, nk)] = 0
        F1[box(ine + 1, ni, 1, nj, 1, nk)] = 0
        F1[box(1, ni, 1, j0e - 1, 1, nk)] = 0
        F1[box(1, ni, jne + 1, nj, 1, nk)] = 0
        F1[box(i0e, ine, j0e, jne, 1, k0d2 - 1)] = 0
    elif setup in (2, 4):                                     # :97-151, :212-268
        src, dst = (F1, F2) if setup == 2 else (F2, F1)
        i0c, inc, j0c, jnc = 1 + pil_w, ni - pil_e, 1 + pil_s, nj - pil_n
        for b in (box(1, i0c - 1, 1, nj, 1, nk), box(inc + 1, ni, 1, nj, 1, nk), box(1, ni, 1, j0c - 1, 1, nk),
                  box(1, ni, jnc + 1, nj, 1, nk), box(i0c, inc, j0c, jnc, 1, k0 - 1)):
            dst[b] = src[b]
    elif setup == 3:                                          # :153-210
        B = 4
        F1[box(1, B, 1, nj, 1, nk)] = 0
        F1[box(ni - B + 1, ni, 1, nj, 1, nk)] = 0
        F1[box(1, ni, 1, B, 1, nk)] = 0
        F1[box(1, ni, nj - B + 1, nj, 1, nk)] = 0
    else:
        raise ValueError("ADZ_BC_LAM_ZLF: SETUP not defined")


def _ilmc_dom(layout, box, sweep_max, min_max, autobar):
    minx, maxx, miny, maxy, nk = layout
    i0, in_, j0, jn, k0 = box
    return np.array([minx, maxx, miny, maxy, nk, i0, in_, j0, jn, k0, sweep_max, int(min_max), int(autobar)], np.int32)


def ilmc(f, mn, mx, air, layout, box, sweep_max=2, min_max=True, autobar=False):
    """ILMC_LAM (ILMC_LAM.F90:18-273, ilmc_lam_loop.inc:16-139) in place on the FP32 flat
    array f of `layout` = (minx, maxx, miny, maxy, nk) over `box` = (i0, in, j0, jn, k0).
    Returns the counts (sweeps that moved mass, min resets, max resets)."""
    for a in (f, mn, mx, air):
        assert a.dtype == np.float32 and a.flags.c_contiguous
    counts = np.zeros(3, np.int64)
    lib().ora_ilmc(_p(f), _p(mn), _p(mx), _p(air), _p(_ilmc_dom(layout, box, sweep_max, min_max, autobar)),
                   _p(counts))
    return counts


def ilmc_shell(layout, box, i, j, k, s, autobar=False):
    """Array offsets of the cells of sweep s around (i,j,k), in the oracle's visiting order."""
    out = np.zeros((2 * s + 1) ** 3, np.int64)
    n = lib().ora_ilmc_shell(_p(_ilmc_dom(layout, box, s, 0, autobar)), i, j, k, s, _p(out))
    return out[:n].copy()


class RefIlmc:
    """The reference's ILMC sweep loop (ilmc_lam_loop.inc) compiled from its own source
    inside oracle/ilmc_ref_host.F90 (oracle/_ref).  No min/max reset."""

    def __init__(self):
        if not os.path.exists(REF_ILMC):
            raise FileNotFoundError(REF_ILMC)
        self.L = C.CDLL(REF_ILMC)
        self.f = self.L.ilmc_ref_loop
        self.f.argtypes = [C.c_void_p] * 6

    def __call__(self, f, mn, mx, air, layout, box, sweep_max=2, autobar=False):
        moved = np.zeros(1, np.int64)
        self.f(_p(f), _p(mn), _p(mx), _p(air), _p(_ilmc_dom(layout, box, sweep_max, 0, autobar)), _p(moved))
        return int(moved[0])


# ------------------------------------------------------------------ air mass (fixer inputs)
def air_mass(pm, hu, hyd, area, zm, layout, tile, k0, dry, wload, grav):
    """get_air_mass_hlt (get_air_mass_hlt.F90:18-103) for one time level, on host arrays of
    the model layout (minx:maxx, miny:maxy, .) as (k, j, i) numpy arrays: pm has nk+1 levels
    (pw_pm_*), hu the humidity, hyd the list of hydrometeor tracers in tracer order, area
    the FP64 geomh_area_8 (j, i); zm = Ver_z_8%m(0:nk+1).  Arithmetic as the reference writes
    it: sumq in FP64 from 0, each hydrometeor then the humidity (sumhydro_hlt.F90:35-60,
    :67-79, only with Schm_dry_mixing_ratio_L, hydrometeors also needing Schm_wload_L); the
    pressure difference in the storage type; dens = diff*(1-sumq)*idz*(1/g) and
    air = dens*area*dz left to right in FP64 (:83-94), Ver_dz_8%t(k) = zm(k+1) - zm(k) and
    Ver_idz_8%t = 1/dz (set_zeta.F90:271-275); zero outside the tile and below k0 (:84-91)."""
    minx, maxx, miny, maxy, nk = layout
    ni, nj = tile
    dt = pm.dtype
    out = np.zeros((nk, maxy - miny + 1, maxx - minx + 1), dt)
    zm = np.asarray(zm, np.float64)
    inv_grav = 1.0 / grav
    sj = slice(1 - miny, nj - miny + 1)
    si = slice(1 - minx, ni - minx + 1)
    for k in range(k0, nk + 1):
        sumq = np.zeros((nj, ni))
        if dry and wload:
            for h in hyd:
                sumq = sumq + h[k - 1, sj, si].astype(np.float64)
        if dry:
            sumq = sumq + hu[k - 1, sj, si].astype(np.float64)
        diff = (pm[k, sj, si] - pm[k - 1, sj, si]).astype(np.float64)   # in the storage type
        dz = zm[k + 1] - zm[k]
        idz = 1.0 / dz
        dens = diff * (1.0 - sumq) * idz * inv_grav
        out[k - 1, sj, si] = (dens * area[sj, si] * dz).astype(dt)
    return out


def pkps(pt, p0, layout, tile):
    """set_post_tr_hlt.F90:62-70: pr_t / p0 at time M on 1..l_ni x 1..l_nj x 1..nk, in the
    storage type (IEEE division); elsewhere the array keeps what it held (zeros here)."""
    minx, maxx, miny, maxy, nk = layout
    ni, nj = tile
    out = np.zeros_like(pt)
    sj = slice(1 - miny, nj - miny + 1)
    si = slice(1 - minx, ni - minx + 1)
    out[:, sj, si] = pt[:, sj, si] / p0[None, sj, si]
    return out
