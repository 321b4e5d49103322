"""GPU parity at the workloads' full sizes, against the oracle.

c2  1024x512x84, halo 5, dt 450 s, seeded rotation + Gaussian jets at Courant numbers up to
    6 in x, y and 2 in z: the plain FP64 drop-in.  The first step's trajectory (max(6, 3)
    iterations, positions stored as the reference's `real` arrays) is bit-exact on every
    position set; the cell indices and weights of the momentum and thermodynamic sets are
    bit-exact; u on Adz_pmu, v on Adz_pmv and rhsc on Adz_pm (momentum levels), T on Adz_pt
    (extended sources) and q + 10 tracers on Adz_pt (model layout), 15 FP64 fields in the
    reference's arithmetic, are bit-exact (the groups of adz_main.F90:32-58 and
    adz_tracers.F90:38-66).
c5  2048x1024x84: the trajectory over the whole tile on the GPU against the oracle over a
    band of levels (exact: each departure point reads only the winds and its own iterate)
    with the bottom and top levels included; the FP32 form of the tracer groups (the
    persistent launch with lin/min/max) against the FP64 oracle at the same positions:
    relative RMS <= 1e-5 per field and bit-exact indices on the band; 50 steps of trajectory
    + interpolation + Bermejo-Conde on 4 tracers: the conserved budget (mass minus carried
    deficit, summed exactly on the host) constant to 1e-13 relative.
c4  16 subdomains from 64x48x84 to 512x384x84, 40 tracers each, the committed autotuned
    tiles: the persistent launch is bit-identical to one adz_tricub_hlt call per
    subdomain, within 1e-5 relative RMS of the FP64 oracle on sampled subdomains, and a
    hipGraph of 20 timesteps (16 trajectories + the launch per step) replays bit for bit
    as the same steps run eagerly.

Inputs are seeded (gem_amd/synth.py).  Each case prints what it runs before it runs it.
"""
import ctypes as C
import json
import math
import os
import sys

import numpy as np
import pytest

from gem_amd import synth as S
from oracle import oracle as O

pytestmark = pytest.mark.gpu
MH = 2
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def say(*a):
    print("[configs]", *a, file=sys.stderr, flush=True)


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint8)


class _Geom:
    """a context's vertical descriptor as the handle gem_amd.tricublin.vcoef takes"""

    def __init__(self, ptr):
        self.h = C.c_void_p(ptr)


@pytest.fixture(scope="module")
def T(cuda):
    import torch

    from gem_amd import _lib as L
    from gem_amd import adz as A
    from gem_amd import tricublin as TC

    return torch, L, A, TC


# ----------------------------------------------------------------------------- c2
@pytest.fixture(scope="module")
def C2(T):
    torch, L, A, TC = T
    cfg = S.CONFIGS["c2"]
    g = S.grid_desc(cfg)
    say(f"c2: {cfg.ni}x{cfg.nj}x{cfg.nk} halo {cfg.halo}: winds, GPU trajectory")
    w = S.winds(cfg, g)
    ctx = A.AdzContext(g, L.ADZ_F32, itraj=cfg.niter)
    wd = {k: v.cuda() for k, v in w.items()}
    ctx.adz_inittraj()
    ctx.adz_traject(cfg.dt, wd["uu_ext"], wd["vv_ext"], wd["ww_ext"], wd["ut0"], wd["vt0"], wd["zdt0"])
    torch.cuda.synchronize()
    del wd
    say("c2: oracle trajectory (whole tile)")
    G = O.Grid(g)
    ref = G.traject({k: v.numpy() for k, v in w.items()}, cfg.dt, storage="f32", niter=cfg.niter)
    yield cfg, g, ctx, ref
    ctx.free()


def test_c2_trajectory_bit_exact(T, C2):
    torch, L, A, TC = T
    cfg, g, ctx, ref = C2
    for name, which in (("pxyzm", A.ADZ_A_PXYZM), ("pm", A.ADZ_A_PM), ("pmu", A.ADZ_A_PMU), ("pmv", A.ADZ_A_PMV),
                        ("pt", A.ADZ_A_PT), ("pb", A.ADZ_A_PB), ("uvw_dep", A.ADZ_A_UVW_DEP)):
        got = ctx.array(which).cpu().numpy()
        assert got.shape == ref[name].shape, name
        assert np.array_equal(_bits(got), _bits(ref[name])), (name, int(np.count_nonzero(got != ref[name])))
    got = ctx.array(A.ADZ_A_WPXYZ).cpu().numpy().reshape(3, cfg.nk, cfg.nj + 4, cfg.ni + 4)[:, :, 2:-2, 2:-2]
    exp = ref["wpxyz"].reshape(3, cfg.nk, cfg.nj + 4, cfg.ni + 4)[:, :, 2:-2, 2:-2]
    assert np.array_equal(_bits(got), _bits(exp))
    # the workload reaches its large Courant numbers: departure points several cells away
    i = ctx.info
    pm = ref["pm"].reshape(cfg.nk - i["k0"] + 1, i["jn"] - i["j0"] + 1, i["in"] - i["i0"] + 1, 3).astype(np.float64)
    ii = np.arange(i["i0"], i["in"] + 1)[None, None, :]
    jj = np.arange(i["j0"], i["jn"] + 1)[None, :, None]
    kk = np.arange(i["k0"], cfg.nk + 1)[:, None, None]
    dx, dy = np.abs(pm[..., 0] - ii).max(), np.abs(pm[..., 1] - jj).max()
    say(f"c2: largest displacements {dx:.2f} cells in x, {dy:.2f} in y")
    assert max(dx, dy) > 4.0
    assert np.abs(pm[..., 2] - kk).max() > 0.5


def test_c2_indices_bit_exact(T, C2):
    """Vcoef_pxyz4_inline on the thermodynamic (Adz_pt) and momentum (Adz_pmu) sets."""
    torch, L, A, TC = T
    cfg, g, ctx, ref = C2
    i = ctx.info
    for name, which, thermo, lev in (("pt", A.ADZ_A_PT, True, g["zt"]), ("pmu", A.ADZ_A_PMU, False, g["zm"])):
        pos = ctx.array(which)
        zt = O.Ztab(lev[1:cfg.nk + 1], i["nit"], i["njt"], 0, 0)
        n = pos.numel() // 3
        chunk = 1 << 22
        for p0 in range(0, n, chunk):
            m = min(chunk, n - p0)
            ijk, zl, coef = TC.vcoef(pos[3 * p0:3 * (p0 + m)], _Geom(ctx.ztab(thermo)), m)
            rijk, rzl, rcoef = O.vcoef(zt, pos[3 * p0:3 * (p0 + m)].cpu().numpy(), zabs_fma=1)
            assert np.array_equal(ijk.cpu().numpy(), rijk), (name, p0)
            assert np.array_equal(zl.cpu().numpy(), rzl), (name, p0)
            assert np.array_equal(_bits(coef.cpu().numpy()), _bits(rcoef)), (name, p0)


def test_c2_fp64_drop_in_bit_exact(T, C2):
    torch, L, A, TC = T
    cfg, g, ctx, ref = C2
    i = ctx.info
    nit, njt = i["nit"], i["njt"]
    h = cfg.halo
    ext_off = (1 - i["lminx"]) + nit * (1 - i["lminy"])
    mod_shape = (cfg.nk, cfg.nj + 2 * MH, cfg.ni + 2 * MH)
    mod_off = MH + (cfg.ni + 2 * MH) * MH
    dl = (1 - MH, cfg.ni + MH, 1 - MH, cfg.nj + MH)
    n_mod = math.prod(mod_shape)
    say("c2: 15 FP64 fields (u, v, rhsc, T extended; q + 10 tracers model layout)")
    ext = S.fields(cfg, 4, seed=cfg.seed + 1)                   # u, v, T as the smooth fields 0..2, rhsc 3
    mod = [np.ascontiguousarray(f.numpy().reshape(cfg.nk, cfg.nj + 2 * h, cfg.ni + 2 * h)
                                [:, h - MH:h + cfg.nj + MH, h - MH:h + cfg.ni + MH]).reshape(-1)
           for f in S.fields(cfg, 11, kind="front", seed=cfg.seed + 2)]
    ext = [f.numpy() for f in ext]
    groups = [  # (name, sources, position set, thermo levels, box, ext layout)
        ("u", [ext[0]], A.ADZ_A_PMU, False, (i["i0u"], i["inu"], i["j0"], i["jn"], i["k0"]), True),
        ("v", [ext[1]], A.ADZ_A_PMV, False, (i["i0"], i["in"], i["j0v"], i["jnv"], i["k0"]), True),
        ("rhsc", [ext[3]], A.ADZ_A_PM, False, (i["i0"], i["in"], i["j0"], i["jn"], i["k0"]), True),
        ("T", [ext[2]], A.ADZ_A_PT, True, (i["i0"], i["in"], i["j0"], i["jn"], i["k0t"]), True),
        ("q+tracers", mod, A.ADZ_A_PT, True, (i["i0"], i["in"], i["j0"], i["jn"], i["k0t"]), False),
    ]
    fill = -777.0
    for name, srcs, which, thermo, box, is_ext in groups:
        say(f"c2: group {name}: {len(srcs)} field(s)")
        xyz = ctx.array(which)
        stack = [dict(src=torch.from_numpy(s).cuda(), dst=torch.full((n_mod,), fill, dtype=torch.float64,
                                                                      device="cuda")) for s in srcs]
        A.adz_tricub_hlt(stack, xyz, ctx.ztab(thermo), *box, cfg.nk, dl, ext=is_ext, math=L.ADZ_MATH_REF,
                         src_off=ext_off if is_ext else mod_off)
        torch.cuda.synchronize()
        lev = (g["zt"] if thermo else g["zm"])[1:cfg.nk + 1]
        zt = O.Ztab(lev, nit, njt, 0, 0) if is_ext else O.Ztab(lev, cfg.ni + 2 * MH, cfg.nj + 2 * MH, 0, 0)
        exp = O.zyx1_m_n(zt, srcs, ext_off if is_ext else mod_off, xyz.cpu().numpy(), O.ORD_AVX2)
        i0, in_, j0, jn, k0 = box
        for f, e in enumerate(stack):
            got = e["dst"].cpu().numpy().reshape(mod_shape)
            inner = got[k0 - 1:, j0 - 1 + MH:jn + MH, i0 - 1 + MH:in_ + MH]
            want = exp[f].reshape(inner.shape)
            assert np.array_equal(_bits(inner), _bits(want)), (name, f, int(np.count_nonzero(inner != want)))
            # nothing outside the target box is written
            outside = got.copy()
            outside[k0 - 1:, j0 - 1 + MH:jn + MH, i0 - 1 + MH:in_ + MH] = fill
            assert np.all(outside == fill), (name, f)
        del stack


# ----------------------------------------------------------------------------- c5
BAND = [(1, 3), (82, 84)]  # bottom and top levels of the 84 (their special cases)


@pytest.fixture(scope="module")
def C5(T):
    torch, L, A, TC = T
    cfg = S.CONFIGS["c5"]
    g = S.grid_desc(cfg)
    say(f"c5: {cfg.ni}x{cfg.nj}x{cfg.nk}: winds, GPU trajectory")
    w = S.winds(cfg, g, device="cuda")
    ctx = A.AdzContext(g, L.ADZ_F32, itraj=cfg.niter)
    ctx.adz_inittraj()
    ctx.adz_traject(cfg.dt, w["uu_ext"], w["vv_ext"], w["ww_ext"], w["ut0"], w["vt0"], w["zdt0"])
    torch.cuda.synchronize()
    yield cfg, g, ctx, w
    ctx.free()


def _band_rows(ctx, name, k0key, klo, khi):
    """slice of a position set ((3, i, j) triplets, levels outermost) holding levels klo..khi"""
    i = ctx.info
    if name == "pmu":
        nx, ny = i["inu"] - i["i0u"] + 1, i["jn"] - i["j0"] + 1
    elif name == "pmv":
        nx, ny = i["in"] - i["i0"] + 1, i["jnv"] - i["j0v"] + 1
    elif name == "pb":
        nx, ny = i["inb"] - i["i0b"] + 1, i["jnb"] - i["j0b"] + 1
    else:
        nx, ny = i["in"] - i["i0"] + 1, i["jn"] - i["j0"] + 1
    k0 = 1 if name == "pb" else i[k0key]
    lo = max(klo, k0)
    per = 3 * nx * ny
    return slice((lo - k0) * per, max(lo - k0, khi - k0 + 1) * per)


def test_c5_trajectory_band_bit_exact(T, C5):
    torch, L, A, TC = T
    cfg, g, ctx, w = C5
    G = O.Grid(g)
    wn = {k: v.cpu().numpy() for k, v in w.items()}
    sets = (("pm", A.ADZ_A_PM, "k0"), ("pmu", A.ADZ_A_PMU, "k0"), ("pmv", A.ADZ_A_PMV, "k0"),
            ("pt", A.ADZ_A_PT, "k0t"), ("pb", A.ADZ_A_PB, "k0"))
    gpu = {name: ctx.array(which).cpu().numpy() for name, which, _ in sets}
    for klo, khi in BAND:
        say(f"c5: oracle trajectory on levels {klo}..{khi}")
        ref = G.traject(wn, cfg.dt, storage="f32", niter=cfg.niter, band=(klo, khi),
                        want=[s[0] for s in sets])
        for name, which, k0key in sets:
            full = _band_rows(ctx, name, k0key, klo, khi)
            got, exp = gpu[name][full], ref[name][full]
            assert got.size > 0 and np.array_equal(_bits(got), _bits(exp)), \
                (name, klo, khi, int(np.count_nonzero(got != exp)))


def _c5_tracers(cfg, n, torch):
    """q-like smooth field and n-1 plume tracers on the model layout (FP64 on the host)"""
    mod_shape = (cfg.nk, cfg.nj + 2 * MH, cfg.ni + 2 * MH)
    out = [S.smooth_field(mod_shape, np.random.default_rng(cfg.seed + 93), "cuda")]
    out[0] = 0.006 + 0.002 * out[0] / 8.0
    for q in range(1, n):
        out.append(S.plume_tracer(mod_shape, np.random.default_rng(cfg.seed + 31 * q), "cuda"))
    return [f.reshape(-1) for f in out]


def _tiles(config, key):
    with open(os.path.join(ROOT, "gem_amd", "tuned_tiles.json")) as fh:
        return json.load(fh)[config][key]


def test_c5_fp32_form_vs_fp64_oracle_on_band(T, C5):
    torch, L, A, TC = T
    cfg, g, ctx, w = C5
    i = ctx.info
    ntr = 6
    mod_shape = (cfg.nk, cfg.nj + 2 * MH, cfg.ni + 2 * MH)
    mod_off = MH + (cfg.ni + 2 * MH) * MH
    dl = (1 - MH, cfg.ni + MH, 1 - MH, cfg.nj + MH)
    box = (i["i0"], i["in"], i["j0"], i["jn"], i["k0t"], cfg.nk)
    say(f"c5: {ntr} tracers, FP32 persistent launch with lin/min/max (committed tiles)")
    src64 = _c5_tracers(cfg, ntr, torch)
    src = [s.float() for s in src64]
    n_mod = src[0].numel()
    stack = [dict(src=s, dst=torch.zeros(n_mod, device="cuda"), lin=torch.zeros(n_mod, device="cuda"),
                  min=torch.zeros(n_mod, device="cuda"), max=torch.zeros(n_mod, device="cuda")) for s in src]
    pt = ctx.array(A.ADZ_A_PT)
    item = dict(geom=ctx.ztab(True), xyz=pt, box=box, dlayout=dl, ext=False,
                fields=A.field_list(stack, mod_off, mod_off))
    plan = A.Plan([item], L.ADZ_F32, math=L.ADZ_MATH_F32, post=A.ADZ_POST_STORE,
                  tune=_tiles("c5", "tracer_groups"))
    plan.execute()
    torch.cuda.synchronize()
    assert plan.status() == 0
    plan.free()
    zt = O.Ztab(g["zt"][1:cfg.nk + 1], cfg.ni + 2 * MH, cfg.nj + 2 * MH, 0, 0)
    # the same descriptor on the device: the model layout, stencils allowed over all of it
    lvm = TC.Ztab(g["zt"][1:cfg.nk + 1], cfg.ni + 2 * MH, cfg.nj + 2 * MH).set_bounds(1 - MH, cfg.ni + MH, 1 - MH,
                                                                                      cfg.nj + MH)
    src_h = [s.cpu().numpy() for s in src64]
    posh = pt.cpu().numpy()
    i0, in_, j0, jn, k0 = box[:5]
    for klo, khi in BAND:
        full = _band_rows(ctx, "pt", "k0t", klo, khi)
        pos = posh[full]
        if pos.size == 0:
            continue
        say(f"c5: oracle FP64 at levels {max(klo, k0)}..{khi}")
        ijk, zl, coef = TC.vcoef(pt[full], lvm, pos.size // 3)
        rijk, rzl, _ = O.vcoef(zt, pos, zabs_fma=1)
        assert np.array_equal(ijk.cpu().numpy(), rijk) and np.array_equal(zl.cpu().numpy(), rzl), (klo, khi)
        exp = O.zyx1_m_n(zt, src_h, mod_off, pos, O.ORD_AVX2)
        ks = slice(max(klo, k0) - 1, khi)
        for f in range(ntr):
            got = stack[f]["dst"].reshape(mod_shape)[ks, j0 - 1 + MH:jn + MH, i0 - 1 + MH:in_ + MH]
            got = got.double().cpu().numpy()
            e = exp[f].reshape(got.shape)
            rms = np.linalg.norm(got - e) / np.linalg.norm(e)
            mx = np.abs(got - e).max() / np.abs(e).max()
            say(f"c5: field {f} levels {ks.start + 1}..{ks.stop}: RMS {rms:.2e}, max {mx:.2e}")
            assert rms <= 1e-5 and mx <= 1e-5, (f, klo, rms, mx)


def test_c5_fifty_steps_budget(T, C5):
    """50 steps of trajectory + FP32 interpolation with lin/min/max + Bermejo-Conde (legacy
    mass = 1 with the min/max reset) on 3 plume tracers over the whole tile, the tracer
    fed back each step.  The fixer's conserved budget, the tracer mass minus the carried
    deficit (summed exactly here, from the stored FP32 tracer), stays at its start value
    to 1e-13 relative; the stored mass itself moves by no more than the FP32 rounding of
    the tracer (1e-6 relative per step)."""
    torch, L, A, TC = T
    cfg, g, ctx, w = C5
    i = ctx.info
    ntr = 3
    mod_off = MH + (cfg.ni + 2 * MH) * MH
    dl = (1 - MH, cfg.ni + MH, 1 - MH, cfg.nj + MH)
    box = (i["i0"], i["in"], i["j0"], i["jn"], i["k0t"], cfg.nk)
    layout5 = dl + (cfg.nk,)
    scope = (1 + cfg.pil, cfg.ni - cfg.pil, 1 + cfg.pil, cfg.nj - cfg.pil, i["k0t"])
    src = [s.float() for s in _c5_tracers(cfg, ntr + 1, torch)[1:]]
    n_mod = src[0].numel()
    dst = [torch.zeros_like(s) for s in src]
    post = [{k: torch.zeros(n_mod, device="cuda") for k in ("lin", "min", "max")} for _ in range(ntr)]
    airm1, airm0, pkps, mask = S.airmass(cfg, g, device="cuda", mh=MH)
    airm1, airm0, pkps = airm1.reshape(-1).contiguous(), airm0.reshape(-1).contiguous(), pkps.reshape(-1).contiguous()
    airm1.copy_(airm0)  # a steady air mass: time P of one step is time M of the previous
    mask = mask.reshape(-1).contiguous()
    stack = [dict(src=src[q], dst=dst[q], lin=post[q]["lin"], min=post[q]["min"], max=post[q]["max"])
             for q in range(ntr)]
    item = dict(geom=ctx.ztab(True), xyz=ctx.array(A.ADZ_A_PT), box=box, dlayout=dl, ext=False,
                fields=A.field_list(stack, mod_off, mod_off))
    plan = A.Plan([item], L.ADZ_F32, math=L.ADZ_MATH_F32, post=A.ADZ_POST_STORE,
                  tune=_tiles("c5", "tracer_groups"))
    bc = A.BermejoConde([dict(p=src[q], m=dst[q], mass=1, **post[q]) for q in range(ntr)], layout5, scope, mod_off,
                        min_max=True)
    deficit = torch.zeros(ntr, dtype=torch.float64, device="cuda")
    am = airm0.cpu().numpy()
    mk = mask.cpu().numpy()

    def budget():
        return [O.mass(d.cpu().numpy(), am, mk, layout5, scope) for d in src]

    start = budget()
    prev, prev_step = start, 0
    say(f"c5: 50 steps of trajectory + interpolation + fixer on {ntr} tracers")
    for step in range(50):
        ctx.adz_traject(cfg.dt, w["uu_ext"], w["vv_ext"], w["ww_ext"], w["ut0"], w["vt0"], w["zdt0"])
        plan.execute()
        bc(airm1, airm0, mask, deficit, pkps=pkps)
        for s, d in zip(src, dst):
            s.copy_(d)  # the corrected tracer is the next step's source
        if step + 1 in (10, 50):
            torch.cuda.synchronize()
            now = budget()
            dfc = deficit.cpu().numpy()
            for q in range(ntr):
                conserved = now[q] - dfc[q]
                rel = abs(conserved - start[q]) / abs(start[q])
                drift = abs(now[q] - prev[q]) / abs(start[q])
                say(f"c5: step {step + 1} tracer {q}: budget moved {rel:.2e}, stored mass moved {drift:.2e} "
                    f"since step {prev_step}")
                assert rel <= 1e-13, (step, q, rel)
                assert drift <= 1e-6 * (step + 1 - prev_step), (step, q, drift)
            prev, prev_step = now, step + 1
    assert plan.status() == 0
    plan.free()
    bc.free()


# ----------------------------------------------------------------------------- c4
@pytest.fixture(scope="module")
def C4(T):
    torch, L, A, TC = T
    subs = []
    say("c4: 16 subdomains x 40 tracers: winds, tracers")
    for cfg in S.c4_subdomains():
        g = S.grid_desc(cfg)
        ctx = A.AdzContext(g, L.ADZ_F32, itraj=cfg.niter)
        w = {k: v.cuda() for k, v in S.winds(cfg, g).items()}
        mod = (cfg.nk, cfg.nj + 2 * MH, cfg.ni + 2 * MH)
        rng = np.random.default_rng(cfg.seed)
        src64 = [S.lognormal_front_tracer(mod, rng, "cuda").reshape(-1) for _ in range(cfg.ntracers)]
        src = [x.float() for x in src64]
        dst = [torch.zeros_like(x) for x in src]
        i = ctx.info
        subs.append(dict(cfg=cfg, g=g, ctx=ctx, w=w, src64=src64, src=src, dst=dst,
                         box=(i["i0"], i["in"], i["j0"], i["jn"], i["k0t"], cfg.nk),
                         dl=(1 - MH, cfg.ni + MH, 1 - MH, cfg.nj + MH), off=MH + (cfg.ni + 2 * MH) * MH))
    yield subs
    for s in subs:
        s["ctx"].free()


def _c4_traj(s):
    w = s["w"]
    s["ctx"].adz_traject(s["cfg"].dt, w["uu_ext"], w["vv_ext"], w["ww_ext"], w["ut0"], w["vt0"], w["zdt0"])


def _c4_items(A, subs):
    return [dict(geom=s["ctx"].ztab(True), xyz=s["ctx"].array(A.ADZ_A_PT), box=s["box"], dlayout=s["dl"], ext=False,
                 fields=A.field_list([dict(src=a, dst=b) for a, b in zip(s["src"], s["dst"])], s["off"], s["off"]))
            for s in subs]


def test_c4_tuned_plan_matches_single_calls_and_oracle(T, C4):
    torch, L, A, TC = T
    subs = C4
    for s in subs:
        s["ctx"].adz_inittraj()
        _c4_traj(s)
    plan = A.Plan(_c4_items(A, subs), L.ADZ_F32, math=L.ADZ_MATH_F32, tune=_tiles("c4", "plan"))
    say(f"c4: persistent launch with tiles {plan.tune()}")
    plan.execute()
    torch.cuda.synchronize()
    assert plan.status() == 0
    plan.free()
    for n, s in enumerate(subs):
        ref = [torch.zeros_like(x) for x in s["src"]]
        A.adz_tricub_hlt([dict(src=a, dst=b) for a, b in zip(s["src"], ref)], s["ctx"].array(A.ADZ_A_PT),
                         s["ctx"].ztab(True), *s["box"][:5], s["box"][5], s["dl"], ext=False, math=L.ADZ_MATH_F32,
                         src_off=s["off"])
        torch.cuda.synchronize()
        for f, (a, b) in enumerate(zip(s["dst"], ref)):
            assert torch.equal(a, b), (n, f)
    for n in (0, 7, 15):  # smallest, middle and largest subdomains
        s = subs[n]
        cfg, box = s["cfg"], s["box"]
        say(f"c4: subdomain {n} ({cfg.ni}x{cfg.nj}) against the FP64 oracle, 8 tracers")
        pos = s["ctx"].array(A.ADZ_A_PT).cpu().numpy()
        mod = (cfg.nk, cfg.nj + 2 * MH, cfg.ni + 2 * MH)
        zt = O.Ztab(s["g"]["zt"][1:cfg.nk + 1], mod[2], mod[1], 0, 0)
        fs = list(range(0, cfg.ntracers, 5))
        exp = O.zyx1_m_n(zt, [s["src64"][f].cpu().numpy() for f in fs], s["off"], pos, O.ORD_AVX2)
        i0, in_, j0, jn, k0, nk = box
        for e, f in zip(exp, fs):
            got = s["dst"][f].reshape(mod)[k0 - 1:, j0 - 1 + MH:jn + MH, i0 - 1 + MH:in_ + MH].double().cpu().numpy()
            e = e.reshape(got.shape)
            rms = np.linalg.norm(got - e) / np.linalg.norm(e)
            assert rms <= 1e-5, (n, f, rms)


def test_c4_graph_of_twenty_steps_matches_eager(T, C4):
    torch, L, A, TC = T
    subs = C4
    nsteps = 20
    plan = A.Plan(_c4_items(A, subs), L.ADZ_F32, math=L.ADZ_MATH_F32, tune=_tiles("c4", "plan"))

    def steps():
        for _ in range(nsteps):
            for s in subs:
                _c4_traj(s)
            plan.execute()

    def reset():
        for s in subs:
            s["ctx"].adz_inittraj()
            for d in s["dst"]:
                d.fill_(-1.0)

    def snapshot():
        return [[d.clone() for d in s["dst"][::7]] + [s["ctx"].array(A.ADZ_A_PT).clone(),
                                                     s["ctx"].array(A.ADZ_A_WPXYZ).clone()] for s in subs]

    say("c4: 20 steps eagerly")
    reset()
    steps()
    torch.cuda.synchronize()
    eager = snapshot()
    say("c4: capture 20 steps")
    graph = torch.cuda.CUDAGraph()
    st = torch.cuda.Stream()
    st.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(st):
        reset()
        torch.cuda.synchronize()
        with torch.cuda.graph(graph, stream=st):
            steps()
    torch.cuda.current_stream().wait_stream(st)
    for rep in range(2):
        say(f"c4: replay {rep}")
        reset()
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        for a, b in zip(snapshot(), eager):
            for x, y in zip(a, b):
                assert torch.equal(x, y), rep
        steps()  # eager work between replays
        torch.cuda.synchronize()
    assert plan.status() == 0
    plan.free()
    del graph
