"""GPU: the Fortran drop-in layer's step routines, adz_main() and adz_tracers(.true.) /
adz_tracers(.false.) (adz_main.F90:16-71, adz_tracers.F90:16-184, as dynstep.F90:93-98
calls them), run through module adz_gpu (libadz_gpu.so) on a tile with every tracer group
of tr3d: tricubic and quintic tracers without post-processing, and both groups with it
(Tr_3CWP tricubic, Tr_BQWP quintic) with clipping, ILMC, the LAM boundary flux and the
Bermejo-Conde fixer over the registry of both groups (adz_post_tr(1), (2), (0)), after
set_post_tr_hlt forms the air masses from the pressures.

Checks:
  * the trajectory of the module equals the C ABI's bit for bit;
  * u on Adz_pmu, v on Adz_pmv, rhsc on Adz_pm (momentum levels), rhst and rhsf on Adz_pt
    (thermodynamic levels) equal the oracle's interpolation at those positions bit for bit
    (the reference's arithmetic);
  * every tracer, the lin/min/max arrays and the carried deficits equal the same sequence
    called through the C ABI bit for bit;
  * the whole step captured in a hipGraph (adz_inittraj + the step) and replayed twice
    equals the eager step bit for bit.
"""
import numpy as np
import pytest

from gem_amd import synth as S
from oracle import oracle as O

pytestmark = pytest.mark.gpu
MH = 2
WP_MASS, WP_MONO = [1, 112, 121, 0], [1, 0, 2, 0]
BQ_MASS, BQ_MONO = [1, 0], [2, 1]


def _uses_bc(m):
    return m == 1 or 111 <= m <= 139


@pytest.fixture(scope="module")
def W(cuda):
    import torch

    from gem_amd import _lib as L
    from gem_amd import adz as A
    from gem_amd import adz_fortran as AF

    cfg = S.Config("fs", 48, 40, 16, 4, 300.0, jets=True, cmax_xy=3.0, cmax_z=1.0, seed=97)
    g = S.grid_desc(cfg)
    nk = cfg.nk
    w = {k: v.cuda() for k, v in S.winds(cfg, g).items()}
    mod = (nk, cfg.nj + 2 * MH, cfg.ni + 2 * MH)
    n_mod = int(np.prod(mod))
    rng = np.random.default_rng(7)
    rhs_src = [f.float().cuda() for f in S.fields(cfg, 5, seed=cfg.seed + 1)]          # u v c t f (extended)
    trc = [S.plume_tracer(mod, rng, "cpu").float().reshape(-1).cuda() for _ in range(9)]
    # pressures of gmm_pw and the area (set_post_tr_hlt's inputs)
    shape2 = mod[1:]
    ps = rng.uniform(0.96, 1.04, shape2)
    pm = [torch.from_numpy((np.exp(g["zm"][1:nk + 2])[:, None, None] * (ps * rng.uniform(0.999, 1.001, shape2))[None])
                           .astype(np.float32).reshape(-1)).cuda() for _ in range(2)]
    pt = torch.from_numpy((np.exp(g["zt"][1:nk + 1])[:, None, None] * ps[None]).astype(np.float32).reshape(-1)).cuda()
    p0 = torch.from_numpy((1e5 * ps).astype(np.float32).reshape(-1)).cuda()
    area = torch.from_numpy(((S.RAYT * g["h"]) ** 2 * rng.uniform(0.9, 1.0, shape2)).reshape(-1)).cuda()
    mask = torch.ones(shape2[0] * shape2[1], dtype=torch.float64, device="cuda")
    yield dict(torch=torch, L=L, A=A, AF=AF, cfg=cfg, g=g, w=w, mod=mod, n_mod=n_mod, rhs_src=rhs_src, trc=trc,
               pm=pm, pt=pt, p0=p0, area=area, mask=mask)


def _fortran(W, tune=None):
    """module adz_gpu set up on fresh destination arrays; returns (layer, arrays)"""
    torch, L, AF, cfg, g = W["torch"], W["L"], W["AF"], W["cfg"], W["g"]
    n_mod, nk = W["n_mod"], cfg.nk
    layout = (1 - MH, cfg.ni + MH, 1 - MH, cfg.nj + MH)
    off = MH + (cfg.ni + 2 * MH) * MH
    F = AF.FortranStep(g, layout, itraj=cfg.niter, pos_dtype=L.ADZ_F32, field_dtype=L.ADZ_F32, math=L.ADZ_MATH_REF)
    i = F.ctx.info
    ext_off = (1 - i["lminx"]) + i["nit"] * (1 - i["lminy"])
    z = lambda: torch.full((n_mod,), -5.0, device="cuda")
    arr = dict(rhs_dst=[z() for _ in range(5)], M=[z() for _ in range(9)],
               post=[{k: z() for k in ("lin", "min", "max")} for _ in range(6)],
               airm1=z(), airm0=z(), pkps=z())
    F.set_time(cfg.dt)
    F.set_winds(W["w"])
    F.set_rhs(W["rhs_src"] + [None], arr["rhs_dst"] + [None], ext_off, off, hydro=True)
    F.set_tracers(W["trc"], arr["M"], off=off)
    F.set_groups(tricub_nt=(1, 2), bichqv_nt=(3, 1), tricub_wp=(4, 4), bichqv_wp=(8, 2))
    F.set_wp(1, WP_MASS, WP_MONO, arr["post"][:4], off=off)
    F.set_wp(2, BQ_MASS, BQ_MONO, arr["post"][4:], off=off)
    F.set_bc_fields(arr["airm1"], arr["airm0"], arr["pkps"], W["mask"], off, MH + (cfg.ni + 2 * MH) * MH)
    F.set_airmass(W["pm"][0], W["pm"][1], W["pt"], W["p0"], W["area"], off, MH + (cfg.ni + 2 * MH) * MH)
    F.set_options()
    if tune:
        F.set_tune(0, tune)
        F.set_tune(1, tune)
    return F, arr, off, ext_off


def _snapshot(F, arr, A):
    out = [t.clone() for t in arr["rhs_dst"] + arr["M"]]
    out += [q[k].clone() for q in arr["post"] for k in ("lin", "min", "max")]
    out += [arr["airm1"].clone(), arr["airm0"].clone(), arr["pkps"].clone(), F.array(A.ADZ_A_PT).clone()]
    return out


def _bits(t):
    return t.cpu().numpy().view(np.uint32)


@pytest.mark.parametrize("tiles", ["default", "stage4"])
def test_fortran_step_matches_oracle_and_c_abi(W, tiles):
    torch, L, A = W["torch"], W["L"], W["A"]
    cfg, g, nk, n_mod = W["cfg"], W["g"], W["cfg"].nk, W["n_mod"]
    tune = None if tiles == "default" else dict(fields_per_pass=2, points_outer=0, blocks_per_cu=4, xcd_chunks=1,
                                                stage=4)
    F, arr, off, ext_off = _fortran(W, tune)
    print("[gpu-case] Fortran adz_main + adz_tracers", tiles, flush=True)
    F.adz_inittraj()
    F.step()
    torch.cuda.synchronize()
    fdef1 = F.deficits(1, 4)
    fdef2 = F.deficits(2, 2)

    # ---- the same through the C ABI
    ctx = A.AdzContext(g, L.ADZ_F32, itraj=cfg.niter)
    w = W["w"]
    ctx.adz_inittraj()
    ctx.adz_traject(cfg.dt, w["uu_ext"], w["vv_ext"], w["ww_ext"], w["ut0"], w["vt0"], w["zdt0"])
    i = ctx.info
    for which in (A.ADZ_A_PM, A.ADZ_A_PMU, A.ADZ_A_PMV, A.ADZ_A_PT, A.ADZ_A_PB):
        assert np.array_equal(_bits(F.array(which)), _bits(ctx.array(which))), which

    # adz_main: every group against the oracle at the module's positions
    mod_shape = W["mod"]
    groups = [(0, A.ADZ_A_PMU, False, (i["i0u"], i["inu"], i["j0"], i["jn"], i["k0"])),
              (1, A.ADZ_A_PMV, False, (i["i0"], i["in"], i["j0v"], i["jnv"], i["k0"])),
              (2, A.ADZ_A_PM, False, (i["i0"], i["in"], i["j0"], i["jn"], i["k0"])),
              (3, A.ADZ_A_PT, True, (i["i0"], i["in"], i["j0"], i["jn"], i["k0t"])),
              (4, A.ADZ_A_PT, True, (i["i0"], i["in"], i["j0"], i["jn"], i["k0t"]))]
    for q, which, thermo, box in groups:
        lev = (g["zt"] if thermo else g["zm"])[1:nk + 1]
        zt = O.Ztab(lev, i["nit"], i["njt"], 0, 0)
        exp = O.zyx1_m_n(zt, [W["rhs_src"][q].cpu().numpy()], ext_off, F.array(which).cpu().numpy(), O.ORD_AVX2)[0]
        i0, in_, j0, jn, k0 = box
        got = arr["rhs_dst"][q].cpu().numpy().reshape(mod_shape)[k0 - 1:, j0 - 1 + MH:jn + MH, i0 - 1 + MH:in_ + MH]
        assert np.array_equal(got.view(np.uint32), exp.reshape(got.shape).view(np.uint32)), q

    # adz_tracers(.true.): the interpolation of the four groups
    dl = (1 - MH, cfg.ni + MH, 1 - MH, cfg.nj + MH)
    box = (i["i0"], i["in"], i["j0"], i["jn"], i["k0t"])
    trc = W["trc"]
    M = [torch.full((n_mod,), -5.0, device="cuda") for _ in range(9)]
    post = [{k: torch.full((n_mod,), -5.0, device="cuda") for k in ("lin", "min", "max")} for _ in range(6)]
    ent = lambda n: dict(src=trc[n], dst=M[n])
    A.adz_tricub_hlt([ent(0), ent(1)], ctx.array(A.ADZ_A_PT), ctx.ztab(True), *box, nk, dl, ext=False,
                     math=L.ADZ_MATH_REF, src_off=off)
    A.adz_tricub_hlt_qv(ctx, [ent(2)], ctx.array(A.ADZ_A_PT), *box, dl, ext=False, post=0, src_off=off)
    wp = [dict(ent(3 + q), **post[q]) for q in range(4)]
    bq = [dict(ent(7 + q), **post[4 + q]) for q in range(2)]
    for clip in (0, 1):
        pf = A.ADZ_POST_STORE | (A.ADZ_POST_CLIP if clip else 0)
        sel = [e for e, m in zip(wp, WP_MONO) if (m == 1) == bool(clip)]
        if sel:
            A.adz_tricub_hlt(sel, ctx.array(A.ADZ_A_PT), ctx.ztab(True), *box, nk, dl, ext=False, post=pf,
                             math=L.ADZ_MATH_REF, src_off=off)
        sel = [e for e, m in zip(bq, BQ_MONO) if (m == 1) == bool(clip)]
        if sel:
            A.adz_tricub_hlt_qv(ctx, sel, ctx.array(A.ADZ_A_PT), *box, dl, ext=False, post=pf, src_off=off)
    # adz_tracers(.false.): air masses, then per group the flux and ILMC, then the fixer
    a1, a0, pk = (torch.full((n_mod,), -5.0, device="cuda") for _ in range(3))
    lay5 = dl + (nk,)
    A.adz_set_post_tr(ctx, lay5, W["pm"][0], W["pm"][1], W["pt"], W["p0"], W["area"], a1, a0, pk)
    scope = (1 + g["pil_w"], cfg.ni - g["pil_e"], 1 + g["pil_s"], cfg.nj - g["pil_n"], i["k0t"])
    fmass = []
    bcs = []
    for grp, mass, mono in ((wp, WP_MASS, WP_MONO), (bq, BQ_MASS, BQ_MONO)):
        sel = [e for e, m in zip(grp, mass) if _uses_bc(m)]
        fx = A.BcLamFlux([dict(src=e["src"]) for e in sel], lay5, (cfg.ni, cfg.nj),
                         (i["i0b"], i["inb"], i["j0b"], i["jnb"]), scope, off, ctx.ztab(True), pos_dtype=L.ADZ_F32)
        fm = torch.zeros(2 * len(sel), dtype=torch.float64, device="cuda")
        fx(ctx.array(A.ADZ_A_PB), a0, W["mask"], fm)
        fmass.append(fm)
        il = [dict(f=e["dst"], min=e["min"], max=e["max"]) for e, m in zip(grp, mono) if m == 2]
        if il:
            ilmc = A.IlmcLam(il, lay5, box, off)
            ilmc(a0)
            torch.cuda.synchronize()
            ilmc.free()
        bcs += [dict(p=e["src"], m=e["dst"], lin=e["lin"], min=e["min"], max=e["max"], mass=int(m))
                for e, m in zip(grp, mass) if _uses_bc(m)]
        torch.cuda.synchronize()
        fx.free()
    bc = A.BermejoConde(bcs, lay5, scope, off, min_max=True)
    deficit = torch.zeros(len(bcs), dtype=torch.float64, device="cuda")
    bc(a1, a0, W["mask"], deficit, pkps=pk, flux=torch.cat(fmass))
    torch.cuda.synchronize()
    bc.free()

    for n in range(9):
        assert np.array_equal(_bits(arr["M"][n]), _bits(M[n])), ("tracer", n)
    for q in range(6):
        for k in ("lin", "min", "max"):
            assert np.array_equal(_bits(arr["post"][q][k]), _bits(post[q][k])), ("post", q, k)
    for got, exp in ((arr["airm1"], a1), (arr["airm0"], a0), (arr["pkps"], pk)):
        assert np.array_equal(_bits(got), _bits(exp))
    exp_def = deficit.cpu().numpy()
    got_def = [fdef1[q] for q in range(4) if _uses_bc(WP_MASS[q])] + \
              [fdef2[q] for q in range(2) if _uses_bc(BQ_MASS[q])]
    assert np.array_equal(np.array(got_def).view(np.uint64), exp_def.view(np.uint64))
    assert np.all(exp_def != 0.0)
    ctx.free()
    F.free()


def test_fortran_step_graph_replay(W):
    """adz_inittraj + the step captured once in a hipGraph on a created stream; two replays
    (with eager work between them) equal the eager step bit for bit"""
    torch, A = W["torch"], W["A"]
    F, arr, off, ext_off = _fortran(W)
    print("[gpu-case] Fortran step: eager", flush=True)
    F.adz_inittraj()
    F.step()
    torch.cuda.synchronize()
    eager = _snapshot(F, arr, A)
    edef = F.deficits(1, 4) + F.deficits(2, 2)
    print("[gpu-case] Fortran step: capture", flush=True)
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    graph = torch.cuda.CUDAGraph()
    F.set_deficits(1, [0.0] * 4)
    F.set_deficits(2, [0.0] * 2)
    with torch.cuda.stream(s):
        F.set_stream(s)
        with torch.cuda.graph(graph, stream=s):
            F.adz_inittraj()
            F.step()
    torch.cuda.current_stream().wait_stream(s)
    F.set_stream(torch.cuda.current_stream())
    for rep in range(2):
        print(f"[gpu-case] Fortran step: replay {rep}", flush=True)
        for t in arr["M"] + arr["rhs_dst"]:
            t.fill_(-5.0)
        F.set_deficits(1, [0.0] * 4)
        F.set_deficits(2, [0.0] * 2)
        graph.replay()
        torch.cuda.synchronize()
        got = _snapshot(F, arr, A)
        for n, (a, b) in enumerate(zip(got, eager)):
            assert torch.equal(a.view(torch.int32), b.view(torch.int32)), (rep, n)
        assert F.deficits(1, 4) + F.deficits(2, 2) == edef
        F.adz_inittraj()  # eager work between the replays
        F.adz_main()
        torch.cuda.synchronize()
    del graph
    F.free()
