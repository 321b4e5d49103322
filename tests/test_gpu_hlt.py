"""GPU parity of adz_tricub_hlt (fused multi-field interpolation into the destination
arrays, with the Adz_post lin/min/max and mono clipping) and of the persistent batched
plan over several subdomains.

Reference arithmetic is bit-exact against the oracle's interpolation followed by the
scatter of adz_interp_hlt.F90:165-180; the FP32 form is held to |err| <= 5e-6*max|f|
against the FP64 oracle; plans give bit-identical results to single-item calls for every
tile parameter set."""
import numpy as np
import pytest

from gem_amd import synth as S
from oracle import oracle as O

pytestmark = pytest.mark.gpu
MH = 2  # model halo of the destination arrays


@pytest.fixture(scope="module")
def T(cuda):
    import torch

    from gem_amd import _lib as L
    from gem_amd import adz as A

    return torch, L, A


def _setup(T, cfg, pos_dtype, fdt, nfields=3, seed=3):
    torch, L, A = T
    g = S.grid_desc(cfg)
    ctx = A.AdzContext(g, pos_dtype)
    w = {k: v.cuda() for k, v in S.winds(cfg, g).items()}
    ctx.adz_inittraj()
    ctx.adz_traject(cfg.dt, w["uu_ext"], w["vv_ext"], w["ww_ext"], w["ut0"], w["vt0"], w["zdt0"])
    src64 = S.fields(cfg, nfields, seed=seed)
    return g, ctx, [f.numpy() for f in src64]


def _expected(cfg, g, ctx, src, pos, mono, thermo=True):
    """oracle interpolation at the target box, scattered into (MH-halo) model arrays."""
    i = ctx.info
    nit, njt = i["nit"], i["njt"]
    lev = g["zt"][1:cfg.nk + 1] if thermo else g["zm"][1:cfg.nk + 1]
    zt = O.Ztab(lev, nit, njt, 0, 0)
    base = (1 - i["lminx"]) + nit * (1 - i["lminy"])
    if mono:
        outs = O.mono_m_n(zt, src, base, pos, O.ORD_AVX2)
    else:
        outs = [O.zyx1_m_n(zt, src, base, pos, O.ORD_AVX2)]
    return outs


def _scatter(vals, cfg, box, fill):
    i0, in_, j0, jn, k0 = box
    ni, nj, nk = cfg.ni + 2 * MH, cfg.nj + 2 * MH, cfg.nk
    out = np.full((len(vals), nk, nj, ni), fill, dtype=vals.dtype)
    nz = nk - k0 + 1
    v = vals.reshape(len(vals), nz, jn - j0 + 1, in_ - i0 + 1)
    out[:, k0 - 1:, j0 - 1 + MH:jn + MH, i0 - 1 + MH:in_ + MH] = v
    return out


def _bits_equal(a, b):
    a, b = np.ascontiguousarray(a), np.ascontiguousarray(b)
    return a.shape == b.shape and np.array_equal(a.view(np.uint8), b.view(np.uint8))


CFG = S.Config("h", 72, 56, 24, 4, 300.0, jets=True, cmax_xy=3.0, cmax_z=1.0, seed=41)


@pytest.mark.parametrize("post", [0, 1, 3])
@pytest.mark.parametrize("fdt", ["f32", "f64"])
def test_hlt_reference_math_bit_exact(T, post, fdt):
    torch, L, A = T
    g, ctx, src64 = _setup(T, CFG, L.ADZ_F32, fdt)
    npdt = np.float32 if fdt == "f32" else np.float64
    tdt = torch.float32 if fdt == "f32" else torch.float64
    src = [s.astype(npdt) for s in src64]
    i = ctx.info
    box = (i["i0"], i["in"], i["j0"], i["jn"], i["k0t"])
    pt = ctx.array(A.ADZ_A_PT)
    n_model = (CFG.ni + 2 * MH) * (CFG.nj + 2 * MH) * CFG.nk
    fill = -777.0
    stack = []
    for s in src:
        e = {"src": torch.from_numpy(s).cuda(), "dst": torch.full((n_model,), fill, dtype=tdt, device="cuda")}
        if post:
            for k in ("lin", "min", "max"):
                e[k] = torch.full((n_model,), fill, dtype=tdt, device="cuda")
        stack.append(e)
    src_off = (1 - i["lminx"]) + i["nit"] * (1 - i["lminy"])
    A.adz_tricub_hlt(stack, pt, ctx.ztab(True), *box[:4], box[4], CFG.nk,
                     (1 - MH, CFG.ni + MH, 1 - MH, CFG.nj + MH), ext=True, post=post, src_off=src_off)
    torch.cuda.synchronize()
    pos = pt.cpu().numpy()
    outs = _expected(CFG, g, ctx, src, pos, mono=bool(post))
    d, = [_scatter(outs[0], CFG, box, fill)] if not post else [None]
    if post:
        cub, lin, mi, ma = outs
        dexp = np.maximum(mi, np.minimum(ma, cub)) if post & 2 else cub
        d = _scatter(dexp, CFG, box, fill)
        lexp = np.maximum(mi, np.minimum(ma, lin))
        for key, arr in (("lin", lexp), ("min", mi), ("max", ma)):
            got = np.stack([e[key].cpu().numpy() for e in stack]).reshape(d.shape)
            assert _bits_equal(got, _scatter(arr, CFG, box, fill)), key
    got = np.stack([e["dst"].cpu().numpy() for e in stack]).reshape(d.shape)
    assert _bits_equal(got, d)
    ctx.free()


def test_hlt_model_layout_source(T):
    """ext = 0: sources on the model layout, read in place."""
    torch, L, A = T
    g, ctx, src64 = _setup(T, CFG, L.ADZ_F32, "f32")
    i = ctx.info
    h = CFG.halo
    # cut the model layout (halo MH) out of the extended (halo h) fields
    src_ext = [s.astype(np.float32).reshape(CFG.nk, CFG.nj + 2 * h, CFG.ni + 2 * h) for s in src64]
    src_mod = [np.ascontiguousarray(s[:, h - MH:h + CFG.nj + MH, h - MH:h + CFG.ni + MH]).reshape(-1) for s in src_ext]
    n_model = src_mod[0].size
    box = (i["i0"], i["in"], i["j0"], i["jn"], i["k0t"])
    stack = [{"src": torch.from_numpy(s).cuda(), "dst": torch.zeros(n_model, device="cuda")} for s in src_mod]
    A.adz_tricub_hlt(stack, ctx.array(A.ADZ_A_PT), ctx.ztab(True), *box[:4], box[4], CFG.nk,
                     (1 - MH, CFG.ni + MH, 1 - MH, CFG.nj + MH), ext=False)
    torch.cuda.synchronize()
    pos = ctx.array(A.ADZ_A_PT).cpu().numpy()
    zt = O.Ztab(g["zt"][1:CFG.nk + 1], CFG.ni + 2 * MH, CFG.nj + 2 * MH, 0, 0)
    base = MH + (CFG.ni + 2 * MH) * MH
    exp = _scatter(O.zyx1_m_n(zt, src_mod, base, pos, O.ORD_AVX2), CFG, box, 0.0)
    got = np.stack([e["dst"].cpu().numpy() for e in stack]).reshape(exp.shape)
    assert _bits_equal(got, exp)
    ctx.free()


def test_hlt_f32_math_error_vs_fp64_oracle(T):
    torch, L, A = T
    g, ctx, src64 = _setup(T, CFG, L.ADZ_F64, "f32", nfields=4)
    i = ctx.info
    box = (i["i0"], i["in"], i["j0"], i["jn"], i["k0t"])
    n_model = (CFG.ni + 2 * MH) * (CFG.nj + 2 * MH) * CFG.nk
    stack = [{"src": torch.from_numpy(s).float().cuda(), "dst": torch.zeros(n_model, device="cuda")} for s in src64]
    src_off = (1 - i["lminx"]) + i["nit"] * (1 - i["lminy"])
    A.adz_tricub_hlt(stack, ctx.array(A.ADZ_A_PT), ctx.ztab(True), *box[:4], box[4], CFG.nk,
                     (1 - MH, CFG.ni + MH, 1 - MH, CFG.nj + MH), math=L.ADZ_MATH_F32, src_off=src_off)
    torch.cuda.synchronize()
    pos = ctx.array(A.ADZ_A_PT).cpu().numpy()
    exp = _scatter(_expected(CFG, g, ctx, src64, pos, mono=False)[0], CFG, box, 0.0)
    got = np.stack([e["dst"].cpu().numpy() for e in stack]).reshape(exp.shape).astype(np.float64)
    for f in range(len(src64)):
        scale = np.abs(src64[f]).max()
        assert np.abs(got[f] - exp[f]).max() <= 5e-6 * scale
    ctx.free()


def _plan_case(T):
    torch, L, A = T
    cfgs = [S.Config(f"p{s}", 40 + 16 * s, 36 + 8 * s, 16, 4, 300.0, jets=True, cmax_xy=2.0, seed=60 + s)
            for s in range(3)]
    items, refs = [], []
    keep = []
    for cfg in cfgs:
        g, ctx, src64 = _setup(T, cfg, L.ADZ_F64, "f32", nfields=5, seed=cfg.seed)
        keep.append(ctx)
        i = ctx.info
        box = (i["i0"], i["in"], i["j0"], i["jn"], i["k0t"], cfg.nk)
        n_model = (cfg.ni + 2 * MH) * (cfg.nj + 2 * MH) * cfg.nk
        srcs = [torch.from_numpy(s).float().cuda() for s in src64]
        stack = [{"src": s, "dst": torch.zeros(n_model, device="cuda")} for s in srcs]
        ref = [{"src": s, "dst": torch.zeros(n_model, device="cuda")} for s in srcs]
        src_off = (1 - i["lminx"]) + i["nit"] * (1 - i["lminy"])
        dl = (1 - MH, cfg.ni + MH, 1 - MH, cfg.nj + MH)
        dst_off = (1 - dl[0]) + (dl[1] - dl[0] + 1) * (1 - dl[2])
        A.adz_tricub_hlt(ref, ctx.array(A.ADZ_A_PT), ctx.ztab(True), *box[:5], box[5], dl, math=L.ADZ_MATH_F32,
                         src_off=src_off)
        items.append(dict(geom=ctx.ztab(True), xyz=ctx.array(A.ADZ_A_PT), box=box, dlayout=dl,
                          fields=A.field_list(stack, src_off, dst_off)))
        refs.append((stack, ref))
    return items, refs, keep


@pytest.mark.parametrize("tune", [None, dict(fields_per_pass=1, points_outer=0, blocks_per_cu=2, xcd_chunks=0),
                                  dict(fields_per_pass=3, points_outer=1, blocks_per_cu=4, xcd_chunks=1),
                                  dict(fields_per_pass=5, points_outer=0, blocks_per_cu=4, xcd_chunks=1, stage=1),
                                  dict(fields_per_pass=2, points_outer=1, blocks_per_cu=8, xcd_chunks=0, stage=1),
                                  dict(fields_per_pass=4, points_outer=1, blocks_per_cu=4, xcd_chunks=1, stage=2,
                                       level_inner=1),
                                  dict(fields_per_pass=3, points_outer=0, blocks_per_cu=4, xcd_chunks=1, stage=1,
                                       level_inner=1),
                                  dict(fields_per_pass=5, points_outer=0, blocks_per_cu=3, xcd_chunks=1, stage=3),
                                  dict(fields_per_pass=2, points_outer=1, blocks_per_cu=2, xcd_chunks=0, stage=3,
                                       level_inner=1),
                                  dict(fields_per_pass=4, points_outer=0, blocks_per_cu=4, xcd_chunks=1, stage=2,
                                       level_inner=4),
                                  dict(fields_per_pass=5, points_outer=1, blocks_per_cu=2, xcd_chunks=0, stage=1,
                                       level_inner=3),
                                  dict(fields_per_pass=3, points_outer=0, blocks_per_cu=3, xcd_chunks=1, stage=3,
                                       level_inner=2),
                                  dict(fields_per_pass=2, points_outer=0, blocks_per_cu=4, xcd_chunks=1, stage=4,
                                       level_inner=4),
                                  dict(fields_per_pass=5, points_outer=0, blocks_per_cu=4, xcd_chunks=1, stage=5),
                                  dict(fields_per_pass=2, points_outer=1, blocks_per_cu=2, xcd_chunks=0, stage=5,
                                       level_inner=3),
                                  dict(fields_per_pass=5, points_outer=0, blocks_per_cu=4, xcd_chunks=1, stage=6,
                                       level_inner=3),
                                  dict(fields_per_pass=1, points_outer=1, blocks_per_cu=2, xcd_chunks=0, stage=6),
                                  dict(fields_per_pass=5, points_outer=0, blocks_per_cu=4, xcd_chunks=1, stage=7,
                                       level_inner=3),
                                  dict(fields_per_pass=2, points_outer=1, blocks_per_cu=2, xcd_chunks=0, stage=7),
                                  dict(fields_per_pass=5, points_outer=0, blocks_per_cu=1, xcd_chunks=1, stage=8,
                                       level_inner=3),
                                  dict(fields_per_pass=1, points_outer=1, blocks_per_cu=2, xcd_chunks=0, stage=8,
                                       level_inner=1)])
def test_plan_matches_single_calls(T, tune):
    torch, L, A = T
    items, refs, keep = _plan_case(T)
    plan = A.Plan(items, L.ADZ_F32, math=L.ADZ_MATH_F32, tune=tune)
    plan.execute()
    plan.execute()  # idempotent
    torch.cuda.synchronize()
    for stack, ref in refs:
        for a, b in zip(stack, ref):
            assert torch.equal(a["dst"], b["dst"])
    assert plan.status() == 0   # no staging wait gave up
    plan.free()
    for c in keep:
        c.free()


def test_stage3_waits_that_give_up_fall_back_exactly(T):
    """The staging hand-off of mode 3 forced to give up from the host (a spin bound of 0
    polls, so every wait that is not already satisfied gives up at once; no kernel is
    made to hang): the plan still equals the single calls bit for bit (the fields whose
    slots were not published are read from global memory), and the plan's status word
    counts the waits that gave up.  With the default bound, a later run reports none."""
    torch, L, A = T
    items, refs, keep = _plan_case(T)
    plan = A.Plan(items, L.ADZ_F32, math=L.ADZ_MATH_F32,
                  tune=dict(fields_per_pass=5, points_outer=0, blocks_per_cu=3, xcd_chunks=1, stage=3))
    plan.set_spin_limit(0)
    for stack, _ in refs:
        for a in stack:
            a["dst"].zero_()
    plan.execute()
    torch.cuda.synchronize()
    for stack, ref in refs:
        for a, b in zip(stack, ref):
            assert torch.equal(a["dst"], b["dst"])
    assert plan.status() > 0
    assert plan.status() == 0   # read and reset
    plan.set_spin_limit(1 << 20)
    plan.execute()
    torch.cuda.synchronize()
    assert plan.status() == 0
    for stack, ref in refs:
        for a, b in zip(stack, ref):
            assert torch.equal(a["dst"], b["dst"])
    plan.free()
    for c in keep:
        c.free()


@pytest.mark.parametrize("post", [0, 1, 3])
@pytest.mark.parametrize("math", ["f32", "ref"])
@pytest.mark.parametrize("pdt", ["f32", "f64"])
def test_shared_columns_match_single_calls(T, post, math, pdt):
    """Staging mode 4 (stencil columns shared across a wave, 61-target tiles) against one
    adz_tricub_hlt call per item (per-lane stencils): bit for bit, with lin/min/max stored
    and clipped, in both arithmetics, with FP32 and FP64 positions, over items whose winds
    reach Courant 3 (targets whose neighbours skip or repeat a cell take the per-lane
    path) and over every tile order."""
    torch, L, A = T
    M = L.ADZ_MATH_F32 if math == "f32" else L.ADZ_MATH_REF
    P = L.ADZ_F32 if pdt == "f32" else L.ADZ_F64
    cfgs = [S.Config("s0", 130, 37, 12, 4, 300.0, jets=True, cmax_xy=3.0, cmax_z=1.0, seed=81),
            S.Config("s1", 61, 26, 9, 4, 300.0, seed=82),
            S.Config("s2", 200, 24, 10, 4, 300.0, jets=True, cmax_xy=1.0, seed=83)]
    items, refs, keep = [], [], []
    for cfg in cfgs:
        g, ctx, src64 = _setup(T, cfg, P, "f32", nfields=3, seed=cfg.seed)
        keep.append(ctx)
        i = ctx.info
        box = (i["i0"], i["in"], i["j0"], i["jn"], i["k0t"], cfg.nk)
        n_model = (cfg.ni + 2 * MH) * (cfg.nj + 2 * MH) * cfg.nk
        srcs = [torch.from_numpy(s).float().cuda() for s in src64]

        def stack_of():
            d = {"dst": torch.full((n_model,), -7.0, device="cuda")}
            if post & 1:
                d.update({k: torch.full((n_model,), -7.0, device="cuda") for k in ("lin", "min", "max")})
            return d
        stack = [dict(src=s, **stack_of()) for s in srcs]
        ref = [dict(src=s, **stack_of()) for s in srcs]
        src_off = (1 - i["lminx"]) + i["nit"] * (1 - i["lminy"])
        dl = (1 - MH, cfg.ni + MH, 1 - MH, cfg.nj + MH)
        dst_off = (1 - dl[0]) + (dl[1] - dl[0] + 1) * (1 - dl[2])
        A.adz_tricub_hlt(ref, ctx.array(A.ADZ_A_PT), ctx.ztab(True), *box[:5], box[5], dl, math=M, post=post,
                         src_off=src_off)
        items.append(dict(geom=ctx.ztab(True), xyz=ctx.array(A.ADZ_A_PT), box=box, dlayout=dl,
                          fields=A.field_list(stack, src_off, dst_off)))
        refs.append((stack, ref))
    for tune in (dict(fields_per_pass=2, points_outer=0, blocks_per_cu=4, xcd_chunks=1, stage=4),
                 dict(fields_per_pass=3, points_outer=1, blocks_per_cu=2, xcd_chunks=0, stage=4, level_inner=1),
                 dict(fields_per_pass=1, points_outer=0, blocks_per_cu=3, xcd_chunks=1, stage=4, level_inner=3),
                 dict(fields_per_pass=3, points_outer=0, blocks_per_cu=4, xcd_chunks=1, stage=5),
                 dict(fields_per_pass=1, points_outer=1, blocks_per_cu=2, xcd_chunks=0, stage=5, level_inner=2),
                 dict(fields_per_pass=3, points_outer=0, blocks_per_cu=4, xcd_chunks=1, stage=6, level_inner=3),
                 dict(fields_per_pass=2, points_outer=1, blocks_per_cu=2, xcd_chunks=0, stage=6),
                 dict(fields_per_pass=3, points_outer=0, blocks_per_cu=4, xcd_chunks=1, stage=7, level_inner=3),
                 dict(fields_per_pass=1, points_outer=1, blocks_per_cu=2, xcd_chunks=0, stage=7, level_inner=1),
                 dict(fields_per_pass=3, points_outer=0, blocks_per_cu=1, xcd_chunks=1, stage=8, level_inner=2),
                 dict(fields_per_pass=2, points_outer=1, blocks_per_cu=2, xcd_chunks=0, stage=8)):
        if tune["stage"] == 8 and M != L.ADZ_MATH_F32:  # the 16-wave form is FP32 arithmetic only
            with pytest.raises(RuntimeError, match="stage 8"):
                A.Plan(items, L.ADZ_F32, math=M, post=post, tune=tune)
            continue
        for stack, _ in refs:
            for a in stack:
                for k in ("dst", "lin", "min", "max"):
                    if k in a:
                        a[k].fill_(-7.0)
        plan = A.Plan(items, L.ADZ_F32, math=M, post=post, tune=tune)
        assert plan.tune()["stage"] == tune["stage"]
        plan.execute()
        torch.cuda.synchronize()
        for stack, ref in refs:
            for a, b in zip(stack, ref):
                for k in ("dst", "lin", "min", "max"):
                    if k in a:
                        assert torch.equal(a[k], b[k]), (tune, k)
        plan.free()
    for c in keep:
        c.free()


@pytest.mark.parametrize("stage", [4, 5, 6, 7, 8])
@pytest.mark.parametrize("math", ["f32", "ref"])
def test_shared_columns_edge_positions(T, math, stage):
    """Staging modes 4 to 8 with positions that break lane sharing and box staging on
    purpose: stencils out of range (NaN targets whose lanes serve no neighbour), positions
    on cell faces, a cell repeated or skipped every few targets, levels clamped at both
    ends, stencils on the first columns of the source.  Bit for bit the single-call result
    (NaN patterns included), with lin/min/max."""
    torch, L, A = T
    M = L.ADZ_MATH_F32 if math == "f32" else L.ADZ_MATH_REF
    if stage == 8 and math == "ref":
        pytest.skip("staging mode 8 is FP32 arithmetic only (refused, checked in the test above)")
    cfg = S.Config("e", 150, 26, 11, 4, 300.0, seed=85)   # 11 levels: tiles of 2 and 4 levels ragged
    g, ctx, src64 = _setup(T, cfg, L.ADZ_F32, "f32", nfields=2, seed=cfg.seed)
    i = ctx.info
    box = (i["i0"], i["in"], i["j0"], i["jn"], i["k0t"], cfg.nk)
    nx, ny, nz = box[1] - box[0] + 1, box[3] - box[2] + 1, box[5] - box[4] + 1
    rng = np.random.default_rng(9)
    ii = np.arange(box[0], box[1] + 1, dtype=np.float64)[None, None, :]
    jj = np.arange(box[2], box[3] + 1, dtype=np.float64)[None, :, None]
    kk = np.arange(box[4], box[5] + 1, dtype=np.float64)[:, None, None]
    px = ii - 0.3 + 0.45 * np.sin(ii * 0.7) + 0.0 * jj + 0.0 * kk          # repeats and skips along i
    py = jj + 0.25 + 0.0 * ii + 0.0 * kk
    pz = kk + 0.4 + 0.0 * ii + 0.0 * jj
    px = np.where((np.arange(nx) % 17 == 5)[None, None, :], np.floor(px), px)   # on cell faces
    pz = np.where((np.arange(nx) % 23 == 3)[None, None, :], -3.0, pz)           # below the first level
    pz = np.where((np.arange(nx) % 29 == 4)[None, None, :], cfg.nk + 5.0, pz)   # above the last
    bad = rng.random((nz, ny, nx)) < 0.03
    px = np.where(bad, 1.0e10, px)                                              # stencil out of range
    px = np.where((np.arange(nx) % 41 == 40)[None, None, :], 1 - 40.0, px)      # left of the source
    # stencils on the first columns, rows and planes of the source, and on its last ones
    lo = np.zeros((nz, ny, nx), bool)
    lo[0, 0, np.arange(nx) % 37 == 11] = True
    hi = np.zeros((nz, ny, nx), bool)
    hi[nz - 1, ny - 1, np.arange(nx) % 37 == 13] = True
    first = lambda l0: l0 + 1.3 if l0 + 1 >= 0 else l0 + 0.7   # (int) truncates toward zero
    px = np.where(lo, first(i["lminx"]), np.where(hi, i["lmaxx"] - 1.7, px))
    py = np.where(lo, first(i["lminy"]), np.where(hi, i["lmaxy"] - 1.7, py))
    pz = np.where(lo, 1.2, np.where(hi, cfg.nk - 0.5, pz))
    pos = np.stack([px, py, pz], -1).astype(np.float32).reshape(-1)
    xyz = torch.from_numpy(pos).cuda()
    n_model = (cfg.ni + 2 * MH) * (cfg.nj + 2 * MH) * cfg.nk
    srcs = [torch.from_numpy(s).float().cuda() for s in src64]
    mk = lambda: {k: torch.full((n_model,), -7.0, device="cuda") for k in ("dst", "lin", "min", "max")}
    stack = [dict(src=s, **mk()) for s in srcs]
    ref = [dict(src=s, **mk()) for s in srcs]
    src_off = (1 - i["lminx"]) + i["nit"] * (1 - i["lminy"])
    dl = (1 - MH, cfg.ni + MH, 1 - MH, cfg.nj + MH)
    dst_off = (1 - dl[0]) + (dl[1] - dl[0] + 1) * (1 - dl[2])
    A.adz_tricub_hlt(ref, xyz, ctx.ztab(True), *box[:5], box[5], dl, math=M, post=3, src_off=src_off)
    item = dict(geom=ctx.ztab(True), xyz=xyz, box=box, dlayout=dl, fields=A.field_list(stack, src_off, dst_off))
    plan = A.Plan([item], L.ADZ_F32, math=M, post=3,
                  tune=dict(fields_per_pass=2, points_outer=0, blocks_per_cu=2, xcd_chunks=0, stage=stage))
    plan.execute()
    torch.cuda.synchronize()
    nan_seen = False
    for a, b in zip(stack, ref):
        for k in ("dst", "lin", "min", "max"):
            assert torch.equal(a[k].view(torch.int32), b[k].view(torch.int32)), k
            nan_seen = nan_seen or bool(torch.isnan(a[k]).any())
    assert nan_seen
    plan.free()
    ctx.free()


@pytest.mark.parametrize("entry", ["hlt", "plan", "mono_m_n"])
def test_f32_form_reproduces_constants(T, entry):
    """The FP32 form on a plume over a constant 1e-12 background (the c5 tracers' shape):
    wherever a target's stencil range is a single value (min == max) the value is that
    constant, so no cell of a constant neighbourhood lies outside [min, max]; elsewhere the
    values stay within the FP32 form's error of the FP64 oracle."""
    torch, L, A = T
    from gem_amd import tricublin as TC

    g, ctx, _ = _setup(T, CFG, L.ADZ_F32, "f32", nfields=1)
    i = ctx.info
    box = (i["i0"], i["in"], i["j0"], i["jn"], i["k0t"], CFG.nk)
    shape = (CFG.nk, i["njt"], i["nit"])
    rng = np.random.default_rng(12)
    src64 = S.plume_tracer(shape, rng, "cpu", background=1e-12, nplumes=2).reshape(-1).numpy()
    src = torch.from_numpy(src64.astype(np.float32)).cuda()
    src_off = (1 - i["lminx"]) + i["nit"] * (1 - i["lminy"])
    pt = ctx.array(A.ADZ_A_PT)
    n = pt.numel() // 3
    if entry == "mono_m_n":
        outs = [torch.zeros(n, device="cuda") for _ in range(4)]
        lv = TC.Ztab(g["zt"][1:CFG.nk + 1], i["nit"], i["njt"])
        TC.Tricublin_mono_zyx_m_n(*outs, [src], pt, lv, n, 1, base_off=src_off, math=L.ADZ_MATH_F32)
        torch.cuda.synchronize()
        d, mi, ma = outs[0], outs[2], outs[3]
    else:
        n_model = (CFG.ni + 2 * MH) * (CFG.nj + 2 * MH) * CFG.nk
        e = {k: torch.zeros(n_model, device="cuda") for k in ("dst", "lin", "min", "max")}
        e["src"] = src
        dl = (1 - MH, CFG.ni + MH, 1 - MH, CFG.nj + MH)
        if entry == "hlt":
            A.adz_tricub_hlt([e], pt, ctx.ztab(True), *box[:5], box[5], dl, post=A.ADZ_POST_STORE,
                             math=L.ADZ_MATH_F32, src_off=src_off)
        else:
            dst_off = (1 - dl[0]) + (dl[1] - dl[0] + 1) * (1 - dl[2])
            item = dict(geom=ctx.ztab(True), xyz=pt, box=box, dlayout=dl, fields=A.field_list([e], src_off, dst_off))
            plan = A.Plan([item], L.ADZ_F32, math=L.ADZ_MATH_F32, post=A.ADZ_POST_STORE)
            plan.execute()
            torch.cuda.synchronize()
            plan.free()
        torch.cuda.synchronize()
        sl = lambda a: a.reshape(CFG.nk, CFG.nj + 2 * MH, CFG.ni + 2 * MH)[
            box[4] - 1:, box[2] - 1 + MH:box[3] + MH, box[0] - 1 + MH:box[1] + MH].reshape(-1)
        d, mi, ma = sl(e["dst"]), sl(e["min"]), sl(e["max"])
    const = mi == ma
    assert int(const.sum()) > n // 4  # the background dominates
    assert torch.equal(d[const], mi[const])
    assert int(((d < mi) | (d > ma))[const].sum()) == 0
    zt = O.Ztab(g["zt"][1:CFG.nk + 1], i["nit"], i["njt"], 0, 0)
    exp = O.zyx1_m_n(zt, [src64], src_off, pt.cpu().numpy(), O.ORD_AVX2)[0]
    got = d.double().cpu().numpy()
    assert np.abs(got - exp).max() <= 5e-6 * np.abs(exp).max()
    ctx.free()
