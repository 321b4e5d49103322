"""GPU: bicubic-horizontal / quintic-vertical interpolation (adz_tricub_hlt_qv, the
F_QV_L path of adz_interp_hlt.F90) against the oracle (bicubHQV_lin_mono.F90:20-172,
bicubHQV_lin_min_max_zyx.F90:20-99), bit for bit.

Setup: the c1 grid (200x150x60) after one trajectory, the Adz_pt positions in FP32 (the
reference's storage) and FP64; three tracers of lognormal values; sources on the model
layout (halo 2) and on the extended layout.  Post-processing: none, lin/min/max stored,
and stored with the mono = 1 clip.  Every target of the box takes one of the three
vertical modes (quintic, cubic, linear); the test checks that each occurs.
"""
import numpy as np
import pytest

from gem_amd import synth as S
from oracle import oracle as O

pytestmark = pytest.mark.gpu
MH = 2


@pytest.fixture(scope="module", params=["pos_f32", "pos_f64"])
def W(request, cuda):
    import torch

    from gem_amd import _lib as L
    from gem_amd import adz as A

    cfg = S.CONFIGS["c1"]
    g = S.grid_desc(cfg)
    pdt = L.ADZ_F32 if request.param == "pos_f32" else L.ADZ_F64
    ctx = A.AdzContext(g, pdt, itraj=cfg.niter)
    w = {k: v.cuda() for k, v in S.winds(cfg, g).items()}
    ctx.adz_inittraj()
    ctx.adz_traject(cfg.dt, w["uu_ext"], w["vv_ext"], w["ww_ext"], w["ut0"], w["vt0"], w["zdt0"])
    yield dict(torch=torch, L=L, A=A, cfg=cfg, g=g, ctx=ctx, tab=O.QvTab(g["zt"], cfg.nk))
    ctx.free()


@pytest.mark.parametrize("post", [0, 1, 3])
@pytest.mark.parametrize("ext", [False, True])
def test_qv_bit_exact(W, ext, post):
    torch, A, cfg, ctx = W["torch"], W["A"], W["cfg"], W["ctx"]
    i = ctx.info
    nk = cfg.nk
    dl = (1 - MH, cfg.ni + MH, 1 - MH, cfg.nj + MH)
    sl = (i["lminx"], i["lminx"] + i["nit"] - 1, i["lminy"], i["lminy"] + i["njt"] - 1) if ext else dl
    n_src = (sl[1] - sl[0] + 1) * (sl[3] - sl[2] + 1) * nk
    n_dst = (dl[1] - dl[0] + 1) * (dl[3] - dl[2] + 1) * nk
    rng = np.random.default_rng(17 + post + 4 * ext)
    src = [(1e-6 * rng.lognormal(0.0, 1.0, n_src)).astype(np.float32) for _ in range(3)]
    stack = [dict(src=torch.from_numpy(s).cuda(), dst=torch.zeros(n_dst, device="cuda"),
                  lin=torch.zeros(n_dst, device="cuda"), min=torch.zeros(n_dst, device="cuda"),
                  max=torch.zeros(n_dst, device="cuda")) for s in src]
    box = (i["i0"], i["in"], i["j0"], i["jn"], i["k0t"])
    pos = ctx.array(A.ADZ_A_PT)
    A.adz_tricub_hlt_qv(ctx, stack, pos, *box, dl, ext=ext, post=post)
    torch.cuda.synchronize()
    p = pos.cpu().numpy()
    qut, lin, mi, ma = O.qv_m_n(W["tab"], src, sl, p, -1, -1)  # Adz_ioff = l_i0 - 2 = -1
    pz = p.reshape(-1, 3)[:, 2]
    kk1 = pz.astype(np.int64)
    modes = {"quintic": int(np.count_nonzero((kk1 >= 3) & (kk1 <= nk - 3))),
             "cubic": int(np.count_nonzero((kk1 == 2) | (kk1 == nk - 2)))}
    modes["linear"] = pz.size - modes["quintic"] - modes["cubic"]
    print(f"[qv] ext={ext} post={post}: targets per vertical mode {modes}", flush=True)
    assert all(v > 0 for v in modes.values())
    shape = (nk, dl[3] - dl[2] + 1, dl[1] - dl[0] + 1)
    bsl = (slice(box[4] - 1, None), slice(box[2] - dl[2], box[3] - dl[2] + 1), slice(box[0] - dl[0], box[1] - dl[0] + 1))

    def got(t, key):
        return t[key].cpu().numpy().reshape(shape)[bsl].reshape(-1)

    for f in range(len(src)):
        exp = qut[f]
        if post & A.ADZ_POST_CLIP:
            exp = np.maximum(mi[f].astype(np.float64), np.minimum(ma[f].astype(np.float64), exp)).astype(np.float32)
        assert np.array_equal(got(stack[f], "dst").view(np.uint32), exp.view(np.uint32)), f
        if post & A.ADZ_POST_STORE:
            el = np.maximum(mi[f].astype(np.float64), np.minimum(ma[f].astype(np.float64), lin[f])).astype(np.float32)
            assert np.array_equal(got(stack[f], "lin").view(np.uint32), el.view(np.uint32)), f
            assert np.array_equal(got(stack[f], "min").view(np.uint32), mi[f].view(np.uint32)), f
            assert np.array_equal(got(stack[f], "max").view(np.uint32), ma[f].view(np.uint32)), f


def test_qv_positions_outside_the_source_give_nan(W):
    """A stencil that would leave the source arrays yields NaN (the reference reads out of
    bounds there); the other targets are unaffected."""
    torch, A, cfg, ctx = W["torch"], W["A"], W["cfg"], W["ctx"]
    i = ctx.info
    dl = (1 - MH, cfg.ni + MH, 1 - MH, cfg.nj + MH)
    n_dst = (dl[1] - dl[0] + 1) * (dl[3] - dl[2] + 1) * cfg.nk
    src = torch.ones(n_dst, device="cuda")
    box = (i["i0"], i["i0"] + 3, i["j0"], i["j0"], cfg.nk)  # 4 targets on the top level
    pos = torch.tensor([[3.5, 5.5, 20.5], [-50.0, 5.5, 20.5], [3.5, 1e12, 20.5], [float("nan"), 5.5, 20.5]],
                       dtype=ctx.array(A.ADZ_A_PT).dtype, device="cuda").reshape(-1)
    stack = [dict(src=src, dst=torch.zeros(n_dst, device="cuda"))]
    A.adz_tricub_hlt_qv(ctx, stack, pos, *box, dl, ext=False, post=0)
    torch.cuda.synchronize()
    shape = (cfg.nk, dl[3] - dl[2] + 1, dl[1] - dl[0] + 1)
    got = stack[0]["dst"].cpu().numpy().reshape(shape)[cfg.nk - 1, box[2] - dl[2], box[0] - dl[0]:box[1] - dl[0] + 1]
    assert got[0] == 1.0
    assert np.all(np.isnan(got[1:]))
