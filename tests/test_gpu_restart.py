"""GPU: restart of the semi-Lagrangian state (adz_ctx_save_state / adz_ctx_load_state,
the arrays adz_set.F90:343-360 registers for restart, with BC_mass_deficit as
wrrstrt.F90:49-63 writes it).

Six steps of trajectory + tracer interpolation (lin/min/max) + LAM boundary flux + fixer
run without interruption, and again as three steps, a save, a fresh context that loads
the state and three more steps.  Positions, tracers and carried deficits agree bit for
bit (the reference's restart test compares whole model states the same way).  A state
loaded into a context of another position storage is refused.
"""
import numpy as np
import pytest

from gem_amd import synth as S

pytestmark = pytest.mark.gpu
MH = 2


def _run(torch, L, A, cfg, g, w, tracers, pos_dtype, nsteps, ctx=None, deficit=None, start=0):
    """Steps start .. start+nsteps-1: even steps read buffer 0 and air mass 0 at time P, odd
    steps the reverse; start = 0 begins with adz_inittraj.  tracers: one array per tracer
    (start = 0) or the pair of buffers a run returned (the model's tracer state)."""
    if ctx is None:
        ctx = A.AdzContext(g, pos_dtype, itraj=cfg.niter)
    if start == 0:
        ctx.adz_inittraj()
    i = ctx.info
    mod = (cfg.nk, cfg.nj + 2 * MH, cfg.ni + 2 * MH)
    n_mod = int(np.prod(mod))
    ntr = len(tracers)
    if start == 0:
        bufs = [[t.clone(), torch.zeros_like(t)] for t in tracers]
    else:
        bufs = [[b[0].clone(), b[1].clone()] for b in tracers]
    post = [{k: torch.zeros(n_mod, device="cuda") for k in ("lin", "min", "max")} for _ in range(ntr)]
    airm1, airm0, pkps, mask = S.airmass(cfg, g, device="cuda", mh=MH)
    air = [airm1.reshape(-1).contiguous(), airm0.reshape(-1).contiguous()]
    mask = mask.reshape(-1).contiguous()
    box = (i["i0"], i["in"], i["j0"], i["jn"], i["k0t"], cfg.nk)
    dl = (1 - MH, cfg.ni + MH, 1 - MH, cfg.nj + MH)
    off = MH + (cfg.ni + 2 * MH) * MH
    scope = (1 + cfg.pil, cfg.ni - cfg.pil, 1 + cfg.pil, cfg.nj - cfg.pil, i["k0t"])
    stacks, bcs, fxs = [], [], []
    for par in range(2):
        st = [dict(src=bufs[t][par], dst=bufs[t][1 - par], lin=post[t]["lin"], min=post[t]["min"],
                   max=post[t]["max"]) for t in range(ntr)]
        stacks.append(st)
        bcs.append(A.BermejoConde([dict(p=e["src"], m=e["dst"], lin=e["lin"], min=e["min"], max=e["max"], mass=1)
                                   for e in st], dl + (cfg.nk,), scope, off, min_max=True))
        fxs.append(A.BcLamFlux([dict(src=e["src"]) for e in st], dl + (cfg.nk,), (cfg.ni, cfg.nj),
                               (i["i0b"], i["inb"], i["j0b"], i["jnb"]), scope, off, ctx.ztab(True),
                               pos_dtype=pos_dtype))
    if deficit is None:
        deficit = torch.zeros(ntr, dtype=torch.float64, device="cuda")
    fmass = torch.zeros(2 * ntr, dtype=torch.float64, device="cuda")
    for n in range(start, start + nsteps):
        par = n % 2
        ctx.adz_traject(cfg.dt, w["uu_ext"], w["vv_ext"], w["ww_ext"], w["ut0"], w["vt0"], w["zdt0"])
        A.adz_tricub_hlt(stacks[par], ctx.array(A.ADZ_A_PT), ctx.ztab(True), *box[:5], box[5], dl, ext=False,
                         post=A.ADZ_POST_STORE, math=L.ADZ_MATH_F32, src_off=off)
        fxs[par](ctx.array(A.ADZ_A_PB), air[1 - par], mask, fmass)
        bcs[par](air[par], air[1 - par], mask, deficit, pkps=pkps.reshape(-1).contiguous(), flux=fmass)
    torch.cuda.synchronize()
    out = [[b[0].clone(), b[1].clone()] for b in bufs]
    for b in bcs + fxs:
        b.free()
    return ctx, out, deficit


@pytest.mark.parametrize("pos", ["f32", "f64"])
def test_restart_is_bit_identical(cuda, pos):
    import torch

    from gem_amd import _lib as L
    from gem_amd import adz as A

    cfg = S.Config("r", 96, 72, 30, 4, 300.0, jets=True, cmax_xy=3.0, cmax_z=1.0, seed=77)
    g = S.grid_desc(cfg)
    pdt = L.ADZ_F32 if pos == "f32" else L.ADZ_F64
    w = {k: v.cuda() for k, v in S.winds(cfg, g).items()}
    rng = np.random.default_rng(cfg.seed)
    mod = (cfg.nk, cfg.nj + 2 * MH, cfg.ni + 2 * MH)
    tracers = [S.plume_tracer(mod, rng, "cuda").float().reshape(-1).contiguous() for _ in range(2)]

    ctx_a, out_a, def_a = _run(torch, L, A, cfg, g, w, tracers, pdt, 6)
    ctx_b, mid, def_b = _run(torch, L, A, cfg, g, w, tracers, pdt, 3)
    state = ctx_b.save_state()
    saved_def = def_b.cpu().numpy().copy()       # BC_mass_deficit, as wrrstrt writes it
    ctx_b.free()
    ctx_c = A.AdzContext(g, pdt, itraj=cfg.niter)  # adz_set of the restarted run
    ctx_c.load_state(state)
    ctx_c, out_c, def_c = _run(torch, L, A, cfg, g, w, mid, pdt, 3, ctx=ctx_c,
                               deficit=torch.from_numpy(saved_def).cuda(), start=3)
    print(f"[restart] {pos}: state of {state.size} bytes", flush=True)
    for which in (A.ADZ_A_PT, A.ADZ_A_PB, A.ADZ_A_PXYZM, A.ADZ_A_WPXYZ):
        assert torch.equal(ctx_a.array(which), ctx_c.array(which)), which
    # the tracers where the step defines them (the interpolation's target box)
    i = ctx_a.info
    sl = (slice(i["k0t"] - 1, None), slice(i["j0"] - 1 + MH, i["jn"] + MH), slice(i["i0"] - 1 + MH, i["in"] + MH))
    for a, c in zip(out_a, out_c):
        a3, c3 = a[0].cpu().numpy().reshape(mod)[sl], c[0].cpu().numpy().reshape(mod)[sl]  # step 6's result
        assert np.array_equal(a3.view(np.uint32), c3.view(np.uint32)), int(np.count_nonzero(a3 != c3))
    assert np.array_equal(def_a.cpu().numpy().view(np.uint64), def_c.cpu().numpy().view(np.uint64))
    assert np.any(def_a.cpu().numpy() != 0.0)
    # a state of the other position storage is refused
    other = A.AdzContext(g, L.ADZ_F64 if pdt == L.ADZ_F32 else L.ADZ_F32, itraj=cfg.niter)
    with pytest.raises(L.AdzError):
        other.load_state(state)
    bad = state.copy()
    bad[:8] = 0
    with pytest.raises(L.AdzError):
        ctx_c.load_state(bad)
    other.free()
    ctx_a.free()
    ctx_c.free()
