"""GPU: a Fortran host program (gem_amd/fortran/adz_f_driver, built with amdflang) runs one
tracer step through the drop-in routines with the reference's signatures -- adz_set,
adz_inittraj, adz_traject(F_dt_8), adz_tricub_hlt(F_stk, ..., F_post), adz_post_tr(1),
adz_post_tr(0) -> Bermejo_Conde -- and its results are bit-identical to the same step
called through the C ABI from Python: departure positions, interpolated and clipped
tracers, Adz_post lin/min/max, fixed tracers and the carried mass deficits (the fixer with
the LAM boundary flux, Adz_BC_LAM_flux = 1).  Both interpolation forms: tricubic, and
bicubic-horizontal / quintic-vertical (F_QV_L = .true.); and the ZLF boundary treatment
(Adz_BC_LAM_flux = 2: interpolation on Adz_pb, piloting values saved and restored, the
fixer on the Adz_pb box); and ILMC (tracers with mono = 2, ILMC_LAM called by
adz_post_tr(1) before the fixer); and the tricubic step captured by the Fortran program in
a hipGraph on a stream it created and replayed (the replay's output is compared)."""
import os
import subprocess

import numpy as np
import pytest

from gem_amd import synth as S

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "gem_amd", "fortran", "adz_f_driver")
MH = 2


@pytest.mark.parametrize("form", ["tricubic", "quintic", "zlf", "ilmc", "graph"])
def test_fortran_step_matches_c_abi(cuda, tmp_path, form):
    if not os.path.exists(DRIVER):
        pytest.fail("gem_amd/fortran/adz_f_driver is not built")
    import torch

    from gem_amd import _lib as L
    from gem_amd import adz as A

    cfg = S.Config("f", 48, 40, 16, 4, 300.0, jets=True, cmax_xy=3.0, cmax_z=1.0, seed=91)
    g = S.grid_desc(cfg)
    w = S.winds(cfg, g)
    nit, njt, nk = cfg.ni + 2 * cfg.halo, cfg.nj + 2 * cfg.halo, cfg.nk
    n_ext = nit * njt * nk
    mod = (nk, cfg.nj + 2 * MH, cfg.ni + 2 * MH)
    n_mod = int(np.prod(mod))
    rng = np.random.default_rng(5)
    trc = [S.plume_tracer(mod, rng, "cpu").float().numpy().reshape(-1) for _ in range(4)]
    airm1, airm0, pkps, mask = S.airmass(cfg, g, mh=MH)
    mass = np.array([1, 112, 121, 0], np.int32)
    mono = np.array([2, 0, 2, 0] if form == "ilmc" else [1, 0, 1, 0], np.int32)
    names = ["uu_ext", "vv_ext", "ww_ext", "ut0", "vt0", "zdt0"]
    for k in names:
        assert w[k].numel() == n_ext and w[k].dtype == torch.float32, k

    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(fin, "wb") as f:
        np.array([cfg.ni, cfg.nj, nk, cfg.halo, cfg.halo, g["pil_w"], cfg.niter, 4, MH, L.ADZ_F32],
                 np.int32).tofile(f)
        np.array([g["inv_hx"], g["inv_hy"], g["inv_rayt"], g["zmin"], g["zmax"], g["wmstar_nk"],
                  g["idz_t_nkm1"], cfg.dt], np.float64).tofile(f)
        for a in (g["zm"], g["zt"], g["zx"], g["lat"]):
            np.asarray(a, np.float64).tofile(f)
        for k in names:
            w[k].numpy().tofile(f)
        for t in trc:
            t.tofile(f)
        for a in (airm1, airm0, pkps):
            a.numpy().astype(np.float32).reshape(-1).tofile(f)
        mask.numpy().astype(np.float64).reshape(-1).tofile(f)
        mass.tofile(f)
        mono.tofile(f)
    print("[gpu-case] running the Fortran driver", flush=True)
    mode = {"tricubic": [], "quintic": ["qv"], "zlf": ["zlf"], "ilmc": [], "graph": ["graph"]}[form]
    r = subprocess.run([DRIVER, str(fin), str(fout)] + mode, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr

    # the same step through the C ABI
    if form == "zlf":
        g["bc_lam_flux"] = 2
    ctx = A.AdzContext(g, L.ADZ_F32, itraj=cfg.niter)
    wd = {k: v.cuda() for k, v in w.items()}
    ctx.adz_inittraj()
    ctx.adz_traject(cfg.dt, *[wd[k] for k in names])
    i = ctx.info
    box = (i["i0"], i["in"], i["j0"], i["jn"], i["k0t"], nk)
    dl = (1 - MH, cfg.ni + MH, 1 - MH, cfg.nj + MH)
    off = MH + (cfg.ni + 2 * MH) * MH
    ent = [dict(src=torch.from_numpy(t).cuda(), dst=torch.zeros(n_mod, device="cuda"),
                lin=torch.zeros(n_mod, device="cuda"), min=torch.zeros(n_mod, device="cuda"),
                max=torch.zeros(n_mod, device="cuda")) for t in trc]
    lay5 = dl + (nk,)
    zlf = dict(layout=lay5, tile=(cfg.ni, cfg.nj), pil=(g["pil_w"], g["pil_e"], g["pil_s"], g["pil_n"]),
               ext=cfg.halo, k0=i["k0t"], off=off)
    pil_c = [torch.zeros(n_mod, device="cuda") for _ in ent]
    if form == "zlf":  # adz_BC_LAM_zlf_0_hlt(n, 0)
        A.adz_bc_lam_zlf(1, [e["src"] for e in ent], None, **zlf)
        A.adz_bc_lam_zlf(2, [e["dst"] for e in ent], pil_c, **zlf)
    for clip in (0, 1):
        grp = [e for e, m in zip(ent, mono) if (m == 1) == bool(clip)]
        post = A.ADZ_POST_STORE | (A.ADZ_POST_CLIP if clip else 0)
        if not grp:
            continue
        if form == "zlf":
            A.adz_tricub_hlt(grp, ctx.array(A.ADZ_A_PB), ctx.ztab(True), i["i0b"], i["inb"], i["j0b"], i["jnb"], 1,
                             nk, dl, ext=False, post=post, math=L.ADZ_MATH_REF, src_off=off)
        elif form == "quintic":
            A.adz_tricub_hlt_qv(ctx, grp, ctx.array(A.ADZ_A_PT), *box[:5], dl, ext=False, post=post, src_off=off)
        else:
            A.adz_tricub_hlt(grp, ctx.array(A.ADZ_A_PT), ctx.ztab(True), *box[:5], box[5], dl, ext=False,
                             post=post, math=L.ADZ_MATH_REF, src_off=off)
    d_air0, d_mask = airm0.reshape(-1).cuda(), mask.reshape(-1).double().cuda()
    if form == "ilmc":  # adz_post_tr(1): ILMC_LAM on Adz_i0:Adz_in x Adz_j0:Adz_jn x Adz_k0t:nk
        for n in np.nonzero(mono == 2)[0]:
            e = ent[n]
            assert int(((e["dst"] < e["min"]) | (e["dst"] > e["max"])).sum()) > 0, n  # ILMC has work to do
            il = A.IlmcLam([dict(f=ent[n]["dst"], min=ent[n]["min"], max=ent[n]["max"])], dl + (nk,), box[:5], off)
            il(d_air0)
            torch.cuda.synchronize()
            il.free()
    bcn = [n for n in range(4) if mass[n] == 1 or 111 <= mass[n] <= 139]
    bcs = [dict(p=ent[n]["src"], m=ent[n]["dst"], lin=ent[n]["lin"], min=ent[n]["min"], max=ent[n]["max"],
                mass=int(mass[n])) for n in bcn]
    scope = (1 + g["pil_w"], cfg.ni - g["pil_e"], 1 + g["pil_s"], cfg.nj - g["pil_n"], i["k0t"])
    if form == "zlf":  # adz_BC_LAM_zlf_0_hlt(n, 1); the fixer on the Adz_pb box from level 1
        A.adz_bc_lam_zlf(3, [e["dst"] for e in ent], None, **zlf)
        scope = (i["i0b"], i["inb"], i["j0b"], i["jnb"], 1)
    bc = A.BermejoConde(bcs, dl + (nk,), scope, off, min_max=True)
    deficit = torch.zeros(len(bcn), dtype=torch.float64, device="cuda")
    fx = None
    if form == "zlf":
        bc(airm1.reshape(-1).cuda(), d_air0, d_mask, deficit, pkps=pkps.reshape(-1).cuda())
        A.adz_bc_lam_zlf(4, [e["dst"] for e in ent], pil_c, **zlf)  # adz_BC_LAM_zlf_0_hlt(n, 2)
    else:
        # the LAM boundary flux (Adz_BC_LAM_flux = 1, the driver's grid) of the fixer's tracers
        fx = A.BcLamFlux([dict(src=ent[n]["src"]) for n in bcn], dl + (nk,), (cfg.ni, cfg.nj),
                         (i["i0b"], i["inb"], i["j0b"], i["jnb"]), scope, off, ctx.ztab(True), pos_dtype=L.ADZ_F32)
        fmass = torch.zeros(2 * len(bcn), dtype=torch.float64, device="cuda")
        fx(ctx.array(A.ADZ_A_PB), d_air0, d_mask, fmass)
        bc(airm1.reshape(-1).cuda(), d_air0, d_mask, deficit, pkps=pkps.reshape(-1).cuda(), flux=fmass)
    torch.cuda.synchronize()

    npt = 3 * (box[1] - box[0] + 1) * (box[3] - box[2] + 1) * (nk - box[4] + 1)
    raw = open(fout, "rb").read()
    o = 0

    def take(n, dt):
        nonlocal o
        a = np.frombuffer(raw, dt, n, o)
        o += a.nbytes
        return a

    pos = take(npt, np.float32)
    assert np.array_equal(pos.view(np.uint32), ctx.array(A.ADZ_A_PT).cpu().numpy().view(np.uint32))
    for n in range(4):
        assert np.array_equal(take(n_mod, np.float32).view(np.uint32), ent[n]["dst"].cpu().numpy().view(np.uint32)), n
    for n in range(4):
        for k in ("lin", "min", "max"):
            assert np.array_equal(take(n_mod, np.float32).view(np.uint32),
                                  ent[n][k].cpu().numpy().view(np.uint32)), (n, k)
    fdef = take(4, np.float64)
    assert o == len(raw)
    exp = np.zeros(4)
    exp[bcn] = deficit.cpu().numpy()
    assert np.array_equal(fdef.view(np.uint64), exp.view(np.uint64))
    assert np.any(exp != 0.0)
    if fx is not None:
        assert np.all(fmass.cpu().numpy() != 0.0)
        fx.free()
    bc.free()
    ctx.free()
