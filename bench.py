#!/usr/bin/env python3
"""Benchmark of the semi-Lagrangian advection step on one MI355X.

Workload (default `c5`): 2048x1024x84 single-tile LAM grid, halo 5, dt = 450 s, winds
with Courant numbers up to 6 (x, y) and 2 (z); 64 interpolated fields: u at the u-grid
positions, v at the v-grid positions, T, q and 60 tracers at the thermodynamic positions;
the Bermejo-Conde fixer on the 60 tracers.  One step = trajectory (adz_traject: FP64
iteration arithmetic and FP64 position history; positions and derived winds stored as the
reference's real arrays by default, --traj-storage f64 for FP64 throughout) + batched FP32
interpolation (persistent plans) + Bermejo-Conde, captured in a hipGraph and replayed.
Inputs are synthetic and already resident in HBM.

Printed JSON line: value = interpolated Gpoints/s of the whole step (points x fields /
step time).  Beside it: the interpolation-kernel-only rate and HBM GB/s, the plain FP64
path's rate at the same positions (the FP64 drop-in), the per-field max / RMS relative
error of the FP32 path against the FP64 path (bit-exact with the FP64 oracle at test
sizes), the roofline of the dominant interpolation kernel, and the reference CPU path
timed on this host on a bounded sample.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

MH = 2          # model halo of destination / tracer arrays
HBM_PEAK = 8000.0  # GB/s, MI355X published HBM3E bandwidth


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=5)
    p.add_argument("--warmup", type=int, default=2)
    p.add_argument("--config", default="c5", help="c5 (default) or c4 (16 uneven subdomains, 40 tracers)")
    p.add_argument("--graph-steps", type=int, default=20, help="c4: timesteps captured in one hipGraph")
    p.add_argument("--bc-group", type=int, default=20, help="tracers per Bermejo-Conde group")
    p.add_argument("--fp64-fields", type=int, default=16, help="fields in the plain FP64 timing leg")
    p.add_argument("--err-fields", type=int, default=0, help="fields checked against FP64 (0 = all)")
    p.add_argument("--no-lam-flux", action="store_true",
                   help="c5: leave out the LAM boundary flux of the fixer (Adz_BC_LAM_flux = 0)")
    p.add_argument("--no-ilmc", action="store_true", help="skip the ILMC (mono = 2) timing beside the step")
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--cpu-seconds", type=float, default=12.0)
    p.add_argument("--profile-step", action="store_true",
                   help="run one step (each of its kernels once), then a 1 GiB copy for calibrating PMC byte "
                        "counters, and exit (for rocprofv3 --pmc)")
    p.add_argument("--profile-kernel", action="store_true",
                   help="only build the workload and run the dominant interpolation kernel 5 times (rocprofv3)")
    p.add_argument("--traj-storage", choices=("f32", "f64"), default="f32",
                   help="storage of the trajectory state: f32 = the reference's real arrays (FP64 iteration "
                        "arithmetic and FP64 position history), f64 = FP64 throughout (timed beside it either way)")
    p.add_argument("--tiles", default=os.path.join(ROOT, "gem_amd", "tuned_tiles.json"),
                   help="JSON file of tile parameters per workload (the autotuner's committed result)")
    p.add_argument("--fortran", action="store_true",
                   help="c5 through the Fortran drop-in layer instead (module adz_gpu: adz_main + adz_tracers "
                        "before and after the pressure adjustment, one hipGraph), timed alone")
    p.add_argument("--pmc-summary", nargs=2, metavar=("FETCH_DIR", "WRITE_DIR"),
                   help="write profiles/k_plan_<config>_pmc.json from two rocprofv3 --pmc runs of "
                        "--profile-kernel (FETCH_SIZE, then WRITE_SIZE; their csv output folders) and exit")
    p.add_argument("--kernel-trace", metavar="KT_DIR",
                   help="with --pmc-summary: the rocprofv3 --kernel-trace --stats folder of the same command, "
                        "for the average launch duration")
    p.add_argument("--retune", action="store_true",
                   help="run adz_plan_autotune instead of using the tiles it found before (gem_amd/tuned_tiles.json)")
    return p.parse_args()


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU: the path shards by horizontal domain "
                                   "decomposition (one tile per GPU), which is not part of this build; "
                                   "run with --gpus 1"}))
        sys.exit(2)
    import numpy as np
    import torch

    from gem_amd import _lib as L
    from gem_amd import adz as A
    from gem_amd import synth as S

    if args.pmc_summary:
        return pmc_summary(args)
    assert torch.cuda.is_available(), "bench.py needs a HIP device"
    dev = torch.device("cuda:0")
    if args.config == "c4":
        return run_c4(args, np, torch, L, A, S, dev)
    if args.fortran:
        return run_fortran(args, np, torch, L, A, S, dev)
    cfg = S.CONFIGS[args.config]
    g = S.grid_desc(cfg)
    nf = cfg.nfields
    ntr = cfg.ntracers
    t_setup = time.time()

    # ------------------------------------------------------------------ state
    pos_dt = L.ADZ_F32 if args.traj_storage == "f32" else L.ADZ_F64
    ctx = A.AdzContext(g, pos_dt, itraj=cfg.niter)
    info = ctx.info
    w = {k: v.to(dev) for k, v in S.winds(cfg, g, device=dev).items()}
    ext_shape = (cfg.nk, cfg.nj + 2 * cfg.halo, cfg.ni + 2 * cfg.halo)
    mod_shape = (cfg.nk, cfg.nj + 2 * MH, cfg.ni + 2 * MH)
    n_ext = math.prod(ext_shape)
    n_mod = math.prod(mod_shape)
    ext_off = (1 - info["lminx"]) + info["nit"] * (1 - info["lminy"])
    mod_layout = (1 - MH, cfg.ni + MH, 1 - MH, cfg.nj + MH)
    mod_off = MH + (cfg.ni + 2 * MH) * MH
    nb = min(args.bc_group, ntr) if ntr else 0
    est = (3 * n_ext + (nf - 3) * n_mod + nf * n_mod + 3 * nb * n_mod) * 4 / 1e9
    log(f"config {cfg.name}: {cfg.ni}x{cfg.nj}x{cfg.nk}, {nf} fields, est. field memory {est:.1f} GB")
    if est > 230:
        raise SystemExit("workload would not leave headroom in HBM")

    # sources: u, v, T on the extended layout (orhs*_ext), q and tracers on the model layout
    def gen_field(q, dtype):
        if q < 3:
            f = S.fields(cfg, 1, device=dev, kind="smooth", seed=cfg.seed + 31 * q)[0]
            f = f.reshape(ext_shape)
            scale, off = [(20.0, 0.0), (20.0, 0.0), (15.0, 250.0)][q]
            return (off + scale * f / 8.0).reshape(-1).to(dtype)
        shape = mod_shape
        rng_seed = cfg.seed + 31 * q
        if q == 3:
            f = S.smooth_field(shape, np.random.default_rng(rng_seed), dev)
            return (0.006 + 0.002 * f / 8.0).reshape(-1).to(dtype)
        return S.plume_tracer(shape, np.random.default_rng(rng_seed), dev).reshape(-1).to(dtype)

    src = [gen_field(q, torch.float32) for q in range(nf)]
    dst = [torch.zeros(n_mod, dtype=torch.float32, device=dev) for _ in range(nf)]
    post = [{k: torch.zeros(n_mod, dtype=torch.float32, device=dev) for k in ("lin", "min", "max")}
            for _ in range(nb)]
    # pressures at times P and M (nk+1 momentum levels), pr_t and p0 at M: set_post_tr_hlt forms
    # the air masses and pr_t/p0 from them in every step, as after the pressure adjustment
    _, _, _, mask = S.airmass(cfg, g, device=dev, mh=MH)
    mask = mask.reshape(-1).contiguous()
    prng = np.random.default_rng(cfg.seed + 7)
    shape2 = mod_shape[1:]
    ps = torch.from_numpy(prng.uniform(0.97, 1.03, shape2)).to(dev)
    pres_m = [(torch.exp(torch.tensor(g["zm"][1:cfg.nk + 2], device=dev))[:, None, None] * ps[None] *
               (1.0 + 1e-3 * q)).float().reshape(-1) for q in range(2)]
    pres_t = (torch.exp(torch.tensor(g["zt"][1:cfg.nk + 1], device=dev))[:, None, None] * ps[None]).float().reshape(-1)
    pres_0 = (1e5 * ps).float().reshape(-1)
    cell_area = ((S.RAYT * g["h"]) ** 2 * torch.ones(shape2, dtype=torch.float64, device=dev)).reshape(-1)
    del ps
    airm1, airm0, pkps = (torch.zeros(n_mod, device=dev) for _ in range(3))

    def air_masses():  # set_post_tr_hlt.F90:53-70
        A.adz_set_post_tr(ctx, mod_layout, pres_m[0], pres_m[1], pres_t, pres_0, cell_area, airm1, airm0, pkps)
    deficit = torch.zeros(max(ntr, 1), dtype=torch.float64, device=dev)
    fmass = torch.zeros(2 * max(ntr, 1), dtype=torch.float64, device=dev)  # (mass_fo, mass_fi) per tracer
    lam_flux = not args.no_lam_flux and g["bc_lam_flux"] == 1

    ctx.adz_inittraj()
    step_traj = lambda: ctx.adz_traject(cfg.dt, w["uu_ext"], w["vv_ext"], w["ww_ext"], w["ut0"], w["vt0"],
                                        w["zdt0"])
    step_traj()
    torch.cuda.synchronize()
    pmu, pmv, pt = ctx.array(A.ADZ_A_PMU), ctx.array(A.ADZ_A_PMV), ctx.array(A.ADZ_A_PT)
    pb = ctx.array(A.ADZ_A_PB)
    npts_b = (info["inb"] - info["i0b"] + 1) * (info["jnb"] - info["j0b"] + 1) * cfg.nk
    boxes = {
        "u": ((info["i0u"], info["inu"], info["j0"], info["jn"], info["k0"], cfg.nk), pmu, ctx.ztab(False)),
        "v": ((info["i0"], info["in"], info["j0v"], info["jnv"], info["k0"], cfg.nk), pmv, ctx.ztab(False)),
        "t": ((info["i0"], info["in"], info["j0"], info["jn"], info["k0t"], cfg.nk), pt, ctx.ztab(True)),
    }

    def npts(box):
        i0, in_, j0, jn, k0, nk = box
        return (in_ - i0 + 1) * (jn - j0 + 1) * (nk - k0 + 1)

    def item(name, fields, ext):
        box, xyz, zt = boxes[name]
        so = ext_off if ext else mod_off
        return dict(geom=zt, xyz=xyz, box=box, dlayout=mod_layout, ext=ext,
                    fields=A.field_list(fields, so, mod_off))

    # plan A: u, v, T (extended sources) and q (model layout), no post-processing
    itemsA = [item("u", [dict(src=src[0], dst=dst[0])], True), item("v", [dict(src=src[1], dst=dst[1])], True),
              item("t", [dict(src=src[2], dst=dst[2])], True),
              item("t", [dict(src=src[3], dst=dst[3])], False)]
    t_tune = time.time()
    cached = None if args.retune else tuned_tiles(args.config, args.tiles)  # the autotuner's earlier result
    tuneA = cached["uvtq"] if cached else A.Plan.autotune(itemsA, L.ADZ_F32, math=L.ADZ_MATH_F32, post=0, reps=2)[0]
    planA = A.Plan(itemsA, L.ADZ_F32, math=L.ADZ_MATH_F32, post=0, tune=tuneA)
    # plans B: tracer groups with lin/min/max for the fixer, and the fixer itself
    groups = []
    tbox = boxes["t"][0]
    scope = (1 + cfg.pil, cfg.ni - cfg.pil, 1 + cfg.pil, cfg.nj - cfg.pil, info["k0t"])  # core (adz_post_tr.F90:141)
    layout5 = mod_layout + (cfg.nk,)
    for g0 in range(4, nf, nb if nb else nf):
        ids = list(range(g0, min(nf, g0 + nb)))
        stack = [dict(src=src[q], dst=dst[q], lin=post[k]["lin"], min=post[k]["min"], max=post[k]["max"])
                 for k, q in enumerate(ids)]
        itemsB = [item("t", stack, False)]
        if not groups:  # the groups share one shape: tune the first, reuse its tiles
            tuneB = cached["tracer_groups"] if cached else A.Plan.autotune(
                itemsB, L.ADZ_F32, math=L.ADZ_MATH_F32, post=A.ADZ_POST_STORE, reps=2)[0]
        plan = A.Plan(itemsB, L.ADZ_F32, math=L.ADZ_MATH_F32, post=A.ADZ_POST_STORE, tune=tuneB)
        bcs = [dict(p=src[q], m=dst[q], lin=post[k]["lin"], min=post[k]["min"], max=post[k]["max"], mass=1)
               for k, q in enumerate(ids)]
        bc = A.BermejoConde(bcs, layout5, scope, mod_off, min_max=True)
        groups.append((ids, plan, bc))
    # Aranami flux of every fixer tracer at Adz_pb (adz_interp_hlt.F90:91-96), one call for all
    # tracers: the targets whose stencil crosses the core's edge are found once per step
    fx = A.BcLamFlux([dict(src=src[q]) for q in range(4, nf)], layout5, (cfg.ni, cfg.nj),
                     (info["i0b"], info["inb"], info["j0b"], info["jnb"]), scope, mod_off, ctx.ztab(True),
                     pos_dtype=pos_dt) if lam_flux and ntr else None
    tuneB = tuneB if groups else None
    log(f"tiles ({'tuned now' if not cached else 'from gem_amd/tuned_tiles.json'}, {time.time() - t_tune:.1f}s): "
        f"u/v/T/q plan {tuneA}, tracer groups {tuneB}")
    pts_total = npts(boxes["u"][0]) + npts(boxes["v"][0]) + (nf - 2) * npts(tbox)
    log(f"setup {time.time() - t_setup:.1f}s; {pts_total / 1e9:.3f} G interpolated point-fields per step")

    def interp():
        planA.execute()
        for ids, plan, bc in groups:
            plan.execute()

    def fixer(ids, bc):
        sl = slice(ids[0] - 4, ids[-1] - 3)
        fm = fmass[2 * (ids[0] - 4):2 * (ids[-1] - 3)] if fx else None
        bc(airm1, airm0, mask, deficit[sl], pkps=pkps, flux=fm)

    def step():  # dynstep.F90:94-98: tracers interpolated, pressure adjusted, air masses, fixed
        step_traj()
        planA.execute()
        for ids, plan, bc in groups:
            plan.execute()
        air_masses()
        if fx:
            fx(pb, airm0, mask, fmass)
        for ids, plan, bc in groups:
            fixer(ids, bc)

    ev = lambda: torch.cuda.Event(enable_timing=True)

    def timed(fn, reps):
        torch.cuda.synchronize()
        a, b = ev(), ev()
        a.record()
        for _ in range(reps):
            fn()
        b.record()
        torch.cuda.synchronize()
        return a.elapsed_time(b) / reps

    if args.profile_step:
        step()  # warm (first-launch effects out of the counted step)
        torch.cuda.synchronize()
        step()
        torch.cuda.synchronize()
        x = torch.empty(2 ** 28, dtype=torch.float32, device=dev)
        y = torch.empty_like(x)
        y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"profile_step": cfg.name, "tiles": {"uvtq": tuneA, "tracer_groups": tuneB},
                          "copy_bytes": 2 * 4 * 2 ** 28}), flush=True)
        return

    if args.profile_kernel:
        ids, plan, bc = groups[0]
        for _ in range(5):
            plan.execute()
        torch.cuda.synchronize()
        # calibration copy of known size for the PMC byte counters (1 GiB read + 1 GiB write)
        x = torch.empty(2 ** 28, dtype=torch.float32, device=dev)
        y = torch.empty_like(x)
        y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"profile_kernel": "k_plan (tracer group)", "tiles": tuneB, "copy_bytes": 2 * 4 * 2 ** 28,
                          "kernel_source_sha16": kernel_source_sha()}), flush=True)
        return

    # ------------------------------------------------------------------ timed step (hipGraph)
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    eager_ms = timed(step, 1)
    graph = torch.cuda.CUDAGraph()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        step()  # warm on the capture stream
        torch.cuda.synchronize()
        with torch.cuda.graph(graph, stream=s):
            step()
    torch.cuda.current_stream().wait_stream(s)
    graph.replay()
    torch.cuda.synchronize()
    t0 = time.time()
    torch.cuda.synchronize()
    a, b = ev(), ev()
    a.record()
    for _ in range(args.steps):
        graph.replay()
    b.record()
    torch.cuda.synchronize()
    ms_step = a.elapsed_time(b) / args.steps
    wall = time.time() - t0
    log(f"step: graph {ms_step:.2f} ms, eager {eager_ms:.2f} ms ({args.steps} replays in {wall:.2f}s)")

    # ------------------------------------------------------------------ phase breakdown
    t_traj = timed(step_traj, 3)
    t_interp = timed(interp, 3)
    ids0, plan0, bc0 = groups[0] if groups else (None, None, None)
    t_plan0 = timed(plan0.execute, 5) if plan0 else float("nan")
    t_bc = sum(timed(lambda: fixer(i, bc), 2) for i, _, bc in groups)
    t_flux = timed(lambda: fx(pb, airm0, mask, fmass), 3) if fx else 0.0
    t_air = timed(air_masses, 3)
    flux_active = fx.active() if fx else 0
    # the trajectory with the other storage, beside it
    other = L.ADZ_F64 if pos_dt == L.ADZ_F32 else L.ADZ_F32
    ctx2 = A.AdzContext(g, other, itraj=cfg.niter)
    ctx2.adz_inittraj()
    traj2 = lambda: ctx2.adz_traject(cfg.dt, w["uu_ext"], w["vv_ext"], w["ww_ext"], w["ut0"], w["vt0"], w["zdt0"])
    traj2()
    t_traj2 = timed(traj2, 2)
    ctx2.free()
    log(f"phases: traject {t_traj:.2f} ms ({args.traj_storage} storage; {t_traj2:.2f} ms with the other), "
        f"interpolation {t_interp:.2f} ms (tracer-group kernel {t_plan0:.2f} ms), Bermejo-Conde {t_bc:.2f} ms, "
        f"LAM flux {t_flux:.2f} ms ({flux_active} of {npts_b} targets interpolated), air masses {t_air:.2f} ms")

    # ILMC (mono = 2) on one tracer group's cubic results, beside the step (not part of it:
    # the workload's tracers are fixed by Bermejo-Conde).  Two inputs: the group interpolated
    # with the reference's arithmetic (what ILMC sees in GEM), and the FP32 form of the plans
    # (whose constant 1e-12 background lands a few ulp outside [min, max]).  The fields are
    # restored before each timed call, so every call starts from the same violations.
    ilmc = None
    ilmc_cpu_in = ilmc_gpu_out = None
    Nt_il = npts(tbox)
    if not args.no_ilmc and groups:
        ids_l = groups[-1][0]
        stk = [dict(src=src[q], dst=dst[q], lin=post[k]["lin"], min=post[k]["min"], max=post[k]["max"])
               for k, q in enumerate(ids_l)]
        trs = [dict(f=dst[q], min=post[k]["min"], max=post[k]["max"]) for k, q in enumerate(ids_l)]
        il = A.IlmcLam(trs, layout5, tbox[:5], mod_off, sweep_max=2, min_max=True)
        counts = torch.zeros(3 * len(ids_l), dtype=torch.int64, device=dev)
        gil = torch.cuda.CUDAGraph()  # the call as one graph (its launches recorded once)
        sg = torch.cuda.Stream()
        sg.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(sg):
            with torch.cuda.graph(gil, stream=sg):
                il(airm0, counts)
        torch.cuda.current_stream().wait_stream(sg)
        ilmc = {"tracers": len(ids_l), "cells_per_tracer": Nt_il, "sweep_max": 2,
                "hyperplane_launches": (tbox[1] - tbox[0]) + 5 * (tbox[3] - tbox[2]) + 25 * (cfg.nk - tbox[4]) + 1}
        for form in ("ref_math", "f32_math"):
            if form == "ref_math":
                A.adz_tricub_hlt(stk, pt, ctx.ztab(True), *tbox[:5], tbox[5], mod_layout, ext=False,
                                 post=A.ADZ_POST_STORE, math=L.ADZ_MATH_REF, src_off=mod_off)
            else:
                groups[-1][1].execute()
            torch.cuda.synchronize()
            keep = [dst[q].clone() for q in ids_l]
            viol = sum(int(((dst[q] < post[k]["min"]) | (dst[q] > post[k]["max"])).sum())
                       for k, q in enumerate(ids_l))
            if form == "ref_math" and not args.no_cpu:  # one tracer's input for the CPU baseline
                ilmc_cpu_in = [dst[ids_l[0]].cpu().numpy(), post[0]["min"].cpu().numpy(),
                               post[0]["max"].cpu().numpy(), airm0.cpu().numpy()]
            t_il, t_ilg = [], []
            for rep in range(2):
                for q, kq in zip(ids_l, keep):
                    dst[q].copy_(kq)
                t_ilg.append(timed(gil.replay, 1))
            for q, kq in zip(ids_l, keep):
                dst[q].copy_(kq)
            t_il.append(timed(lambda: il(airm0, counts), 1))
            cnt = counts.reshape(-1, 3).sum(0).tolist()
            if form == "ref_math" and not args.no_cpu:
                ilmc_gpu_out = dst[ids_l[0]].cpu().numpy()
            del keep
            ilmc[form] = {"ms": round(min(t_ilg), 3), "ms_eager": round(min(t_il), 3), "cells_out_of_bounds": viol,
                          "sweeps_moving_mass": cnt[0], "min_resets": cnt[1], "max_resets": cnt[2]}
            log(f"ILMC beside the step ({form}): {min(t_ilg):.2f} ms in a graph ({min(t_il):.2f} eager) for "
                f"{len(ids_l)} tracers, {viol} cells out of bounds, {cnt[0]} sweeps moved mass, "
                f"{ilmc['hyperplane_launches']} hyperplane launches")
        del gil
        il.free()
        torch.cuda.empty_cache()

    # algorithmic bytes of the tracer-group kernel (one launch): sources + targets + lin/min/max
    # (3 extra FP32 outputs) + the positions once, at their storage width
    Nt = npts(tbox)
    ng = len(ids0) if ids0 else 0
    pos_b = 12 if pos_dt == L.ADZ_F32 else 24
    alg_plan0 = ng * 4 * (n_mod + Nt) + ng * 3 * 4 * Nt + pos_b * Nt
    achieved = alg_plan0 / (t_plan0 * 1e-3) / 1e9
    traffic = measured_traffic(args.config, tuneB)
    # interpolation-only throughput over all 64 fields
    alg_interp = (3 * 4 * n_ext + (nf - 3) * 4 * n_mod + 4 * pts_total + ntr * 12 * Nt +
                  pos_b * (npts(boxes["u"][0]) + npts(boxes["v"][0]) + Nt))
    interp_gpts = pts_total / (t_interp * 1e-3) / 1e9

    # one tracer's fixer inputs (after its group's interpolation) for the CPU baseline
    bc_cpu_in = None
    if not args.no_cpu and groups:
        ids0_, plan0_, _ = groups[0]
        plan0_.execute()
        torch.cuda.synchronize()
        q0 = ids0_[0]
        bc_cpu_in = dict(p=src[q0].cpu().numpy(), m=dst[q0].cpu().numpy(), lin=post[0]["lin"].cpu().numpy(),
                         min=post[0]["min"].cpu().numpy(), max=post[0]["max"].cpu().numpy(),
                         air1=airm1.cpu().numpy(), air0=airm0.cpu().numpy(), mask=mask.cpu().numpy(),
                         flux=fmass[2 * (q0 - 4):2 * (q0 - 4) + 2].cpu().numpy() if fx else None)

    # ------------------------------------------------------------------ accuracy vs FP64
    interp()  # dst now holds the pure FP32 interpolation (no fixer)
    torch.cuda.synchronize()
    check = range(nf) if args.err_fields <= 0 else range(min(nf, args.err_fields))
    errs = {}
    worst_rms, worst_max = 0.0, 0.0
    d64 = torch.zeros(n_mod, dtype=torch.float64, device=dev)
    for q in check:
        s64 = gen_field(q, torch.float64)
        name = "u" if q == 0 else ("v" if q == 1 else "t")
        box, xyz, zt = boxes[name]
        d64.zero_()
        A.adz_tricub_hlt([dict(src=s64, dst=d64)], xyz, zt, *box[:5], box[5], mod_layout, ext=q < 3,
                         math=L.ADZ_MATH_REF, src_off=ext_off if q < 3 else mod_off)
        i0, in_, j0, jn, k0, nk = box
        a64 = d64.reshape(mod_shape)[k0 - 1:, j0 - 1 + MH:jn + MH, i0 - 1 + MH:in_ + MH]
        a32 = dst[q].reshape(mod_shape)[k0 - 1:, j0 - 1 + MH:jn + MH, i0 - 1 + MH:in_ + MH].double()
        diff = a32 - a64
        rms = float(torch.linalg.vector_norm(diff) / torch.linalg.vector_norm(a64))
        mx = float(diff.abs().max() / a64.abs().max())
        errs[str(q)] = [mx, rms]
        worst_rms, worst_max = max(worst_rms, rms), max(worst_max, mx)
        del s64
    log(f"FP32 vs FP64: worst max rel {worst_max:.3e}, worst RMS rel {worst_rms:.3e} over {len(errs)} fields")

    # ------------------------------------------------------------------ plain FP64 leg
    nfp = min(args.fp64_fields, nf - 4) if nf > 4 else 1
    del post
    torch.cuda.empty_cache()
    s64 = [gen_field(4 + q, torch.float64) for q in range(nfp)]
    d64s = [torch.zeros(n_mod, dtype=torch.float64, device=dev) for _ in range(nfp)]
    stack64 = [dict(src=a, dst=b) for a, b in zip(s64, d64s)]
    box, xyz, zt = boxes["t"]
    run64 = lambda: A.adz_tricub_hlt(stack64, xyz, zt, *box[:5], box[5], mod_layout, ext=False,
                                     math=L.ADZ_MATH_REF, src_off=mod_off)
    run64()
    t64 = timed(run64, 3)
    plain_gpts = nfp * Nt / (t64 * 1e-3) / 1e9
    # the FP32 form on the same fields, same positions, same entry point
    stack32 = [dict(src=src[4 + q], dst=dst[4 + q]) for q in range(nfp)]
    run32 = lambda: A.adz_tricub_hlt(stack32, xyz, zt, *box[:5], box[5], mod_layout, ext=False,
                                     math=L.ADZ_MATH_F32, src_off=mod_off)
    run32()
    t32 = timed(run32, 3)
    f32_gpts = nfp * Nt / (t32 * 1e-3) / 1e9
    log(f"plain FP64: {plain_gpts:.1f} Gpt/s; FP32 same call: {f32_gpts:.1f} Gpt/s ({nfp} fields)")
    del s64, d64s, stack64
    torch.cuda.empty_cache()

    # ------------------------------------------------------------------ CPU baseline
    cpu = bc_cpu = traj_cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(np, torch, cfg, g, src, pt, tbox, mod_shape, nf, args.cpu_seconds)
        traj_cpu = traject_baseline(np, torch, L, A, S, dev)
        traj_cpu["gpu_mpoints_per_s_" + cfg.name] = round(cfg.ni * cfg.nj * cfg.nk / (t_traj * 1e-3) / 1e6, 1)
        if bc_cpu_in is not None:
            bc_cpu = bc_cpu_baseline(np, bc_cpu_in, layout5, scope)
            bc_cpu_in = None
            log(f"fixer CPU baseline (the oracle's Bermejo_Conde in the reference's sequential order, 1 thread): "
                f"{bc_cpu['cpu_ms_per_tracer_reference_order']:.0f} ms per tracer ({bc_cpu['sample']}); "
                f"GPU {t_bc / max(ntr, 1):.2f} ms per tracer")
        if ilmc_cpu_in is not None:
            ilmc["ref_math"].update(ilmc_cpu_baseline(np, ilmc_cpu_in, ilmc_gpu_out, layout5, tbox[:5]))
            log(f"ILMC CPU baseline (the oracle's sequential loop, 1 thread, one tracer): "
                f"{ilmc['ref_math']['cpu_ms_per_tracer']:.0f} ms; GPU result bit-exact: "
                f"{ilmc['ref_math']['gpu_bit_exact_vs_cpu']}")

    value = pts_total / (ms_step * 1e-3) / 1e9
    out = {
        "metric": "interpolated Gpoints/s + HBM GB/s: FP32 batched vs FP64 plain at 1e-5 RMS",
        "value": round(value, 3),
        "unit": "Gpoints/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded analytic winds, sinusoid and plume fields, generated in HBM)",
        "config": {"workload": f"{cfg.name}: {cfg.ni}x{cfg.nj}x{cfg.nk} LAM tile, halo {cfg.halo}, dt {cfg.dt:g} s, "
                               f"{nf} fields (u, v, T, q + {ntr} tracers), FP32 storage, FP64 trajectory arithmetic "
                               f"with {args.traj_storage.upper()} position storage, "
                               f"Bermejo-Conde on {ntr} tracers"
                               f"{' with the LAM boundary flux (Aranami)' if lam_flux else ''}, "
                               f"one step = trajectory + interpolation + fixer in one hipGraph",
                   "points_x_fields_per_step": pts_total, "bc_group": nb,
                   "tiles": {"uvtq": tuneA, "tracer_groups": tuneB}},
        "phases_ms": {"traject": round(t_traj, 3),
                      "traject_" + ("f64" if args.traj_storage == "f32" else "f32") + "_storage": round(t_traj2, 3),
                      "interpolation": round(t_interp, 3),
                      "bermejo_conde": round(t_bc, 3), "lam_flux": round(t_flux, 3), "air_masses": round(t_air, 3),
                      "eager_step": round(eager_ms, 3)},
        "lam_flux_targets_interpolated": flux_active,
        "ilmc_beside_step": ilmc,
        "fixer_cpu_baseline": bc_cpu,
        "traject_cpu_baseline": traj_cpu,
        "interp_gpoints_per_s": round(interp_gpts, 3),
        "interp_alg_GBps": round(alg_interp / (t_interp * 1e-3) / 1e9, 1),
        "fp64_plain_gpoints_per_s": round(plain_gpts, 3),
        "fp32_same_call_gpoints_per_s": round(f32_gpts, 3),
        "speedup_fp32_vs_fp64_plain": round(f32_gpts / plain_gpts, 3),
        "error_vs_fp64": {"worst_max_rel": worst_max, "worst_rms_rel": worst_rms, "per_field_max_rms": errs},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK, 4), "traffic": traffic,
                     "kernel": f"{plan_kernel_name(tuneB, 'float' if pos_dt == L.ADZ_F32 else 'double', 'post')} "
                               f"tracer group ({ng} fields), "
                               f"{t_plan0:.3f} ms/launch, {alg_plan0 / 1e9:.2f} GB algorithmic"},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def run_fortran(args, np, torch, L, A, S, dev):
    """The c5 step through the Fortran drop-in layer (gem_amd/fortran/adz_gpu.F90), set up
    the way GEM's modules hold the state: adz_main() (trajectory; u, v, rhsc on the
    momentum sets, rhst and rhsf on Adz_pt, extended sources), then adz_tracers(.true.)
    (q without post-processing, the 60 tracers with lin/min/max and their boundary flux
    deferred) and adz_tracers(.false.) (set_post_tr_hlt's air masses from the pressures,
    the flux masses, Bermejo-Conde on all 60), captured in one hipGraph.  The reference
    keeps lin/min/max for every tracer of the group until the fixer runs, so this needs
    about 250 GB of HBM at c5."""
    from gem_amd import adz_fortran as AF

    cfg = S.CONFIGS[args.config]
    g = S.grid_desc(cfg)
    ntr = cfg.ntracers
    t_setup = time.time()
    layout = (1 - MH, cfg.ni + MH, 1 - MH, cfg.nj + MH)
    mod_shape = (cfg.nk, cfg.nj + 2 * MH, cfg.ni + 2 * MH)
    n_mod = math.prod(mod_shape)
    mod_off = MH + (cfg.ni + 2 * MH) * MH
    pos_dt = L.ADZ_F32 if args.traj_storage == "f32" else L.ADZ_F64
    F = AF.FortranStep(g, layout, itraj=cfg.niter, pos_dtype=pos_dt, field_dtype=L.ADZ_F32, math=L.ADZ_MATH_F32)
    info = F.ctx.info
    ext_off = (1 - info["lminx"]) + info["nit"] * (1 - info["lminy"])
    ext_shape = (cfg.nk, cfg.nj + 2 * cfg.halo, cfg.ni + 2 * cfg.halo)
    w = {k: v.to(dev) for k, v in S.winds(cfg, g, device=dev).items()}

    def smooth(seed, shape, scale, off):
        f = S.smooth_field(shape, np.random.default_rng(seed), dev)
        return (off + scale * f / 8.0).reshape(-1).float()

    rhs_src = [smooth(cfg.seed + 31 * q, ext_shape, *sc) for q, sc in
               enumerate([(20.0, 0.0), (20.0, 0.0), (1.0, 0.0), (15.0, 250.0), (1.0, 0.0)])]
    rhs_dst = [torch.zeros(n_mod, device=dev) for _ in range(5)]
    P = [smooth(cfg.seed + 93, mod_shape, 0.002, 0.006)]
    P += [S.plume_tracer(mod_shape, np.random.default_rng(cfg.seed + 31 * (4 + q)), dev).reshape(-1).float()
          for q in range(ntr)]
    M = [torch.zeros(n_mod, device=dev) for _ in P]
    post = [{k: torch.zeros(n_mod, device=dev) for k in ("lin", "min", "max")} for _ in range(ntr)]
    rng = np.random.default_rng(cfg.seed + 7)
    shape2 = mod_shape[1:]
    ps = torch.from_numpy(rng.uniform(0.97, 1.03, shape2)).to(dev)
    pm = [(torch.exp(torch.tensor(g["zm"][1:cfg.nk + 2], device=dev))[:, None, None] * ps[None] *
           (1.0 + 1e-3 * q)).float().reshape(-1) for q in range(2)]
    pt = (torch.exp(torch.tensor(g["zt"][1:cfg.nk + 1], device=dev))[:, None, None] * ps[None]).float().reshape(-1)
    p0 = (1e5 * ps).float().reshape(-1)
    area = ((S.RAYT * g["h"]) ** 2 * torch.ones(shape2, dtype=torch.float64, device=dev)).reshape(-1)
    mask = torch.ones(shape2[0] * shape2[1], dtype=torch.float64, device=dev)
    airm1, airm0, pkps = (torch.zeros(n_mod, device=dev) for _ in range(3))
    F.set_time(cfg.dt)
    F.set_winds(w)
    F.set_rhs(rhs_src + [None], rhs_dst + [None], ext_off, mod_off, hydro=True)
    F.set_tracers(P, M, off=mod_off)
    F.set_groups(tricub_nt=(1, 1), tricub_wp=(2, ntr))
    F.set_wp(1, [1] * ntr, [0] * ntr, post, off=mod_off)
    F.set_bc_fields(airm1, airm0, pkps, mask, mod_off, mod_off)
    F.set_airmass(pm[0], pm[1], pt, p0, area, mod_off, mod_off)
    F.set_options(bc_min_max=True)
    cached = tuned_tiles(args.config, args.tiles)
    if cached:
        F.set_tune(0, cached["uvtq"])
        F.set_tune(1, cached["tracer_groups"])
    nfields = 5 + 1 + ntr
    box_u = (info["inu"] - info["i0u"] + 1) * (info["jn"] - info["j0"] + 1) * (cfg.nk - info["k0"] + 1)
    box_v = (info["in"] - info["i0"] + 1) * (info["jnv"] - info["j0v"] + 1) * (cfg.nk - info["k0"] + 1)
    box_m = (info["in"] - info["i0"] + 1) * (info["jn"] - info["j0"] + 1) * (cfg.nk - info["k0"] + 1)
    box_t = (info["in"] - info["i0"] + 1) * (info["jn"] - info["j0"] + 1) * (cfg.nk - info["k0t"] + 1)
    pts_total = box_u + box_v + box_m + (2 + 1 + ntr) * box_t
    log(f"fortran layer: {cfg.ni}x{cfg.nj}x{cfg.nk}, {nfields} fields, setup {time.time() - t_setup:.1f}s, "
        f"{torch.cuda.memory_allocated() / 1e9:.1f} GB allocated")
    ev = lambda: torch.cuda.Event(enable_timing=True)
    F.adz_inittraj()
    for _ in range(max(1, args.warmup)):
        F.step()  # the first call creates and caches the plans
    torch.cuda.synchronize()
    a, b = ev(), ev()
    a.record()
    F.step()
    b.record()
    torch.cuda.synchronize()
    eager_ms = a.elapsed_time(b)
    graph = torch.cuda.CUDAGraph()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        F.set_stream(s)
        with torch.cuda.graph(graph, stream=s):
            F.step()
    torch.cuda.current_stream().wait_stream(s)
    F.set_stream(torch.cuda.current_stream())
    graph.replay()
    torch.cuda.synchronize()
    a, b = ev(), ev()
    a.record()
    for _ in range(args.steps):
        graph.replay()
    b.record()
    torch.cuda.synchronize()
    ms_step = a.elapsed_time(b) / args.steps
    dfc = F.deficits(1, ntr)
    log(f"fortran layer step: graph {ms_step:.2f} ms, eager {eager_ms:.2f} ms; deficits finite: "
        f"{all(math.isfinite(x) for x in dfc)}")
    out = {"fortran_layer": True, "config": cfg.name, "fields": nfields, "points_x_fields_per_step": pts_total,
           "ms_per_step_graph": round(ms_step, 3), "ms_per_step_eager": round(eager_ms, 3),
           "gpoints_per_s": round(pts_total / (ms_step * 1e-3) / 1e9, 3),
           "tiles": cached, "steps": args.steps}
    print(json.dumps(out), flush=True)
    del graph
    F.free()


def run_c4(args, np, torch, L, A, S, dev):
    """Batching workload: 16 subdomains of uneven size (64x48x84 .. 512x384x84), each
    with its own seeded winds and trajectory; 40 FP32 tracers per subdomain share its
    thermodynamic departure points.  All tracers of all subdomains go through ONE
    persistent launch per timestep (tile parameters from adz_plan_autotune); a hipGraph
    holds `graph_steps` timesteps of (16 x adz_traject + the launch).  Plain per-field
    FP64 launches (adz_tricub_hlt, one field per call) run beside it."""
    subs = S.c4_subdomains()
    ntr = subs[0].ntracers
    t_setup = time.time()
    state = []
    for cfg in subs:
        g = S.grid_desc(cfg)
        ctx = A.AdzContext(g, L.ADZ_F32 if args.traj_storage == "f32" else L.ADZ_F64, itraj=cfg.niter)
        w = {k: v.to(dev) for k, v in S.winds(cfg, g, device=dev).items()}
        mod = (cfg.nk, cfg.nj + 2 * MH, cfg.ni + 2 * MH)
        rng = np.random.default_rng(cfg.seed)
        src = [S.lognormal_front_tracer(mod, rng, dev).reshape(-1).float() for _ in range(ntr)]
        dst = [torch.zeros_like(x) for x in src]
        i = ctx.info
        box = (i["i0"], i["in"], i["j0"], i["jn"], i["k0t"], cfg.nk)
        dl = (1 - MH, cfg.ni + MH, 1 - MH, cfg.nj + MH)
        off = MH + (cfg.ni + 2 * MH) * MH
        ctx.adz_inittraj()
        state.append(dict(cfg=cfg, ctx=ctx, w=w, src=src, dst=dst, box=box, dl=dl, off=off, n_mod=math.prod(mod)))

    def traj(st):
        w = st["w"]
        st["ctx"].adz_traject(st["cfg"].dt, w["uu_ext"], w["vv_ext"], w["ww_ext"], w["ut0"], w["vt0"], w["zdt0"])

    for st in state:
        traj(st)
    items = []
    for st in state:
        stack = [dict(src=a, dst=b) for a, b in zip(st["src"], st["dst"])]
        items.append(dict(geom=st["ctx"].ztab(True), xyz=st["ctx"].array(A.ADZ_A_PT), box=st["box"], dlayout=st["dl"],
                          ext=False, fields=A.field_list(stack, st["off"], st["off"])))
    npts = lambda b: (b[1] - b[0] + 1) * (b[3] - b[2] + 1) * (b[5] - b[4] + 1)
    pts = sum(npts(st["box"]) for st in state)
    pf = pts * ntr
    log(f"c4: {len(state)} subdomains, {pts / 1e6:.1f} M target points, {pf / 1e9:.2f} G point-fields per step, "
        f"setup {time.time() - t_setup:.1f}s")
    t0 = time.time()
    cached = None if args.retune else tuned_tiles("c4", args.tiles)
    if cached:
        tune = cached["plan"]
        log(f"tiles from gem_amd/tuned_tiles.json: {tune}")
    else:
        tune, tune_ms = A.Plan.autotune(items, L.ADZ_F32, math=L.ADZ_MATH_F32, reps=2)
        log(f"autotune: {tune} -> {tune_ms:.2f} ms per launch ({time.time() - t0:.1f}s)")
    plan = A.Plan(items, L.ADZ_F32, math=L.ADZ_MATH_F32, tune=tune)
    pos_b = 12 if args.traj_storage == "f32" else 24  # positions at their storage width
    alg = sum(ntr * 4 * (st["n_mod"] + npts(st["box"])) + pos_b * npts(st["box"]) for st in state)
    if args.profile_kernel:
        for _ in range(5):
            plan.execute()
        torch.cuda.synchronize()
        # calibration copy of known size for the PMC byte counters (1 GiB read + 1 GiB write)
        x = torch.empty(2 ** 28, dtype=torch.float32, device=dev)
        y = torch.empty_like(x)
        y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"profile_kernel": "k_plan (c4, 16 items)", "tiles": tune, "copy_bytes": 2 * 4 * 2 ** 28,
                          "algorithmic_bytes_per_launch": alg, "kernel_source_sha16": kernel_source_sha()}),
              flush=True)
        return
    default_plan = A.Plan(items, L.ADZ_F32, math=L.ADZ_MATH_F32)

    def step():
        for st in state:
            traj(st)
        plan.execute()

    ev = lambda: torch.cuda.Event(enable_timing=True)

    def timed(fn, reps):
        torch.cuda.synchronize()
        a, b = ev(), ev()
        a.record()
        for _ in range(reps):
            fn()
        b.record()
        torch.cuda.synchronize()
        return a.elapsed_time(b) / reps

    for _ in range(args.warmup):
        step()
    nst = args.graph_steps
    graph = torch.cuda.CUDAGraph()
    cs = torch.cuda.Stream()
    cs.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(cs):
        step()
        torch.cuda.synchronize()
        with torch.cuda.graph(graph, stream=cs):
            for _ in range(nst):
                step()
    torch.cuda.current_stream().wait_stream(cs)
    graph.replay()
    torch.cuda.synchronize()
    reps = max(1, args.steps // nst)
    t_graph = timed(graph.replay, reps)
    ms_step = t_graph / nst
    t_plan = timed(plan.execute, 5)
    t_plan_default = timed(default_plan.execute, 5)
    t_traj = timed(lambda: [traj(st) for st in state], 2)
    log(f"step {ms_step:.2f} ms (graph of {nst} steps x {reps}); persistent launch {t_plan:.2f} ms "
        f"(default tiles {t_plan_default:.2f} ms); 16 trajectories {t_traj:.2f} ms")

    # plain per-field FP64 launches beside it, and the FP32 error against them
    plan.execute()
    torch.cuda.synchronize()
    t64, worst_rms, worst_max, nerr = 0.0, 0.0, 0.0, 0
    for n, st in enumerate(state):
        nf64 = 4 if n % 5 else ntr  # every field on a few subdomains, 4 on the others (time scales per field)
        for f in range(nf64):
            s64 = st["src"][f].double()
            d64 = torch.zeros_like(s64)
            run = lambda: A.adz_tricub_hlt([dict(src=s64, dst=d64)], st["ctx"].array(A.ADZ_A_PT), st["ctx"].ztab(True),
                                           *st["box"][:5], st["box"][5], st["dl"], ext=False, math=L.ADZ_MATH_REF,
                                           src_off=st["off"])
            run()
            t64 += timed(run, 1) * (ntr / nf64)
            i0, in_, j0, jn, k0, nk = st["box"]
            cfg = st["cfg"]
            mod = (cfg.nk, cfg.nj + 2 * MH, cfg.ni + 2 * MH)
            sl = (slice(k0 - 1, None), slice(j0 - 1 + MH, jn + MH), slice(i0 - 1 + MH, in_ + MH))
            a64 = d64.reshape(mod)[sl]
            a32 = st["dst"][f].reshape(mod)[sl].double()
            diff = a32 - a64
            worst_rms = max(worst_rms, float(torch.linalg.vector_norm(diff) / torch.linalg.vector_norm(a64)))
            worst_max = max(worst_max, float(diff.abs().max() / a64.abs().max()))
            nerr += 1
            del s64, d64
    plain_gpts = pf / (t64 * 1e-3) / 1e9
    log(f"plain FP64 per-field launches: {t64:.1f} ms per step ({plain_gpts:.2f} Gpt/s); FP32 vs FP64 worst "
        f"RMS {worst_rms:.2e}, max {worst_max:.2e} over {nerr} fields")
    # the reference CPU interpolator on a bounded sample of the largest subdomain
    cpu = None
    if not args.no_cpu:
        big = state[-1]
        bc = big["cfg"]
        cpu = cpu_baseline(np, torch, bc, S.grid_desc(bc), big["src"], big["ctx"].array(A.ADZ_A_PT), big["box"],
                           (bc.nk, bc.nj + 2 * MH, bc.ni + 2 * MH), ntr + 4, args.cpu_seconds)
    achieved = alg / (t_plan * 1e-3) / 1e9
    out = {
        "metric": "interpolated Gpoints/s + HBM GB/s: FP32 batched vs FP64 plain at 1e-5 RMS",
        "value": round(pf / (ms_step * 1e-3) / 1e9, 3),
        "unit": "Gpoints/s",
        "n_gpus": 1, "steps": reps * nst, "warmup": args.warmup, "ms_per_step": round(ms_step, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded per-subdomain winds, lognormal front tracers, generated in HBM)",
        "config": {"workload": f"c4: 16 subdomains 64x48x84..512x384x84, {ntr} FP32 tracers each, one persistent "
                               f"launch per step, {nst} steps (trajectories + launch) per hipGraph, "
                               f"{args.traj_storage.upper()} trajectory storage",
                   "points_x_fields_per_step": pf, "tune": tune},
        "phases_ms": {"traject_16": round(t_traj, 3), "persistent_launch": round(t_plan, 3),
                      "persistent_launch_default_tiles": round(t_plan_default, 3),
                      "plain_fp64_per_field_launches": round(t64, 3)},
        "interp_gpoints_per_s": round(pf / (t_plan * 1e-3) / 1e9, 3),
        "fp64_plain_gpoints_per_s": round(plain_gpts, 3),
        "speedup_batched_fp32_vs_fp64_plain": round(t64 / t_plan, 3),
        "error_vs_fp64": {"worst_max_rel": worst_max, "worst_rms_rel": worst_rms, "fields": nerr},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK, 4), "traffic": measured_traffic("c4", tune),
                     "kernel": f"{plan_kernel_name(tune, 'float' if pos_b == 12 else 'double', '0')} over 16 items, "
                               f"{t_plan:.3f} ms/launch, "
                               f"{alg / 1e9:.2f} GB algorithmic"},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    plan.free()
    default_plan.free()


def pmc_summary(args):
    """profiles/k_plan_<config>_pmc.json: HBM bytes per launch of the batched interpolation
    kernel (c5: the 20-tracer group; c4: the launch over 16 subdomains) from rocprofv3
    FETCH_SIZE and WRITE_SIZE runs of --profile-kernel (with the same --config), each counter
    scaled by its reading for the run's closing 1 GiB device copy (1 GiB read, 1 GiB written)."""
    import csv
    import glob

    def counters(folder, name):
        (path,) = glob.glob(os.path.join(folder, "*counter_collection.csv"))
        per = {}
        for r in csv.DictReader(open(path)):
            if r["Counter_Name"] != name:
                continue
            d = per.setdefault(int(r["Dispatch_Id"]), [r["Kernel_Name"], 0.0, int(r["Start_Timestamp"]),
                                                       int(r["End_Timestamp"])])
            d[1] += float(r["Counter_Value"])
        plan = [v for v in per.values() if "k_plan" in v[0]]
        copy = [v for v in per.values() if "copy" in v[0].lower()]
        assert plan and copy, f"{path}: no k_plan or copy dispatch"
        return plan, copy[-1]

    fplan, fcopy = counters(args.pmc_summary[0], "FETCH_SIZE")
    wplan, wcopy = counters(args.pmc_summary[1], "WRITE_SIZE")
    kib = 2 ** 30 / 1024  # the copy moves 1 GiB each way
    fscale, wscale = kib / fcopy[1], kib / wcopy[1]
    fk = sum(v[1] for v in fplan) / len(fplan)
    wk = sum(v[1] for v in wplan) / len(wplan)
    dur = sum(v[3] - v[2] for v in fplan) / len(fplan)
    if args.kernel_trace:
        (st,) = glob.glob(os.path.join(args.kernel_trace, "*kernel_stats.csv"))
        for r in csv.DictReader(open(st)):
            if "k_plan" in r["Name"]:
                dur = float(r["AverageNs"])
    c4 = args.config == "c4"
    tiles = (tuned_tiles(args.config, args.tiles) or {}).get("plan" if c4 else "tracer_groups")
    out = {"kernel": fplan[0][0],
           "workload": ("c4: one persistent launch over 16 subdomains (64x48x84 .. 512x384x84), 40 FP32 tracers "
                        "each (model layout) at their thermodynamic departure points (FP32 positions), no lin/min/max"
                        if c4 else
                        "c5 tracer group: 20 FP32 tracers (model layout 2052x1028x84) at the thermodynamic departure "
                        "points of 2048x1024x84 targets, lin/min/max stored"),
           "tiles": tiles, "kernel_source_sha16": kernel_source_sha(),
           "fetch_size_kib_per_launch": fk, "write_size_kib_per_launch": wk,
           "copy_fetch_size_kib": fcopy[1], "copy_write_size_kib": wcopy[1],
           "fetch_scale": fscale, "write_scale": wscale,
           "hbm_read_bytes_per_launch": int(fk * fscale * 1024), "hbm_write_bytes_per_launch": int(wk * wscale * 1024),
           "hbm_bytes_per_launch": int((fk * fscale + wk * wscale) * 1024),
           # algorithmic bytes: c5, 20 x 4 B x (source + targets) + 20 x 12 B lin/min/max + 12 B FP32
           # positions per target; c4, the sum over the 16 subdomains of 40 x 4 B x (source + targets) +
           # 12 B x targets (the workloads as --profile-kernel builds them, FP32 trajectory storage)
           "algorithmic_bytes_per_launch": 32.41e9 if c4 else 71.04e9, "average_launch_ns": dur}
    path = os.path.join(ROOT, "profiles", f"k_plan_{args.config}_pmc.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps({"pmc_summary": path, "hbm_bytes_per_launch": out["hbm_bytes_per_launch"]}))


def plan_kernel_name(tune, pos, post):
    """the batched kernel a plan with these tiles launches for FP32 fields and arithmetic"""
    stage = (tune or {}).get("stage") or 5
    if stage in (5, 6, 7, 8):
        rows, levels = {5: (4, 1), 6: (8, 1), 7: (4, 2), 8: (4, 4)}[stage]
        return f"k_plan5<{pos},F32,{post},{rows},{levels}>"
    if stage == 3:
        return f"k_plan_pc<{pos},F32,{post}>"
    return f"k_plan<float,{pos},F32,{post},stage {stage}>"


def tuned_tiles(config, path):
    """Tile parameters adz_plan_autotune found for this workload (committed beside the
    package), or None."""
    try:
        with open(path) as f:
            return json.load(f).get(config)
    except (OSError, ValueError):
        return None


def kernel_source_sha():
    """sha256 (first 16 hex digits) of the sources the batched interpolation kernel is built
    from: a PMC measurement is used only for the build it was taken on"""
    import hashlib

    h = hashlib.sha256()
    for fn in ("adz_batched.hip", "adz_common.h", "adz_eval.h"):
        with open(os.path.join(ROOT, "gem_amd", "csrc", fn), "rb") as f:
            h.update(f.read())
    return h.hexdigest()[:16]


def measured_traffic(config, tiles):
    """HBM bytes per launch of the tracer-group kernel from the committed rocprofv3 PMC
    summary (FETCH_SIZE and WRITE_SIZE, each scaled by the same counter's reading for a
    copy of known size), when it was measured on this workload with these tiles and a
    kernel built from the current sources."""
    path = os.path.join(ROOT, "profiles", f"k_plan_{config}_pmc.json")
    try:
        with open(path) as f:
            m = json.load(f)
    except (OSError, ValueError):
        return None
    if m.get("tiles") != tiles or m.get("kernel_source_sha16") != kernel_source_sha():
        return None
    return m.get("hbm_bytes_per_launch")


def bc_cpu_baseline(np, d, layout, scope, levels=8):
    """Bermejo_Conde of one tracer by the oracle with the reference's sequential FP64 sums,
    one thread as in the reference (adz_tracers.F90:151-153), on a bounded sample: the top
    `levels` levels of the scope, scaled to the whole scope."""
    from oracle import oracle as O

    i0, in_, j0, jn, k0 = scope
    nk = layout[4]
    ks = max(k0, nk - levels + 1)
    tr = [dict(p=d["p"], m=d["m"].copy(), lin=d["lin"], min=d["min"], max=d["max"], mass=1)]
    flux = None if d["flux"] is None else np.asarray(d["flux"], np.float64)
    t0 = time.time()
    O.bermejo_conde(tr, layout, (i0, in_, j0, jn, ks), d["air1"], d["air0"], d["mask"], np.zeros(1),
                    min_max=True, sum_mode=0, flux=flux)
    ms = (time.time() - t0) * 1e3 * (nk - k0 + 1) / (nk - ks + 1)
    return {"cpu_ms_per_tracer_reference_order": round(ms, 1), "cpu_kind": "port", "cpu_cores": 1,
            "sample": f"levels {ks}..{nk} of {k0}..{nk}, scaled"}


def ilmc_cpu_baseline(np, inp, gpu_out, layout, box):
    """ILMC of one tracer by the oracle's restatement of the reference's sequential loop
    (ilmc_lam_loop.inc), one thread as in the reference, on the same input as the GPU; also
    checks the GPU result against it bit for bit at full size."""
    from oracle import oracle as O

    f, mn, mx, air = (np.ascontiguousarray(a) for a in inp)
    t0 = time.time()
    O.ilmc(f, mn, mx, air, layout, box, sweep_max=2, min_max=True)
    ms = (time.time() - t0) * 1e3
    return {"cpu_ms_per_tracer": round(ms, 1), "cpu_kind": "port", "cpu_cores": 1,
            "gpu_bit_exact_vs_cpu": bool(np.array_equal(f.view(np.uint32), gpu_out.view(np.uint32)))}


def host_info():
    """The host's CPU model, the logical CPUs this process may run on, and the threads the
    CPU baselines use: the process's CPU share (OMP_NUM_THREADS where it is set, as on a
    one-GPU box) within the CPUs visible to it."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        visible = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        visible = os.cpu_count() or 1
    share = int(os.environ.get("OMP_NUM_THREADS") or visible)
    return {"cpu_model": model, "cpus_visible": visible, "threads": max(1, min(share, visible))}


def _rate_loop(work, npf, seconds):
    """passes of work() until `seconds` have passed (at least one): point-fields per second"""
    reps, t0 = 0, time.time()
    while reps == 0 or time.time() - t0 < seconds:
        work()
        reps += 1
    return reps * npf / (time.time() - t0), reps


def cpu_baseline(np, torch, cfg, g, src, pt, tbox, mod_shape, nf, seconds):
    """The reference CPU interpolator timed on this host on a bounded sample of the same
    workload: 8 tracers at the thermodynamic departure points of 12 mid-levels,
    Tricublin_zyx1_m_n per level slice as adz_interp_hlt.F90:104-107 shares them out.
    Timed with all the process's CPU share and with one core, for the reference's own
    turbo_tricublin.c (the AVX2 build under oracle/_ref, when present) and for the oracle's
    restatement of it (the port); `value` is the reference's all-core rate where it ran,
    else the port's."""
    import concurrent.futures as cf

    from oracle import oracle as O

    host = host_info()
    i0, in_, j0, jn, k0, nk = tbox
    nx, ny = in_ - i0 + 1, jn - j0 + 1
    klo, khi = 36, 47                      # target levels of the sample
    plo, phi = klo - 6, khi + 6            # source planes copied (displacement <= 2 + stencil + margin)
    nfs = min(8, nf - 4)
    pos = pt.reshape(nk - k0 + 1, ny, nx, 3)[klo - k0:khi - k0 + 1].cpu().numpy().astype(np.float32)
    pos[..., 2] -= (plo - 1)
    assert pos[..., 2].min() >= 2.0 and pos[..., 2].max() <= phi - plo - 1.0
    fields = [np.ascontiguousarray(src[4 + q].reshape(mod_shape)[plo - 1:phi].cpu().numpy()).reshape(-1)
              for q in range(nfs)]
    levels = g["zt"][plo:phi + 1]
    ni_m, nj_m = mod_shape[2], mod_shape[1]
    base = MH + ni_m * MH
    nthreads = host["threads"]
    slices = [np.ascontiguousarray(pos[k].reshape(-1)) for k in range(pos.shape[0])]
    npf = sum(s.size // 3 for s in slices) * nfs
    legs = {}
    try:
        R = O.RefKernel("avx2")
        lv = R.setup(levels, ni_m, nj_m, 0, 0)
    except (FileNotFoundError, OSError):
        R = None
    if R is not None:
        for nth in (nthreads, 1):
            with cf.ThreadPoolExecutor(nth) as ex:  # one work item per level slice, as the OpenMP do
                rate, reps = _rate_loop(lambda: list(ex.map(lambda s: R.zyx1_m_n(lv, fields, base, s), slices)), npf,
                                        seconds * 0.3)
            legs[f"reference_{'all' if nth == nthreads else 'one'}"] = rate / 1e9
    zt = O.Ztab(levels, ni_m, nj_m, 0, 0)
    allpos = np.concatenate(slices)
    for nth in (nthreads, 1):
        O.set_threads(nth)
        rate, reps = _rate_loop(lambda: O.zyx1_m_n(zt, fields, base, allpos), npf, seconds * 0.2)
        legs[f"port_{'all' if nth == nthreads else 'one'}"] = rate / 1e9
    O.set_threads(nthreads)
    kind = "reference" if R is not None else "port"
    value = legs[f"{kind}_all"]
    log(f"cpu baseline: {', '.join(f'{k} {v:.3f}' for k, v in legs.items())} Gpt/s; {host}")
    out = {"value": round(value, 4), "unit": "Gpoints/s", "cores": nthreads, "kind": kind,
           "single_core": round(legs[f"{kind}_one"], 4),
           "legs_gpoints_per_s": {k: round(v, 4) for k, v in legs.items()},
           "host": host,
           "sample": f"{nfs} tracers x {npf // nfs} departure points (levels {klo}-{khi} of {cfg.name}), "
                     f"Tricublin_zyx1_m_n per level slice, FP32 fields; each leg timed for about "
                     f"{seconds * (0.3 if kind == 'reference' else 0.2):.1f} s"}
    if R is not None:
        out["port_over_reference"] = round(legs["port_all"] / legs["reference_all"], 3)
    return out


def traject_baseline(np, torch, L, A, S, dev):
    """The trajectory of c1 (200x150x60, the CPU-runnable workload): the oracle's
    restatement of adz_traject (FP64 iteration, the reference's FP32 storage) over the
    whole tile on all the process's CPU share and on one core, and the GPU's adz_traject on
    the same winds; departure points (all levels, every position set) per second of a
    step after the first (3 iterations)."""
    from oracle import oracle as O

    host = host_info()
    cfg = S.CONFIGS["c1"]
    g = S.grid_desc(cfg)
    w = S.winds(cfg, g)
    wn = {k: v.numpy() for k, v in w.items()}
    G = O.Grid(g)
    first = G.traject(wn, cfg.dt, storage="f32", niter=cfg.niter, want=["pt"])
    state = dict(pxyzm=first["pxyzm"], wpxyz=first["wpxyz"])
    npts = cfg.ni * cfg.nj * cfg.nk
    rates = {}
    for nth in (host["threads"], 1):
        O.set_threads(nth)
        t0, reps = time.time(), 0
        while reps == 0 or time.time() - t0 < 2.0:
            G.traject(wn, cfg.dt, storage="f32", state=state, niter=cfg.niter, want=["pt"])
            reps += 1
        rates["all" if nth == host["threads"] else "one"] = reps * npts / (time.time() - t0)
    O.set_threads(host["threads"])
    ctx = A.AdzContext(g, L.ADZ_F32, itraj=cfg.niter)
    wd = {k: v.to(dev) for k, v in w.items()}
    ctx.adz_inittraj()
    run = lambda: ctx.adz_traject(cfg.dt, wd["uu_ext"], wd["vv_ext"], wd["ww_ext"], wd["ut0"], wd["vt0"], wd["zdt0"])
    run()
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(20):
        run()
    b.record()
    torch.cuda.synchronize()
    gpu_rate = 20 * npts / (a.elapsed_time(b) * 1e-3)
    ctx.free()
    log(f"trajectory c1: CPU port {rates['all'] / 1e6:.1f} Mpts/s ({host['threads']} threads), "
        f"{rates['one'] / 1e6:.2f} Mpts/s (1 thread); GPU {gpu_rate / 1e6:.0f} Mpts/s")
    return {"workload": "c1 200x150x60, 3 iterations, FP32 position storage", "kind": "port",
            "cpu_mpoints_per_s": round(rates["all"] / 1e6, 3), "cores": host["threads"],
            "cpu_mpoints_per_s_one_core": round(rates["one"] / 1e6, 3),
            "gpu_mpoints_per_s": round(gpu_rate / 1e6, 1)}


if __name__ == "__main__":
    main()
