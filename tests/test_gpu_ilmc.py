"""GPU: ILMC shape preservation (adz_ilmc_lam; ILMC_LAM.F90:18-273, ilmc_lam_loop.inc:16-139)
against the oracle's sequential (k, j, i) visit, bit for bit, with the counts of sweeps that
moved mass and of min/max resets.

Cases:
  * random fields with dense violations (about half the cells) on a 40x30x12 domain inside
    a halo-2 layout: sweep_max 1..4, with and without Schm_autobar_L, min/max reset on and
    off; three tracers in one call;
  * the realistic case: lognormal fronts interpolated with lin/min/max at the Adz_pt
    positions of a c1 trajectory (200x150x60), then ILMC on the cubic result over the
    Adz_i0..Adz_in x Adz_j0..Adz_jn x Adz_k0t..nk domain with the c1 air mass;
  * every cell out of bounds (the largest bucket load);
  * two calls in a row on one plan, and the call captured in a graph and replayed three
    times, with eager work between the replays.
"""
import numpy as np
import pytest

from gem_amd import synth as S
from oracle import oracle as O

pytestmark = pytest.mark.gpu
MH = 2


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


def _random_case(ni, nj, nk, ntr, seed, spread=0.3):
    layout = (1 - MH, ni + MH, 1 - MH, nj + MH, nk)
    n = (ni + 2 * MH) * (nj + 2 * MH) * nk
    rng = np.random.default_rng(seed)
    trs = []
    for _ in range(ntr):
        src = rng.lognormal(0.0, 1.0, n).astype(np.float32)
        mn = (src * rng.uniform(1.0 - spread, 1.0, n)).astype(np.float32)
        mx = (src * rng.uniform(1.0, 1.0 + spread, n)).astype(np.float32)
        f = (src * rng.uniform(1.0 - 2 * spread, 1.0 + 2 * spread, n)).astype(np.float32)
        trs.append((f, mn, mx))
    air = rng.uniform(0.5, 2.0, n).astype(np.float32)
    return layout, trs, air, MH + (ni + 2 * MH) * MH


def _run_both(torch, A, layout, trs, air, off, box, sweep_max, min_max, autobar):
    dev = [dict(f=torch.from_numpy(f.copy()).cuda(), min=torch.from_numpy(mn).cuda(), max=torch.from_numpy(mx).cuda())
           for f, mn, mx in trs]
    il = A.IlmcLam(dev, layout, box, off, sweep_max=sweep_max, min_max=min_max, autobar=autobar)
    counts = torch.full((3 * len(trs),), -1, dtype=torch.int64, device="cuda")
    il(torch.from_numpy(air).cuda(), counts)
    torch.cuda.synchronize()
    il.free()
    for q, (f, mn, mx) in enumerate(trs):
        exp = f.copy()
        c = O.ilmc(exp, mn, mx, air, layout, box, sweep_max=sweep_max, min_max=min_max, autobar=autobar)
        got = dev[q]["f"].cpu().numpy()
        assert np.array_equal(_bits(got), _bits(exp)), (q, int(np.count_nonzero(got != exp)))
        assert counts[3 * q:3 * q + 3].cpu().tolist() == c.tolist(), q
    return counts.cpu().numpy()


@pytest.mark.parametrize("sweep_max", [1, 2, 3, 4])
@pytest.mark.parametrize("autobar", [False, True])
def test_ilmc_random_bit_exact(cuda, sweep_max, autobar):
    import torch

    from gem_amd import adz as A

    layout, trs, air, off = _random_case(40, 30, 12, 3, 100 + 2 * sweep_max + autobar)
    box = (2, 39, 3, 29, 2)
    counts = _run_both(torch, A, layout, trs, air, off, box, sweep_max, True, autobar)
    print(f"[ilmc] S={sweep_max} autobar={autobar}: counts {counts.tolist()}", flush=True)
    assert counts[0] > 1000


def test_ilmc_without_min_max_reset(cuda):
    import torch

    from gem_amd import adz as A

    layout, trs, air, off = _random_case(40, 30, 12, 2, 7)
    _run_both(torch, A, layout, trs, air, off, (1, 40, 1, 30, 1), 2, False, False)


def test_ilmc_every_cell_out_of_bounds(cuda):
    import torch

    from gem_amd import adz as A

    layout, trs, air, off = _random_case(24, 20, 8, 2, 3)
    trs = [(np.where(np.arange(f.size) % 2 == 0, mx * 1.5, mn * 0.5).astype(np.float32), mn, mx) for f, mn, mx in trs]
    counts = _run_both(torch, A, layout, trs, air, off, (1, 24, 1, 20, 1), 2, True, False)
    assert counts[0] > 0


@pytest.fixture(scope="module")
def C1(cuda):
    import torch

    from gem_amd import _lib as L
    from gem_amd import adz as A

    cfg = S.CONFIGS["c1"]
    g = S.grid_desc(cfg)
    ctx = A.AdzContext(g, L.ADZ_F32, itraj=cfg.niter)
    w = {k: v.cuda() for k, v in S.winds(cfg, g).items()}
    ctx.adz_inittraj()
    ctx.adz_traject(cfg.dt, w["uu_ext"], w["vv_ext"], w["ww_ext"], w["ut0"], w["vt0"], w["zdt0"])
    nx, ny = cfg.ni + 2 * MH, cfg.nj + 2 * MH
    layout = (1 - MH, cfg.ni + MH, 1 - MH, cfg.nj + MH, cfg.nk)
    off = MH + nx * MH
    rng = np.random.default_rng(41)
    stack = []
    for _ in range(3):
        src = S.lognormal_front_tracer((cfg.nk, ny, nx), rng, "cuda").float().reshape(-1).contiguous()
        stack.append(dict(src=src, dst=torch.zeros_like(src), lin=torch.zeros_like(src), min=torch.zeros_like(src),
                          max=torch.zeros_like(src)))
    i = ctx.info
    box = (i["i0"], i["in"], i["j0"], i["jn"], i["k0t"])
    A.adz_tricub_hlt(stack, ctx.array(A.ADZ_A_PT), ctx.ztab(True), *box, cfg.nk, layout[:4], ext=False,
                     post=A.ADZ_POST_STORE, math=L.ADZ_MATH_REF, src_off=off)
    _, airm0, _, _ = S.airmass(cfg, g, mh=MH)
    torch.cuda.synchronize()
    yield dict(torch=torch, A=A, layout=layout, off=off, box=box, stack=stack,
               air=airm0.reshape(-1).contiguous().cuda())
    ctx.free()


def test_ilmc_after_interpolation_bit_exact(C1):
    torch, A = C1["torch"], C1["A"]
    trs = [(e["dst"].cpu().numpy(), e["min"].cpu().numpy(), e["max"].cpu().numpy()) for e in C1["stack"]]
    air = C1["air"].cpu().numpy()
    nviol = sum(int(np.count_nonzero((f < mn) | (f > mx))) for f, mn, mx in trs)
    counts = _run_both(torch, A, C1["layout"], trs, air, C1["off"], C1["box"], 2, True, False)
    print(f"[ilmc] c1 fronts: {nviol} cells out of bounds, counts {counts.tolist()}", flush=True)
    assert nviol > 1000 and counts[0] > 0


def test_ilmc_two_calls_and_graph_replays(C1):
    torch, A = C1["torch"], C1["A"]
    base = [e["dst"].clone() for e in C1["stack"]]
    f = [b.clone() for b in base]
    dev = [dict(f=f[q], min=e["min"], max=e["max"]) for q, e in enumerate(C1["stack"])]
    il = A.IlmcLam(dev, C1["layout"], C1["box"], C1["off"], sweep_max=2)
    counts = torch.zeros(9, dtype=torch.int64, device="cuda")
    il(C1["air"], counts)                       # first call
    torch.cuda.synchronize()
    once = [x.clone() for x in f]
    c_once = counts.clone()
    il(C1["air"], counts)                       # second call on its own output
    torch.cuda.synchronize()
    for q in range(3):
        exp = once[q].cpu().numpy()
        O.ilmc(exp, C1["stack"][q]["min"].cpu().numpy(), C1["stack"][q]["max"].cpu().numpy(), C1["air"].cpu().numpy(),
               C1["layout"], C1["box"], sweep_max=2)
        assert np.array_equal(_bits(f[q].cpu().numpy()), _bits(exp)), q
    graph = torch.cuda.CUDAGraph()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    for q in range(3):
        f[q].copy_(base[q])
    with torch.cuda.stream(s):
        with torch.cuda.graph(graph, stream=s):
            il(C1["air"], counts)
    torch.cuda.current_stream().wait_stream(s)
    for rep in range(3):
        for q in range(3):
            f[q].copy_(base[q])
        counts.fill_(-1)
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        for q in range(3):
            assert torch.equal(f[q], once[q]), (rep, q)
        assert torch.equal(counts, c_once), rep
        il(C1["air"], counts)                   # eager work between the replays
        torch.cuda.synchronize()
    il.free()
