"""GPU: the batching workload at test size -- several subdomains of uneven size, each with
its own winds and trajectory, all tracers of all subdomains interpolated by ONE
persistent launch (adz_plan) whose tile parameters come from the project autotuner
(adz_plan_autotune), and a multi-step hipGraph of (every subdomain's adz_traject + the
persistent launch).

Checks: the tuned plan is bit-identical to one adz_tricub_hlt call per subdomain (same
FP32 arithmetic); the FP32 form is within 1e-5 relative RMS of the FP64 oracle at the
same departure points, with the oracle's cell indices; every replay of the captured
steps equals the same steps run eagerly from the same initial state."""
import numpy as np
import pytest

from gem_amd import synth as S
from oracle import oracle as O

pytestmark = pytest.mark.gpu
MH = 2
NTR = 12


@pytest.fixture(scope="module")
def W(cuda):
    """Six subdomains, a linear ramp from 40x32 to 120x90 (x 16 levels)."""
    import torch

    from gem_amd import _lib as L
    from gem_amd import adz as A

    subs = []
    for s in range(6):
        f = s / 5.0
        cfg = S.Config(f"b{s}", int(40 + f * 80), int(32 + f * 58), 16, 4, 300.0, jets=True, cmax_xy=3.0,
                       cmax_z=1.0, seed=700 + 13 * s)
        g = S.grid_desc(cfg)
        ctx = A.AdzContext(g, L.ADZ_F64, itraj=cfg.niter)
        w = {k: v.cuda() for k, v in S.winds(cfg, g).items()}
        mod = (cfg.nk, cfg.nj + 2 * MH, cfg.ni + 2 * MH)
        rng = np.random.default_rng(cfg.seed)
        src64 = [S.lognormal_front_tracer(mod, rng, "cpu").reshape(-1).numpy() for _ in range(NTR)]
        src = [torch.from_numpy(x).float().cuda() for x in src64]
        dst = [torch.zeros_like(x) for x in src]
        i = ctx.info
        box = (i["i0"], i["in"], i["j0"], i["jn"], i["k0t"], cfg.nk)
        dl = (1 - MH, cfg.ni + MH, 1 - MH, cfg.nj + MH)
        off = MH + (cfg.ni + 2 * MH) * MH
        subs.append(dict(cfg=cfg, g=g, ctx=ctx, w=w, src64=src64, src=src, dst=dst, box=box, dl=dl, off=off))
    yield torch, L, A, subs
    for s in subs:
        s["ctx"].free()


def _traj(s):
    w = s["w"]
    s["ctx"].adz_traject(s["cfg"].dt, w["uu_ext"], w["vv_ext"], w["ww_ext"], w["ut0"], w["vt0"], w["zdt0"])


def _items(A, subs, dsts=None):
    items = []
    for n, s in enumerate(subs):
        d = dsts[n] if dsts else s["dst"]
        stack = [dict(src=a, dst=b) for a, b in zip(s["src"], d)]
        items.append(dict(geom=s["ctx"].ztab(True), xyz=s["ctx"].array(A.ADZ_A_PT), box=s["box"], dlayout=s["dl"],
                          ext=False, fields=A.field_list(stack, s["off"], s["off"])))
    return items


def test_autotuned_plan_matches_single_calls_and_fp64(W):
    torch, L, A, subs = W
    for s in subs:
        s["ctx"].adz_inittraj()
        _traj(s)
    items = _items(A, subs)
    tune, ms = A.Plan.autotune(items, L.ADZ_F32, math=L.ADZ_MATH_F32, reps=2)
    assert ms > 0 and tune["fields_per_pass"] >= 1 and tune["stage"] in (1, 2, 3, 4, 5, 6, 7, 8)
    plan = A.Plan(items, L.ADZ_F32, math=L.ADZ_MATH_F32, tune=tune)
    for s in subs:
        for d in s["dst"]:
            d.zero_()
    plan.execute()
    ref = [[torch.zeros_like(x) for x in s["src"]] for s in subs]
    for s, r in zip(subs, ref):
        A.adz_tricub_hlt([dict(src=a, dst=b) for a, b in zip(s["src"], r)], s["ctx"].array(A.ADZ_A_PT),
                         s["ctx"].ztab(True), *s["box"][:5], s["box"][5], s["dl"], ext=False, math=L.ADZ_MATH_F32,
                         src_off=s["off"])
    torch.cuda.synchronize()
    for s, r in zip(subs, ref):
        for a, b in zip(s["dst"], r):
            assert torch.equal(a, b)
    # FP32 form against the FP64 oracle at the same positions (RMS over the target box)
    for s in subs[::2]:
        cfg, box = s["cfg"], s["box"]
        pos = s["ctx"].array(A.ADZ_A_PT).cpu().numpy()
        mod = (cfg.nk, cfg.nj + 2 * MH, cfg.ni + 2 * MH)
        zt = O.Ztab(s["g"]["zt"][1:cfg.nk + 1], mod[2], mod[1], 0, 0)
        exp = O.zyx1_m_n(zt, s["src64"], s["off"], pos, O.ORD_AVX2)
        i0, in_, j0, jn, k0, nk = box
        for f in range(NTR):
            got = s["dst"][f].reshape(mod)[k0 - 1:, j0 - 1 + MH:jn + MH, i0 - 1 + MH:in_ + MH].double().cpu().numpy()
            e = exp[f].reshape(got.shape)
            rms = np.linalg.norm(got - e) / np.linalg.norm(e)
            assert rms <= 1e-5, (s["cfg"].name, f, rms)
    plan.free()


def test_graph_of_steps_matches_eager(W):
    torch, L, A, subs = W
    nsteps = 4
    plan = A.Plan(_items(A, subs), L.ADZ_F32, math=L.ADZ_MATH_F32)

    def steps():
        for _ in range(nsteps):
            for s in subs:
                _traj(s)
            plan.execute()

    def snapshot():
        return [[d.clone() for d in s["dst"]] + [s["ctx"].array(A.ADZ_A_PT).clone()] for s in subs]

    for s in subs:
        s["ctx"].adz_inittraj()
        for d in s["dst"]:
            d.fill_(-1.0)  # cells outside the target boxes keep this value
    steps()  # eager from the initial state
    torch.cuda.synchronize()
    eager = snapshot()
    graph = torch.cuda.CUDAGraph()
    st = torch.cuda.Stream()
    st.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(st):
        for s in subs:
            s["ctx"].adz_inittraj()
        torch.cuda.synchronize()
        with torch.cuda.graph(graph, stream=st):
            steps()
    torch.cuda.current_stream().wait_stream(st)
    for rep in range(3):
        for s in subs:
            s["ctx"].adz_inittraj()
            for d in s["dst"]:
                d.fill_(-1.0)
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        got = snapshot()
        for a, b in zip(got, eager):
            for x, y in zip(a, b):
                assert torch.equal(x, y), rep
        steps()  # the program's own eager work between replays
        torch.cuda.synchronize()
    plan.free()
