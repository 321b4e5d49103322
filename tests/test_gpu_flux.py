"""GPU: the LAM boundary flux of the fixer (adz_bc_lam_flux, adz_bermejo_conde_lam)
against the oracle (adz_BC_LAM_Aranami_hlt.F90:18-236, mass_tr_PM.F90:101-127,
Bermejo_Conde.F90:92-105).

Setup:
- the c1 grid (200x150x60, halo 4, pilot zone 10) after one trajectory, with the Adz_pb
  positions stored in FP32 (the reference's storage) and in FP64;
- three tracers of lognormal values on the model layout (halo 2), FP32 and FP64 storage;
- the core's top level at 1 (the c1 grid) and at 3 (a top pilot zone);
- a mask with a band of 0.5.

Checks:
- FLUX_out / FLUX_in arrays equal the oracle's value for value (interpolated targets bit
  for bit; zeros may differ in sign where the reference interpolates zeros);
- the flux masses equal the oracle's bit for bit;
- the fixer with the flux equals the oracle's fixer given the same flux masses, bit for
  bit, over two calls;
- the flux captured in a graph and replayed three times gives the eager result each time.
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
    ni, nj, nk = cfg.ni, cfg.nj, cfg.nk
    nx, ny = ni + 2 * MH, nj + 2 * MH
    rng = np.random.default_rng(cfg.seed + 77)
    tr = [(1e-6 * rng.lognormal(0.0, 1.0, nx * ny * nk)) for _ in range(3)]
    airm1, airm0, pkps, mask = S.airmass(cfg, g, device="cuda", mh=MH)
    mask[:, :12] = 0.5
    info = ctx.info
    yield dict(torch=torch, A=A, L=L, cfg=cfg, g=g, ctx=ctx, pdt=pdt, ni=ni, nj=nj, nk=nk, nx=nx, ny=ny, tr=tr,
               layout=(1 - MH, ni + MH, 1 - MH, nj + MH, nk), off=MH + nx * MH,
               bbox=(info["i0b"], info["inb"], info["j0b"], info["jnb"]), pil=cfg.pil, k0t=info["k0t"],
               pb=ctx.array(A.ADZ_A_PB), airm1=airm1.reshape(-1).contiguous(), airm0=airm0.reshape(-1).contiguous(),
               mask=mask.reshape(-1).contiguous())
    ctx.free()


def _core(W, kc0):
    p = W["pil"]
    return (1 + p, W["ni"] - p, 1 + p, W["nj"] - p, kc0)


def _oracle_flux(W, trs, core, air, mask=None, autobar=False):
    zt = O.Ztab(W["g"]["zt"][1:W["nk"] + 1], W["nx"], W["ny"], 0, 0)
    fo, fi = O.bc_lam_aranami(trs, W["layout"], (W["ni"], W["nj"]), W["bbox"], core, zt, W["pb"].cpu().numpy())
    mk = (W["mask"] if mask is None else mask).cpu().numpy()
    masses = O.flux_masses(fo, fi, air.cpu().numpy(), mk, W["layout"], W["bbox"], autobar=autobar)
    return fo, fi, masses


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint8)


@pytest.mark.parametrize("kc0", [1, 3])
@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_flux_arrays_and_masses(W, dt, kc0):
    torch, A = W["torch"], W["A"]
    core = _core(W, kc0)
    trs = [t.astype(dt) for t in W["tr"]]
    dev = [dict(src=torch.from_numpy(t).cuda(), fo=torch.zeros(t.size, dtype=torch.from_numpy(t).dtype, device="cuda"),
                fi=torch.zeros(t.size, dtype=torch.from_numpy(t).dtype, device="cuda")) for t in trs]
    fx = A.BcLamFlux(dev, W["layout"], (W["ni"], W["nj"]), W["bbox"], core, W["off"], W["ctx"].ztab(True),
                     pos_dtype=W["pdt"])
    air = W["airm0"].to(dev[0]["src"].dtype)  # the air mass has the tracers' storage type
    mass = torch.zeros(2 * len(trs), dtype=torch.float64, device="cuda")
    fx(W["pb"], air, W["mask"], mass)
    torch.cuda.synchronize()
    active = fx.active()
    fo, fi, masses = _oracle_flux(W, trs, core, air)
    npts = (W["bbox"][1] - W["bbox"][0] + 1) * (W["bbox"][3] - W["bbox"][2] + 1) * W["nk"]
    print(f"[flux] {dt.__name__} kc0={kc0}: {active} of {npts} targets interpolated; masses {masses}", flush=True)
    for n in range(len(trs)):
        go, gi = dev[n]["fo"].cpu().numpy(), dev[n]["fi"].cpu().numpy()
        assert np.array_equal(go, fo[n]), (n, int(np.count_nonzero(go != fo[n])))
        assert np.array_equal(gi, fi[n]), (n, int(np.count_nonzero(gi != fi[n])))
        nz = fo[n] != 0
        assert np.array_equal(_bits(go[nz]), _bits(fo[n][nz]))
    assert np.array_equal(_bits(mass.cpu().numpy()), _bits(masses))
    assert 0 < active < npts // 4
    assert np.all(masses != 0.0)
    fx.free()


@pytest.mark.parametrize("autobar", [False, True])
def test_fixer_with_flux_bit_exact(W, autobar):
    """the fixer with the flux against the oracle; with Schm_autobar_L every mass is the
    level-1 sum weighted by geomh_area_mask_8 (mass_tr_PM.F90:79-85,105-112,
    mass_tr_X.F90:73-79), given here as the mask"""
    torch, A = W["torch"], W["A"]
    core = _core(W, 1 if autobar else W["k0t"])
    rng = np.random.default_rng(41)
    trs = []
    for t in W["tr"]:
        p = t.astype(np.float32)
        lin = (p * rng.uniform(0.9, 1.1, p.size)).astype(np.float32)
        m = (lin * rng.uniform(0.95, 1.07, p.size)).astype(np.float32)
        trs.append(dict(p=p, m=m, lin=lin, min=(np.minimum(lin, m) * np.float32(0.97)).astype(np.float32),
                        max=(np.maximum(lin, m) * np.float32(1.03)).astype(np.float32), mass=1))
    dev = [{k: (torch.from_numpy(v.copy()).cuda() if isinstance(v, np.ndarray) else v) for k, v in t.items()}
           for t in trs]
    fx = A.BcLamFlux([dict(src=d["p"]) for d in dev], W["layout"], (W["ni"], W["nj"]), W["bbox"], core, W["off"],
                     W["ctx"].ztab(True), pos_dtype=W["pdt"], autobar=autobar)
    bc = A.BermejoConde(dev, W["layout"], core, W["off"], min_max=True, autobar=autobar)
    mass = torch.zeros(2 * len(trs), dtype=torch.float64, device="cuda")
    deficit = torch.tensor([0.0, 3e-12, -2e-12], dtype=torch.float64, device="cuda")
    diag = torch.zeros(6 * len(trs), dtype=torch.float64, device="cuda")
    host_def = deficit.cpu().numpy()
    # geomh_area_mask_8 under autobar: cell areas (m2) times the mask
    wmask = W["mask"] * torch.linspace(1.1e8, 0.9e8, W["mask"].numel(), dtype=torch.float64, device="cuda") \
        if autobar else W["mask"]
    a1, a0, mk = (x.cpu().numpy() for x in (W["airm1"], W["airm0"], wmask))
    _, _, hmass = _oracle_flux(W, [t["p"] for t in trs], core, W["airm0"], wmask, autobar)
    for call in range(2):
        fx(W["pb"], W["airm0"], wmask, mass)
        bc(W["airm1"], W["airm0"], wmask, deficit, diag=diag, flux=mass)
        torch.cuda.synchronize()
        assert np.array_equal(_bits(mass.cpu().numpy()), _bits(hmass)), call
        host_def, hdiag = O.bermejo_conde(trs, W["layout"], core, a1, a0, mk, host_def, flux=hmass,
                                          autobar=autobar)
        for n, t in enumerate(trs):
            got = dev[n]["m"].cpu().numpy()
            assert np.array_equal(_bits(got), _bits(t["m"])), (call, n, int(np.count_nonzero(got != t["m"])))
        assert np.array_equal(_bits(deficit.cpu().numpy()), _bits(host_def)), call
        assert np.array_equal(_bits(diag.cpu().numpy().reshape(-1, 6)), _bits(hdiag)), call
    fx.free()
    bc.free()


def test_flux_graph_replays(W):
    torch, A = W["torch"], W["A"]
    core = _core(W, 3)
    dev = [dict(src=torch.from_numpy(t.astype(np.float32)).cuda()) for t in W["tr"]]
    fx = A.BcLamFlux(dev, W["layout"], (W["ni"], W["nj"]), W["bbox"], core, W["off"], W["ctx"].ztab(True),
                     pos_dtype=W["pdt"])
    eager = torch.zeros(2 * len(dev), dtype=torch.float64, device="cuda")
    fx(W["pb"], W["airm0"], W["mask"], eager)
    torch.cuda.synchronize()
    out = torch.full_like(eager, -1.0)
    graph = torch.cuda.CUDAGraph()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        with torch.cuda.graph(graph, stream=s):
            fx(W["pb"], W["airm0"], W["mask"], out)
    torch.cuda.current_stream().wait_stream(s)
    for rep in range(3):
        out.fill_(-1.0)
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        assert torch.equal(out, eager), rep
        fx(W["pb"], W["airm0"], W["mask"], eager)  # eager work between the replays
        torch.cuda.synchronize()
    fx.free()


def test_flux_with_no_active_target(W):
    """A core that covers the whole tile: no target outside it, no stencil reaching outside it.
    The masses are exactly +0 and the stored arrays stay zero on the target box."""
    torch, A = W["torch"], W["A"]
    core = (1, W["ni"], 1, W["nj"], 1)
    dev = [dict(src=torch.from_numpy(t.astype(np.float32)).cuda(), fo=torch.full((t.size,), 7.0, device="cuda"),
                fi=torch.full((t.size,), 7.0, device="cuda")) for t in W["tr"]]
    fx = A.BcLamFlux(dev, W["layout"], (W["ni"], W["nj"]), W["bbox"], core, W["off"], W["ctx"].ztab(True),
                     pos_dtype=W["pdt"])
    mass = torch.full((2 * len(dev),), -1.0, dtype=torch.float64, device="cuda")
    fx(W["pb"], W["airm0"], W["mask"], mass)
    torch.cuda.synchronize()
    assert fx.active() == 0
    assert np.array_equal(_bits(mass.cpu().numpy()), _bits(np.zeros(2 * len(dev))))
    i0b, inb, j0b, jnb = W["bbox"]
    shape = (W["nk"], W["ny"], W["nx"])
    sl = (slice(None), slice(j0b - 1 + MH, jnb + MH), slice(i0b - 1 + MH, inb + MH))
    for d in dev:
        assert np.all(d["fo"].cpu().numpy().reshape(shape)[sl] == 0.0)
        assert np.all(d["fi"].cpu().numpy().reshape(shape)[sl] == 0.0)
        assert np.all(d["fo"].cpu().numpy().reshape(shape)[:, :MH, :] == 7.0)  # outside the box: untouched
    fx.free()
