"""GPU parity of adz_set_post_tr (get_air_mass_hlt at times P and M, pr_t/pr_s;
get_air_mass_hlt.F90:18-103, set_post_tr_hlt.F90:53-70) against the oracle: bit-identical
air masses and pkps, FP32 and FP64 storage, moist / dry / dry with water loading, k0 > 1,
and the c5 tile (2048 x 1024 x 84) in FP32 storage."""
import numpy as np
import pytest

from gem_amd import synth as S
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def T(cuda):
    import torch

    from gem_amd import adz as A

    return torch, A


def _inputs(rng, cfg, g, mh, dt, nhyd):
    nk = cfg.nk
    shape2 = (cfg.nj + 2 * mh, cfg.ni + 2 * mh)
    pbase = np.exp(g["zm"][1:nk + 2])
    ps = rng.uniform(0.96, 1.04, shape2)
    pm = [(pbase[:, None, None] * (ps * rng.uniform(0.999, 1.001, shape2))[None]).astype(dt) for _ in range(2)]
    hu = [rng.uniform(0.0, 0.02, (nk,) + shape2).astype(dt) for _ in range(2)]
    hyd = [[rng.uniform(0.0, 0.002, (nk,) + shape2).astype(dt) for _ in range(nhyd)] for _ in range(2)]
    pt = (np.exp(g["zt"][1:nk + 1])[:, None, None] * ps[None]).astype(dt)
    p0 = (1e5 * ps).astype(dt)
    area = (S.RAYT * g["h"]) ** 2 * rng.uniform(0.9, 1.0, shape2)
    return pm, hu, hyd, pt, p0, area


def _run(T, cfg, gbpil, dt, dry, wload, nhyd, seed, mh=2):
    torch, A = T
    g = S.grid_desc(cfg)
    g["gbpil_t"] = gbpil
    ctx = A.AdzContext(g)
    k0 = ctx.info["k0t"]
    rng = np.random.default_rng(seed)
    pm, hu, hyd, pt, p0, area = _inputs(rng, cfg, g, mh, dt, nhyd)
    layout = (1 - mh, cfg.ni + mh, 1 - mh, cfg.nj + mh, cfg.nk)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a).reshape(-1)).cuda()
    tdt = torch.float32 if dt == np.float32 else torch.float64
    n3 = cfg.nk * (cfg.nj + 2 * mh) * (cfg.ni + 2 * mh)
    out = [torch.full((n3,), 7.0, dtype=tdt, device="cuda") for _ in range(3)]
    A.adz_set_post_tr(ctx, layout, dev(pm[0]), dev(pm[1]), dev(pt), dev(p0), dev(area), out[0], out[1], out[2],
                      hu_p=dev(hu[0]), hu_m=dev(hu[1]), hyd_p=[dev(h) for h in hyd[0]],
                      hyd_m=[dev(h) for h in hyd[1]], dry=dry, wload=wload)
    torch.cuda.synchronize()
    tile = (cfg.ni, cfg.nj)
    for t in range(2):
        ref = O.air_mass(pm[t], hu[t], hyd[t], area, g["zm"], layout, tile, k0, dry, wload, A.GRAV)
        got = out[t].cpu().numpy().reshape(ref.shape)
        assert np.array_equal(got.view(np.uint8), ref.view(np.uint8)), f"air mass at time {'PM'[t]}"
    ref = O.pkps(pt, p0, layout, tile)
    got = out[2].cpu().numpy().reshape(ref.shape)
    inner = (slice(None), slice(1 - layout[2], cfg.nj - layout[2] + 1), slice(1 - layout[0], cfg.ni - layout[0] + 1))
    assert np.array_equal(got[inner].view(np.uint8), ref[inner].view(np.uint8))
    outside = np.ones(got.shape, bool)
    outside[inner] = False
    assert np.all(got[outside] == 7.0)   # the halo of pkps is left as it was
    ctx.free()


@pytest.mark.parametrize("dt", [np.float32, np.float64])
@pytest.mark.parametrize("flags", [(0, 0), (1, 0), (1, 1)])
def test_airmass_bit_exact_small(T, dt, flags):
    cfg = S.Config("am", 72, 56, 24, 3, 300.0, seed=5)
    _run(T, cfg, 0, dt, flags[0], flags[1], nhyd=3, seed=31 + flags[0] + 2 * flags[1])


def test_airmass_k0_above_one(T):
    cfg = S.Config("am", 64, 48, 20, 3, 300.0, seed=6)
    _run(T, cfg, 4, np.float32, 1, 1, nhyd=2, seed=41)   # Lam_gbpil_T = 4: levels below Adz_k0t are zero


def test_airmass_c5_tile(T):
    cfg = S.CONFIGS["c5"]
    _run(T, cfg, 0, np.float32, 1, 0, nhyd=0, seed=505)
