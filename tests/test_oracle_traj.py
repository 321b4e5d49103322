"""CPU: pin the oracle's adz_set / level search / trajectory restatement (oracle_adz.h).

The Fortran adz_* routines cannot be built here (they need rpn_comm, rmn and MPI, which
are not in the tree), so these are pinned by mathematics:
  * level search: the bracketing contract z(kk) <= z <= z(kk+1) everywhere, and equality
    with a brute-force search away from exact-level ties;
  * Lagrange vertical weights: partition of unity and reproduction of cubics;
  * trajectory: a uniform wind with no vertical motion has the closed-form departure
    point x_D = x_A - u*dt/(a*cos(lat)*h) at every iteration, z_D = the arrival level, and
    the staggered sets are its cubic midpoints;
  * convergence of the fixed-point iteration on a smooth rotating flow.
"""
import math

import numpy as np
import pytest

from gem_amd import synth as S
from oracle import oracle as O


def small_cfg(**kw):
    base = dict(name="t", ni=64, nj=48, nk=20, halo=4, dt=300.0, seed=7)
    base.update(kw)
    return S.Config(**base)


@pytest.fixture(scope="module")
def grid():
    cfg = small_cfg()
    return cfg, S.grid_desc(cfg), O.Grid(S.grid_desc(cfg))


def test_bounds_match_adz_set_formulas(grid):
    cfg, g, G = grid
    i = G.info
    pil = cfg.pil
    assert (i["lminx"], i["lmaxx"]) == (1 - cfg.halo, cfg.ni + cfg.halo)
    assert (i["i0"], i["in"]) == (1 + pil - 2, cfg.ni - pil + 1)          # adz_set.F90:99-100
    assert (i["i0u"], i["inu"]) == (1 + pil, cfg.ni - 1 - pil)            # :106-109
    assert (i["i0b"], i["inb"]) == (5, cfg.ni - 4)                        # :161-164
    assert i["iminposx"] == 1 + 4 + 2e-5 and i["imaxposx"] == 1 + cfg.ni - 4 - 1 - 2e-5  # :156-159
    assert (i["k0"], i["k0t"], i["k0m"]) == (1, 1, 1)


@pytest.mark.parametrize("which", [0, 1])
def test_level_search_contract(grid, which):
    cfg, g, G = grid
    z = g["zm"] if which == 0 else g["zt"]
    nk = cfg.nk
    rng = np.random.default_rng(1)
    pos = np.concatenate([rng.uniform(z[0], z[nk + 1], 20000), z[1:nk + 1], np.nextafter(z[1:nk + 1], 0),
                          np.nextafter(z[1:nk + 1], 99), [z[0], z[nk + 1]]])
    kk = G.level_search(which, pos)
    assert kk.min() >= 0 and kk.max() <= nk + 1
    kc = np.clip(kk, 0, nk)
    assert np.all(z[kc] <= pos) and np.all(pos <= z[np.minimum(kc + 1, nk + 1)])
    # brute force away from ties: number of levels strictly below
    interior = rng.uniform(z[0], z[nk + 1], 20000)
    brute = np.searchsorted(z, interior, side="right") - 1
    assert np.array_equal(G.level_search(which, interior), brute)


def test_search_table_is_a_bin_lookup(grid):
    cfg, g, G = grid
    for which, z in ((0, g["zm"]), (1, g["zt"])):
        t = G.table(which)
        assert t[0] == 0 and np.all(np.diff(t) >= 0) and np.all(np.diff(t) <= 1)
        assert t[-1] <= cfg.nk + 1


def test_vertical_weights_reproduce_cubics(grid):
    cfg, g, G = grid
    zm, zt, zx = g["zm"], g["zt"], g["zx"]
    nk = cfg.nk
    vm = [G.table(w) for w in (2, 3, 4, 5)]
    vt = [G.table(w) for w in (6, 7, 8, 9)]
    for p in range(4):
        for k in range(2, nk):
            nodes = zx[k - 2:k + 2]
            assert math.isclose(sum(vm[q][k] * nodes[q] ** p for q in range(4)), zm[k] ** p, rel_tol=1e-10)
        for k in range(2, nk - 1):
            nodes = zm[k - 1:k + 3]
            assert math.isclose(sum(vt[q][k] * nodes[q] ** p for q in range(4)), zt[k] ** p, rel_tol=1e-10)


def _uniform_winds(cfg, g, u_ms, v_ms):
    n = (cfg.ni + 2 * cfg.halo) * (cfg.nj + 2 * cfg.halo) * cfg.nk
    lat = g["lat"]
    # u_ext is the wind in m/s; with u = U*cos(lat) the angular velocity is uniform
    coslat = np.cos(lat)[None, :, None]
    shape = (cfg.nk, cfg.nj + 2 * cfg.halo, cfg.ni + 2 * cfg.halo)
    u = np.broadcast_to(u_ms * coslat, shape).astype(np.float32).reshape(-1)
    v = np.full(n, v_ms, np.float32)
    z = np.zeros(n, np.float32)
    return dict(uu_ext=u, vv_ext=v, ww_ext=z, ut0=u, vt0=v, zdt0=z)


@pytest.mark.parametrize("storage", ["f32", "f64"])
def test_uniform_wind_closed_form(grid, storage):
    cfg, g, G = grid
    U, V = 25.0, -12.0
    r = G.traject(_uniform_winds(cfg, g, U, V), cfg.dt, storage=storage)
    i = G.info
    nx, ny = i["in"] - i["i0"] + 1, i["jn"] - i["j0"] + 1
    pm = r["pm"].reshape(cfg.nk, ny, nx, 3).astype(np.float64)
    dx = U * cfg.dt / (S.RAYT * g["h"])   # uniform angular velocity U/a
    dy = V * cfg.dt / (S.RAYT * g["h"])
    ii = np.arange(i["i0"], i["in"] + 1)[None, None, :]
    jj = np.arange(i["j0"], i["jn"] + 1)[None, :, None]
    tol = 2e-5 if storage == "f32" else 1e-7  # inputs are FP32 winds (6e-8 relative)
    xe = np.clip(ii - dx, i["iminposx"], i["imaxposx"])
    ye = np.clip(jj - dy, i["iminposy"], i["imaxposy"])
    assert np.abs(pm[..., 0] - xe).max() < tol
    assert np.abs(pm[..., 1] - ye).max() < tol
    kk = np.arange(1, cfg.nk + 1)[:, None, None]
    assert np.abs(pm[..., 2] - kk).max() < 1e-5
    # u-grid set: cubic midpoint in x minus 1/2 -> the same displacement
    nxu = i["inu"] - i["i0u"] + 1
    pmu = r["pmu"].reshape(cfg.nk, ny, nxu, 3).astype(np.float64)
    iu = np.arange(i["i0u"], i["inu"] + 1)[None, None, :]
    assert np.abs(pmu[..., 0] - np.clip(iu - dx, i["iminposx"], i["imaxposx"])).max() < tol


def test_iteration_converges_on_rotation():
    cfg = small_cfg(jets=True, cmax_xy=3.0, cmax_z=1.0)
    g = S.grid_desc(cfg)
    G = O.Grid(g)
    w = {k: v.numpy() for k, v in S.winds(cfg, g).items()}
    r6 = G.traject(w, cfg.dt, storage="f64", state=None, niter=3)             # 6 iterations (inittraj)
    st = dict(pxyzm=r6["pxyzm"], wpxyz=r6["wpxyz"])
    r9 = G.traject(w, cfg.dt, storage="f64", state=st, niter=3)               # +3 from converged start
    d = np.abs(r9["pm"] - r6["pm"]).max()
    assert d < 1e-3, d   # fixed point reached to well below a cell


def test_traject_deterministic_and_storage_consistent(grid):
    cfg, g, G = grid
    w = {k: v.numpy() for k, v in S.winds(cfg, g).items()}
    a = G.traject(w, cfg.dt, storage="f32")
    b = G.traject(w, cfg.dt, storage="f32")
    assert all(np.array_equal(a[k], b[k]) for k in a)
    c = G.traject(w, cfg.dt, storage="f64")
    assert np.abs(a["pt"].astype(np.float64) - c["pt"]).max() < 1e-3


def test_quadratic_wind_closed_form():
    """u = U(x) cos(lat), U quadratic in the grid index x, v = w = 0.  The midpoint
    iteration's fixed point solves x_D + k U(x_D) = x_A - k U(x_A), k = dt/2 /(a h), and
    the cubic interpolation of the departure wind is exact for a quadratic, so every
    departure point has a closed form.  A quadratic (unlike a constant or linear wind)
    also pins the 4-point destaggering of the arrival winds ut0 (a symmetric stencil with
    swapped inner/outer weights reproduces linears, not quadratics), the 1/cos(lat)
    factor and the staggered u-grid midpoints."""
    cfg = small_cfg()
    g = S.grid_desc(cfg)
    G = O.Grid(g)
    i = G.info
    nit, njt, nk = cfg.ni + 2 * cfg.halo, cfg.nj + 2 * cfg.halo, cfg.nk
    xc, u0, u1, u2 = cfg.ni / 2.0, 20.0, 0.2, 0.004
    U = lambda x: u0 + u1 * (x - xc) + u2 * (x - xc) ** 2
    xs = np.arange(i["lminx"], i["lmaxx"] + 1, dtype=np.float64)
    cosl = np.cos(np.asarray(g["lat"]))[None, :, None]
    shape = (nk, njt, nit)
    uu = np.broadcast_to(U(xs)[None, None, :] * cosl, shape).astype(np.float32).reshape(-1)
    ut = np.broadcast_to(U(xs + 0.5)[None, None, :] * cosl, shape).astype(np.float32).reshape(-1)
    z = np.zeros(uu.size, np.float32)
    r = G.traject(dict(uu_ext=uu, vv_ext=z, ww_ext=z, ut0=ut, vt0=z, zdt0=z), cfg.dt, storage="f64")
    kap = 0.5 * cfg.dt * g["inv_rayt"] * g["inv_hx"]

    def x_dep(xa):  # Newton on s + kap U(s) = xa - kap U(xa)
        rhs = xa - kap * U(xa)
        s = xa.copy()
        for _ in range(40):
            s = s - (s + kap * U(s) - rhs) / (1.0 + kap * (u1 + 2 * u2 * (s - xc)))
        return s

    ny, nx = i["jn"] - i["j0"] + 1, i["in"] - i["i0"] + 1
    pm = r["pm"].reshape(nk, ny, nx, 3)
    xe = x_dep(np.arange(i["i0"], i["in"] + 1, dtype=np.float64))
    inside = (xe > i["iminposx"] + 0.1) & (xe < i["imaxposx"] - 0.1)
    assert inside.sum() > nx // 2
    assert np.abs(pm[..., 0] - xe[None, None, :])[..., inside].max() < 1e-6
    assert np.abs(pm[..., 1] - np.arange(i["j0"], i["jn"] + 1)[None, :, None]).max() < 1e-9
    nxu = i["inu"] - i["i0u"] + 1
    pmu = r["pmu"].reshape(nk, ny, nxu, 3)
    xu = x_dep(np.arange(i["i0u"], i["inu"] + 1, dtype=np.float64) + 0.5) - 0.5
    inu = (xu > i["iminposx"] + 0.1) & (xu < i["imaxposx"] - 0.1)
    assert np.abs(pmu[..., 0] - xu[None, None, :])[..., inu].max() < 1e-6
