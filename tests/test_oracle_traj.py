"""CPU: pin the oracle's adz_set / level search / trajectory restatement (oracle_adz.h).

The Fortran adz_* routines cannot be built here (they need rpn_comm, rmn and MPI, which
are not in the tree), so these are pinned by mathematics:
  * level search: the bracketing contract z(kk) <= z <= z(kk+1) everywhere, and equality
    with a brute-force search away from exact-level ties;
  * Lagrange vertical weights: partition of unity and reproduction of cubics;
  * trajectory: a uniform wind with no vertical motion has the closed-form departure
    point x_D = x_A - u*dt/(a*cos(lat)*h) at every iteration, z_D = the arrival level, and
    the staggered sets are its cubic midpoints;
  * a wind quadratic in x (resp. y) has a closed-form fixed point, which also pins the
    destaggering of ut0 (resp. vt0) and the u-grid (resp. v-grid) midpoints;
  * vertical motion: a zeta-dot linear in z at the arrival time only, and at the
    departure time only, gives closed-form departure heights on the momentum and the
    thermodynamic levels, including the top and bottom treatments (vw5, vw6, the dropped
    top node, the special levels 1, nk-1 and nk, wmstar), the level search and the
    fractional indices;
  * convergence of the fixed-point iteration on a smooth rotating flow.
Each closed-form test also shows that a plausible mistake (a swapped stencil, a wrong
level index, the two time levels' terms exchanged) misses its tolerance.
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


def test_quadratic_wind_in_y_closed_form():
    """v = V(y), V quadratic in the grid index y, u = w = 0: the mirror of the test above
    along y.  It pins the 4-point destaggering of the arrival winds vt0 (staggered at
    j + 1/2), the meridional update (no 1/cos(lat) factor, Adz_cy_8 multiplies u only)
    and the v-grid midpoints Adz_pmv (cubic midpoint of the departure rows j-1..j+2, minus
    1/2 in y, adz_interp_traj.F90:62-78)."""
    cfg = small_cfg()
    g = S.grid_desc(cfg)
    G = O.Grid(g)
    i = G.info
    nit, njt, nk = cfg.ni + 2 * cfg.halo, cfg.nj + 2 * cfg.halo, cfg.nk
    yc, v0, v1, v2 = cfg.nj / 2.0, -15.0, 0.25, -0.006
    V = lambda y: v0 + v1 * (y - yc) + v2 * (y - yc) ** 2
    ys = np.arange(i["lminy"], i["lmaxy"] + 1, dtype=np.float64)
    shape = (nk, njt, nit)
    vv = np.broadcast_to(V(ys)[None, :, None], shape).astype(np.float32).reshape(-1)
    vt = np.broadcast_to(V(ys + 0.5)[None, :, None], shape).astype(np.float32).reshape(-1)
    z = np.zeros(vv.size, np.float32)
    r = G.traject(dict(uu_ext=z, vv_ext=vv, ww_ext=z, ut0=z, vt0=vt, zdt0=z), cfg.dt, storage="f64")
    kap = 0.5 * cfg.dt * g["inv_rayt"] * g["inv_hy"]

    def y_dep(ya):  # Newton on s + kap V(s) = ya - kap V(ya)
        rhs = ya - kap * V(ya)
        s = ya.copy()
        for _ in range(40):
            s = s - (s + kap * V(s) - rhs) / (1.0 + kap * (v1 + 2 * v2 * (s - yc)))
        return s

    ny, nx = i["jn"] - i["j0"] + 1, i["in"] - i["i0"] + 1
    pm = r["pm"].reshape(nk, ny, nx, 3)
    ye = y_dep(np.arange(i["j0"], i["jn"] + 1, dtype=np.float64))
    inside = (ye > i["iminposy"] + 0.1) & (ye < i["imaxposy"] - 0.1)
    assert inside.sum() > ny // 2
    assert np.abs(pm[..., 1] - ye[None, :, None])[:, inside, :].max() < 1e-6
    assert np.abs(pm[..., 0] - np.arange(i["i0"], i["in"] + 1)[None, None, :]).max() < 1e-9
    nyv = i["jnv"] - i["j0v"] + 1
    pmv = r["pmv"].reshape(nk, nyv, nx, 3)
    yv = y_dep(np.arange(i["j0v"], i["jnv"] + 1, dtype=np.float64) + 0.5) - 0.5
    inv = (yv > i["iminposy"] + 0.1) & (yv < i["imaxposy"] - 0.1)
    assert inv.sum() > nyv // 2
    assert np.abs(pmv[..., 1] - yv[None, :, None])[:, inv, :].max() < 1e-6
    assert np.abs(pmv[..., 0] - np.arange(i["i0"], i["in"] + 1)[None, None, :]).max() < 1e-9
    # the arrival wind itself: vv_arr = V(j) / a, exact for a quadratic up to the FP32
    # rounding of vt0; a destaggering with swapped inner and outer weights (which still
    # reproduces constants and linears) is 1e-3 away
    nxy = cfg.ni * cfg.nj
    vva = r["warr"][nxy * nk:2 * nxy * nk].reshape(nk, cfg.nj, cfg.ni)
    jj = np.arange(1, cfg.nj + 1, dtype=np.float64)
    ref = (V(jj) * g["inv_rayt"])[None, :, None]
    assert np.abs(vva - ref).max() < 2e-7 * np.abs(ref).max()
    vt32 = V(ys + 0.5).astype(np.float32).astype(np.float64)
    at = lambda d: vt32[jj.astype(int) - 1 + d - i["lminy"]]
    swapped = ((at(-2) + at(1)) * (9.0 / 16.0) + (at(0) + at(-1)) * (-1.0 / 16.0)) * g["inv_rayt"]
    assert np.abs(vva[0, :, 0] - swapped).max() > 1e-4 * np.abs(ref).max()


# ---------------------------------------------------------------- vertical motion
def _lagrange(x, nodes, q):
    v = 1.0
    for r, n in enumerate(nodes):
        if r != q:
            v *= (x - n) / (nodes[q] - n)
    return v


def _frac_index(z, lev, nk):
    """Fractional level index of z on Ver_z_8 levels lev(0:nk+1): the interval
    [lev(nb), lev(nb+1)) that holds z, nb clamped to 1..nk-1 (extrapolating outside), plus
    the linear position inside it (adz_traject.F90:99-103, zindx adz_interp_traj.F90:189-204)."""
    z = np.asarray(z, dtype=np.float64)
    nb = np.clip(np.searchsorted(lev[1:nk + 1], z, side="right"), 1, nk - 1)
    return nb + (z - lev[nb]) / (lev[nb + 1] - lev[nb])


def _vertical_winds(cfg, ww_prof, zdt_prof):
    """u = v = 0; zeta-dot profiles on the x levels 1..nk (ww_ext for the departure wind,
    zdt0 for the arrival wind), horizontally uniform."""
    nit, njt, nk = cfg.ni + 2 * cfg.halo, cfg.nj + 2 * cfg.halo, cfg.nk
    shape = (nk, njt, nit)
    z = np.zeros(nk * njt * nit, np.float32)
    ww = np.broadcast_to(np.asarray(ww_prof)[:, None, None], shape).astype(np.float32).reshape(-1)
    zd = np.broadcast_to(np.asarray(zdt_prof)[:, None, None], shape).astype(np.float32).reshape(-1)
    return dict(uu_ext=z, vv_ext=z, ww_ext=ww, ut0=z, vt0=z, zdt0=zd)


def _split(dt, bA):
    return (1.0 - bA) * dt, bA * dt   # Schm_advec = 2: dtzD = dt - dt*bA, dtzA = dt*bA


def test_arrival_zetadot_closed_form():
    """zeta-dot at the arrival time only (zdt0 = c (z - x0) on the x levels, vanishing at the
    model top x0; ww_ext = 0, so the departure wind is zero).  Then every departure height is
    z_A - dtzA * w_A with w_A the arrival zeta-dot at the arrival level:
      * momentum levels: w_A(m_k) = c (m_k - x0) for k < nk (the cubic of
        adz_prepareWinds.F90:74-85 reproduces a linear profile, the top level's
        interpolation vw5 and the dropped x0 node at k = 2 assume zero at x0), and at
        k = nk the linear interpolation between (x_{nk-1}, w) and (x_nk, 0) (vw6);
      * thermodynamic levels: zdt0 at the same level for k < nk (adz_interp_traj.F90:83-174,
        k = 1, 2..nk-2, nk-1), and wmstar * zdt0(nk-1) at k = nk (:150-155);
    then the FP64 level search and the fractional index of the m and t position sets."""
    cfg = small_cfg()
    g = S.grid_desc(cfg)
    G = O.Grid(g)
    i = G.info
    nk = cfg.nk
    zx, zm, zt = g["zx"], g["zm"], g["zt"]
    dtzD, dtzA = _split(cfg.dt, 0.7)
    c = 0.04 / cfg.dt   # dtzA * c * (m - x0) stays below a few levels
    prof = c * (zx[1:nk + 1] - zx[0])
    prof = prof.astype(np.float32).astype(np.float64)  # what the FP32 zdt0 carries
    r = G.traject(_vertical_winds(cfg, np.zeros(nk), prof), cfg.dt, storage="f64", advec=2, bA_m=0.6, bA=0.7)
    # arrival zeta-dot on the momentum levels (Adz_ww_arr)
    wA = np.empty(nk + 1)
    for k in range(1, nk):
        wA[k] = sum(prof[q - 1] * _lagrange(zm[k], [zx[q] for q in (k - 2, k - 1, k, k + 1)], m)
                    for m, q in enumerate((k - 2, k - 1, k, k + 1)) if q >= 1) if k >= 2 else \
            prof[0] * (zm[1] - zx[0]) / (zx[1] - zx[0])
    wA[nk] = prof[nk - 2] * (zx[nk] - zm[nk]) / (zx[nk] - zx[nk - 1])
    assert np.allclose(wA[1:nk], c * (zm[1:nk] - zx[0]), rtol=1e-6)   # linear profile reproduced
    nxy = cfg.ni * cfg.nj
    wwa = r["warr"][2 * nxy * nk:].reshape(nk, nxy)
    assert np.abs(wwa - wA[1:, None]).max() <= 1e-13 * np.abs(wA[1:]).max()
    # momentum departure points (Adz_pm): fractional index of z_A - dtzA w_A
    ny, nx = i["jn"] - i["j0"] + 1, i["in"] - i["i0"] + 1
    pm = r["pm"].reshape(nk, ny, nx, 3)
    zdep = zm[1:nk + 1] - dtzA * wA[1:]
    pz_e = _frac_index(zdep, zm, nk)
    assert np.abs(pm[..., 2] - pz_e[:, None, None]).max() < 1e-11
    assert np.abs(pm[..., 0] - np.clip(np.arange(i["i0"], i["in"] + 1), i["iminposx"], i["imaxposx"])).max() < 1e-12
    # thermodynamic set (Adz_pb, Adz_pt)
    zt_dep = zt[1:nk + 1] - dtzA * np.concatenate([prof[:nk - 1], [g["wmstar_nk"] * prof[nk - 2]]])
    pb_e = _frac_index(np.clip(zt_dep, g["zmin"], g["zmax"]), zt, nk)
    nyb, nxb = i["jnb"] - i["j0b"] + 1, i["inb"] - i["i0b"] + 1
    pb = r["pb"].reshape(nk, nyb, nxb, 3)
    assert np.abs(pb[..., 2] - pb_e[:, None, None]).max() < 1e-11
    xb = np.clip(np.arange(i["i0b"], i["inb"] + 1), i["iminposx"], i["imaxposx"])
    assert np.abs(pb[..., 0] - xb).max() < 1e-12
    # the staggered sets carry the same z (cubic midpoints of a horizontally uniform field)
    nxu = i["inu"] - i["i0u"] + 1
    assert np.abs(r["pmu"].reshape(nk, ny, nxu, 3)[..., 2] - pz_e[:, None, None]).max() < 1e-12
    # planted mistakes: zdt0(nk) at the bottom level, no wmstar, w_A(m_nk) without the zero at
    # the surface -- each moves a position well beyond the tolerance
    bad = zt_dep.copy(); bad[nk - 1] = zt[nk] - dtzA * g["wmstar_nk"] * prof[nk - 1]
    assert np.abs(_frac_index(bad, zt, nk)[nk - 1] - pb[nk - 1, 0, 0, 2]) > 1e-6
    bad = zt_dep.copy(); bad[nk - 1] = zt[nk] - dtzA * prof[nk - 2]
    assert np.abs(_frac_index(bad, zt, nk)[nk - 1] - pb[nk - 1, 0, 0, 2]) > 1e-6
    assert np.abs(_frac_index(zm[nk] - dtzA * c * (zm[nk] - zx[0]), zm, nk) - pm[nk - 1, 0, 0, 2]) > 1e-6


def test_departure_zetadot_closed_form():
    """zeta-dot at the departure time only (ww_ext = c (z - x0) on the x levels, zdt0 = 0):
    the iteration z <- z_A - dtzD * w(z) (adz_traject.F90:81-83) has the fixed point
    z_D = (z_A + dtzD c x0) / (1 + dtzD c) as long as the vertical interpolation of the
    departure wind reproduces the linear profile, i.e. away from the extrapolated surface
    level.  Checks the departure heights on the momentum levels, the departure winds
    Adz_uvw_dep(3), and the thermodynamic heights, whose departure wind is the vertical
    Lagrange interpolation of Adz_uvw_dep(3) (adz_interp_traj.F90:99-102) or, at k = 1, the
    mean of levels 1 and 2 (:165-168).  dtzA = 0.7 dt and dtzD = 0.3 dt differ, so swapping
    the two terms fails."""
    cfg = small_cfg()
    g = S.grid_desc(cfg)
    G = O.Grid(g)
    i = G.info
    nk = cfg.nk
    zx, zm, zt = g["zx"], g["zm"], g["zt"]
    dtzD, dtzA = _split(cfg.dt, 0.7)
    c = 0.05 / dtzD     # contraction 0.05 per iteration
    prof = (c * (zx[1:nk + 1] - zx[0])).astype(np.float32).astype(np.float64)
    w = _vertical_winds(cfg, prof, np.zeros(nk))
    r = G.traject(w, cfg.dt, storage="f64", advec=2, bA_m=0.6, bA=0.7)                   # 6 iterations
    r = G.traject(w, cfg.dt, storage="f64", advec=2, bA_m=0.6, bA=0.7, niter=3,
                  state=dict(pxyzm=r["pxyzm"], wpxyz=r["wpxyz"]))                           # 3 more
    # the profile as stored in FP32 is c' (z - x0) only up to rounding: fit c' from it
    cc = np.polyfit(zx[1:nk] - zx[0], prof[:nk - 1], 1)
    assert abs(cc[1]) < 1e-7 * prof.max() and abs(cc[0] / c - 1) < 1e-6
    c1 = cc[0]
    zdep = lambda za: (za + dtzD * c1 * zx[0]) / (1.0 + dtzD * c1)
    # momentum levels whose departure stencil stays off the extrapolated level nk
    zD = zdep(zm[1:nk + 1])
    ok = zD < zm[nk - 3]
    assert ok.sum() >= nk - 8
    ny, nx = i["jn"] - i["j0"] + 1, i["in"] - i["i0"] + 1
    pm = r["pm"].reshape(nk, ny, nx, 3)
    tol = 1e-6   # the profile's FP32 rounding (relative 6e-8 of w) and the fitted slope
    assert np.abs(pm[ok][..., 2] - _frac_index(zD[ok], zm, nk)[:, None, None]).max() < tol
    p3 = 3 * cfg.ni * cfg.nj
    wdep = r["uvw_dep"].reshape(nk, cfg.nj, cfg.ni, 3)[..., 2]
    assert np.abs(wdep[ok] - (c1 * (zD[ok] - zx[0]))[:, None, None]).max() < 1e-6 * c1 * (zx[nk] - zx[0])
    # thermodynamic levels: k = 1 (mean of departure winds 1, 2) and 2..nk-2 (cubic) where
    # every departure wind used is exact
    nyb, nxb = i["jnb"] - i["j0b"] + 1, i["inb"] - i["i0b"] + 1
    pb = r["pb"].reshape(nk, nyb, nxb, 3)
    kt = [k for k in range(1, nk - 1) if (k == 1 and ok[0] and ok[1]) or (k >= 2 and ok[k - 2:k + 2].all())]
    assert len(kt) >= nk - 10
    zT = zdep(zt[kt])
    assert np.abs(pb[np.array(kt) - 1][..., 2] - _frac_index(zT, zt, nk)[:, None, None]).max() < tol
    # planted: the departure term with dtzA instead of dtzD
    wrong = (zm[1:nk + 1] + dtzA * c1 * zx[0]) / (1.0 + dtzA * c1)
    assert np.abs(pm[ok][:, 0, 0, 2] - _frac_index(wrong[ok], zm, nk)).max() > 1e-3


def test_level_band_equals_full_run():
    """The oracle's level band (used at the largest tile to keep the CPU time bounded):
    each point's iteration reads only the winds and its own iterate, and the
    thermodynamic sets read levels k-1..k+2, so the position sets of a band equal those of
    the whole-tile run bit for bit, at the top, inside and at the bottom."""
    cfg = small_cfg(jets=True, cmax_xy=3.0, cmax_z=2.0, seed=9)
    g = S.grid_desc(cfg)
    G = O.Grid(g)
    w = {k: v.numpy() for k, v in S.winds(cfg, g).items()}
    full = G.traject(w, cfg.dt, storage="f32")
    i, nk = G.info, cfg.nk
    for klo, khi in ((1, 4), (7, 12), (nk - 3, nk)):
        b = G.traject(w, cfg.dt, storage="f32", band=(klo, khi), want=("pm", "pmu", "pmv", "pt", "pb"))
        for name, k0 in (("pm", i["k0"]), ("pmu", i["k0"]), ("pmv", i["k0"]), ("pt", i["k0t"]), ("pb", 1)):
            per = full[name].size // (nk - k0 + 1)
            sl = slice((klo - k0) * per, (khi - k0 + 1) * per)
            assert np.array_equal(full[name][sl], b[name][sl]), (name, klo, khi)
