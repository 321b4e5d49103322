"""Synthetic, seeded inputs of the five workloads of the SL advection step (regular
lat-lon LAM grids, hybrid log-pressure levels, analytic winds, smooth or frontal fields).

There is no network and no GEM data archive, so every input is generated: the grid and
vertical coordinate replace what GEM's set_geomh / set_zeta derive from vgrid
(set_zeta.F90:231-300: Ver_z_8%m/t/x, dz, idz, wmstar), the winds replace the dynamics
state, and the fields replace GEM's RHS / tracer arrays.  Generation runs on any torch
device (the bench builds multi-GB inputs directly in HBM).

Workloads (sizes are ni x nj x nk of one tile = one GPU):
  c1  200x150x60, halo 4, dt=300 s, 0.1 deg: solid-body rotation 40 m/s with axis tilted
      30 deg (Williamson case 1 winds), zeta-dot = 0.1 * sin(pi*s) levels per step
      (s = normalised height); u, v, T and one tracer; 3 trajectory iterations.
  c2  1024x512x84, halo 5, dt=450 s: rotation plus 3-D Gaussian jets scaled to Courant
      numbers up to 6 in x and y and 2 in z; u, v, T, q and 10 tracers.
  c3  the c1 grid with tracers lognormal in [1e-9, 1e-3] with tanh fronts 2 cells wide.
  c4  16 subdomains from 64x48x84 to 512x384x84 with their own winds; 40 tracers.
  c5  2048x1024x84: u, v, T, q and 60 tracers (non-negative Gaussian plumes on a 1e-12
      background) with the Bermejo-Conde fixer.
"""
import math
from dataclasses import dataclass, field

import numpy as np
import torch

RAYT = 6371220.0          # Earth radius (m), as quoted in GEM's canonical cases
GRAV = 9.80616            # gravity (m s-2)
PTOP, PREF = 1000.0, 1.0e5


@dataclass
class Config:
    name: str
    ni: int
    nj: int
    nk: int
    halo: int
    dt: float
    dx_deg: float = 0.1
    lat_c: float = 0.0        # latitude of the domain centre (deg)
    lon_c: float = 0.0
    niter: int = 3
    nfields: int = 4          # interpolated fields (u, v, T, q + tracers)
    ntracers: int = 0
    cmax_xy: float = 0.0      # 0: winds not rescaled
    cmax_z: float = 0.1       # peak vertical displacement in levels per step
    jets: bool = False
    seed: int = 0
    subdomains: list = field(default_factory=list)

    @property
    def maxcfl(self):
        return self.halo - 1

    @property
    def pil(self):
        # Grd_extension = maxcfl + bsc_base (5) + bsc_ext1 (2), gem_grid_param.F90:52-59
        return self.maxcfl + 5 + 2


CONFIGS = {
    "c1": Config("c1", 200, 150, 60, 4, 300.0, nfields=4, ntracers=1, seed=101),
    "c2": Config("c2", 1024, 512, 84, 5, 450.0, nfields=14, ntracers=10, cmax_xy=6.0, cmax_z=2.0, jets=True, seed=202),
    "c3": Config("c3", 200, 150, 60, 4, 300.0, nfields=4, ntracers=4, seed=303),
    "c5": Config("c5", 2048, 1024, 84, 5, 450.0, nfields=64, ntracers=60, cmax_xy=6.0, cmax_z=2.0, jets=True,
                 seed=505),
}


def c4_subdomains(seed=404):
    """16 uneven subdomains, a linear ramp from 64x48 to 512x384 (x 84 levels)."""
    out = []
    for s in range(16):
        f = s / 15.0
        ni = int(round(64 + f * (512 - 64)))
        nj = int(round(48 + f * (384 - 48)))
        out.append(Config(f"c4_{s}", ni, nj, 84, 5, 450.0, nfields=40, ntracers=40, cmax_xy=4.0, cmax_z=1.0,
                          jets=True, seed=seed + 17 * s))
    return out


# ------------------------------------------------------------------ vertical coordinate
def vertical(nk):
    """Hybrid log-pressure levels (GEM-P, sig = +1): zeta = ln p_ref from Ztop to Zsrf.
    Returns Ver_z_8%m(0:nk+1), %t(0:nk+1), %x(0:nk) and the derived scalars adz_set and
    adz_interp_traj read (set_zeta.F90:231-335)."""
    ztop, zsrf = math.log(PTOP), math.log(PREF)
    eta = ((np.arange(1, nk + 1) - 0.5) / nk) ** 1.25
    zm = np.empty(nk + 2)
    zm[0] = ztop
    zm[1:nk + 1] = np.log(PTOP + eta * (PREF - PTOP))
    zm[nk + 1] = zsrf
    zt = np.empty(nk + 2)
    zt[0] = ztop
    zt[1:nk] = 0.5 * (zm[1:nk] + zm[2:nk + 1])
    zt[nk] = 0.5 * (zm[nk] + zsrf)
    zt[nk + 1] = zsrf
    zx = np.empty(nk + 1)
    zx[0] = ztop
    zx[1:nk] = zt[1:nk]
    zx[nk] = zsrf
    dz_m = zx[1:] - zx[:-1]                  # Ver_dz_8%m(1:nk)
    dz_t = zm[2:nk + 2] - zm[1:nk + 1]       # Ver_dz_8%t(1:nk)
    wmstar_nk = 0.5 * dz_t[nk - 1] / dz_m[nk - 1]
    idz_t_nkm1 = 1.0 / dz_t[nk - 2]
    return dict(zm=zm, zt=zt, zx=zx, zmin=ztop, zmax=zsrf, wmstar_nk=wmstar_nk, idz_t_nkm1=idz_t_nkm1)


def grid_desc(cfg):
    """Everything adz_set needs for a single-tile LAM of `cfg` (the adz_grid struct)."""
    v = vertical(cfg.nk)
    h = math.radians(cfg.dx_deg)
    lat0 = math.radians(cfg.lat_c) - 0.5 * (cfg.nj - 1) * h
    lminy, lmaxy = 1 - cfg.halo, cfg.nj + cfg.halo
    lat = lat0 + (np.arange(lminy, lmaxy + 1) - 1) * h
    return dict(ni=cfg.ni, nj=cfg.nj, nk=cfg.nk, i0g=1, j0g=1, halox=cfg.halo, haloy=cfg.halo, west=1, east=1,
                south=1, north=1, pil_w=cfg.pil, pil_e=cfg.pil, pil_s=cfg.pil, pil_n=cfg.pil, gbpil_t=0,
                bc_lam_flux=1, niu=cfg.ni - 1, njv=cfg.nj - 1, inv_hx=1.0 / h, inv_hy=1.0 / h, inv_rayt=1.0 / RAYT,
                zm=v["zm"], zt=v["zt"], zx=v["zx"], lat=lat, zmin=v["zmin"], zmax=v["zmax"],
                wmstar_nk=v["wmstar_nk"], idz_t_nkm1=v["idz_t_nkm1"], h=h, lat0=lat0,
                lon0=math.radians(cfg.lon_c) - 0.5 * (cfg.ni - 1) * h)


# ------------------------------------------------------------------ winds
def _ext_coords(cfg, g, device, stag_x=0.0, stag_y=0.0):
    """lon, lat (radians) and normalised height of the extended layout (i fastest)."""
    i = torch.arange(1 - cfg.halo, cfg.ni + cfg.halo + 1, device=device, dtype=torch.float64)
    j = torch.arange(1 - cfg.halo, cfg.nj + cfg.halo + 1, device=device, dtype=torch.float64)
    lon = g["lon0"] + (i - 1 + stag_x) * g["h"]
    lat = g["lat0"] + (j - 1 + stag_y) * g["h"]
    return lon, lat


def _jets(cfg, lon, lat, s, rng):
    """Sum of 3-D Gaussian jets (u and v components) on (k, j, i) grids."""
    u = torch.zeros((s.numel(), lat.numel(), lon.numel()), dtype=torch.float64, device=lon.device)
    v = torch.zeros_like(u)
    L = lon[None, None, :]
    P = lat[None, :, None]
    S = s[:, None, None]
    span_x = float(lon[-1] - lon[0])
    span_y = float(lat[-1] - lat[0])
    for _ in range(4):
        cx = float(lon[0]) + span_x * rng.uniform(0.2, 0.8)
        cy = float(lat[0]) + span_y * rng.uniform(0.2, 0.8)
        cz = rng.uniform(0.2, 0.8)
        sx, sy, sz = span_x * rng.uniform(0.08, 0.2), span_y * rng.uniform(0.08, 0.2), rng.uniform(0.1, 0.3)
        amp, ang = rng.uniform(0.5, 1.0), rng.uniform(0, 2 * math.pi)
        bump = amp * torch.exp(-((L - cx) / sx) ** 2 - ((P - cy) / sy) ** 2 - ((S - cz) / sz) ** 2)
        u = u + bump * math.cos(ang)
        v = v + bump * math.sin(ang)
    return u, v


def winds(cfg, g, device="cpu", seed=None):
    """FP32 winds on the extended layout (lminx:lmaxx, lminy:lmaxy, nk), flattened:
    uu_ext, vv_ext (time t1, scalar points, m/s), ww_ext (zeta-dot on the x levels, s-1),
    ut0, vt0 (time t0, staggered), zdt0 (zeta-dot at time t0)."""
    rng = np.random.default_rng(cfg.seed if seed is None else seed)
    nk = cfg.nk
    zx = torch.tensor(g["zx"], dtype=torch.float64, device=device)
    s = (zx[1:nk + 1] - zx[0]) / (zx[nk] - zx[0])           # height of x level k=1..nk in [0,1]
    dzl = torch.tensor(np.diff(g["zm"][1:nk + 2]), dtype=torch.float64, device=device)  # local spacing
    alpha, u0 = math.radians(30.0), 40.0

    def rotation(lon, lat):
        L = lon[None, None, :]
        P = lat[None, :, None]
        u = u0 * (torch.cos(P) * math.cos(alpha) + torch.sin(P) * torch.cos(L) * math.sin(alpha))
        v = -u0 * torch.sin(L) * math.sin(alpha) + 0 * P
        shape = (nk, lat.numel(), lon.numel())
        return u.expand(shape).clone(), v.expand(shape).clone()

    out = {}
    for key, sx, sy in (("ext", 0.0, 0.0), ("t0", 0.5, 0.5)):
        lon, lat = _ext_coords(cfg, g, device, sx, sy)
        u, v = rotation(lon, lat)
        if cfg.jets:
            ju, jv = _jets(cfg, lon, lat, s, np.random.default_rng(rng.integers(1 << 31)))
            u = u + 60.0 * ju
            v = v + 60.0 * jv
        out[key] = (u, v, lon, lat)
    # horizontal Courant number of the t1 winds: |u| dt / (a cos(lat) h)
    u, v, lon, lat = out["ext"]
    if cfg.cmax_xy > 0:
        cxy = torch.maximum((u.abs() / torch.cos(lat)[None, :, None]), v.abs()) * cfg.dt / (RAYT * g["h"])
        scale = cfg.cmax_xy / float(cxy.max())
    else:
        scale = 1.0
    # zeta-dot: peak displacement of cmax_z local level spacings per step, with a smooth
    # horizontal modulation when cmax_z > 0.1
    def zdot(lon, lat):
        L = lon[None, None, :]
        P = lat[None, :, None]
        prof = torch.sin(math.pi * s)[:, None, None] * dzl[:, None, None] / cfg.dt
        if cfg.cmax_z > 0.1:
            mod = torch.sin(3 * (L - lon[0]) / (lon[-1] - lon[0]) * math.pi) * torch.cos(
                2 * (P - lat[0]) / (lat[-1] - lat[0]) * math.pi)
        else:
            mod = torch.ones_like(L * P)
        return cfg.cmax_z * prof * mod

    res = {}
    ue, ve, lone, late = out["ext"]
    u0_, v0_, lon0, lat0 = out["t0"]
    res["uu_ext"] = (scale * ue).float().reshape(-1)
    res["vv_ext"] = (scale * ve).float().reshape(-1)
    res["ww_ext"] = zdot(lone, late).float().reshape(-1)
    res["ut0"] = (scale * u0_).float().reshape(-1)
    res["vt0"] = (scale * v0_).float().reshape(-1)
    res["zdt0"] = zdot(lon0, late).float().reshape(-1)
    return res


# ------------------------------------------------------------------ fields
def smooth_field(shape, rng, device, dtype=torch.float64, nwaves=8):
    """Sum of `nwaves` random-phase sinusoids with integer wavenumbers 1..6 over the box."""
    nk, nj, ni = shape
    z = torch.linspace(0, 1, nk, device=device, dtype=torch.float64)[:, None, None]
    y = torch.linspace(0, 1, nj, device=device, dtype=torch.float64)[None, :, None]
    x = torch.linspace(0, 1, ni, device=device, dtype=torch.float64)[None, None, :]
    f = torch.zeros(shape, device=device, dtype=torch.float64)
    for _ in range(nwaves):
        kx, ky, kz = rng.integers(1, 7, 3)
        a, ph = rng.uniform(0.3, 1.0), rng.uniform(0, 2 * math.pi)
        f = f + a * torch.sin(2 * math.pi * (kx * x + ky * y + kz * z) + ph)
    return f.to(dtype)


def lognormal_front_tracer(shape, rng, device, lo=1e-9, hi=1e-3, width=2.0):
    """Tracer in [lo, hi] with a sharp tanh front `width` cells wide across a random,
    gently curved surface."""
    nk, nj, ni = shape
    z = torch.arange(nk, device=device, dtype=torch.float64)[:, None, None]
    y = torch.arange(nj, device=device, dtype=torch.float64)[None, :, None]
    x = torch.arange(ni, device=device, dtype=torch.float64)[None, None, :]
    th = rng.uniform(0, 2 * math.pi)
    nx, ny = math.cos(th), math.sin(th)
    c = rng.uniform(0.3, 0.7) * (ni * abs(nx) + nj * abs(ny))
    curve = 0.08 * ni * torch.sin(2 * math.pi * y / nj * rng.integers(1, 4)) + 0.3 * z
    d = nx * x + ny * y + curve - c
    s = 0.5 + 0.5 * torch.tanh(d / width)
    return torch.exp(math.log(lo) + (math.log(hi) - math.log(lo)) * s)


def plume_tracer(shape, rng, device, background=1e-12, nplumes=3):
    """Non-negative compact Gaussian plumes on a 1e-12 background."""
    nk, nj, ni = shape
    z = torch.arange(nk, device=device, dtype=torch.float64)[:, None, None]
    y = torch.arange(nj, device=device, dtype=torch.float64)[None, :, None]
    x = torch.arange(ni, device=device, dtype=torch.float64)[None, None, :]
    f = torch.full(shape, background, device=device, dtype=torch.float64)
    for _ in range(nplumes):
        cx, cy, cz = rng.uniform(0.2, 0.8) * ni, rng.uniform(0.2, 0.8) * nj, rng.uniform(0.2, 0.8) * nk
        sx, sy, sz = rng.uniform(0.03, 0.08) * ni, rng.uniform(0.03, 0.08) * nj, rng.uniform(0.05, 0.15) * nk
        a = 10 ** rng.uniform(-6, -3)
        f = f + a * torch.exp(-((x - cx) / sx) ** 2 - ((y - cy) / sy) ** 2 - ((z - cz) / sz) ** 2)
    return f


def fields(cfg, n, device="cpu", kind="smooth", seed=None, dtype=torch.float64):
    """n source fields on the extended layout of cfg (flattened, FP64 unless dtype)."""
    rng = np.random.default_rng((cfg.seed if seed is None else seed) + 7919)
    shape = (cfg.nk, cfg.nj + 2 * cfg.halo, cfg.ni + 2 * cfg.halo)
    out = []
    for q in range(n):
        if kind == "smooth":
            f = smooth_field(shape, rng, device)
            scale, off = [(20.0, 0.0), (20.0, 0.0), (15.0, 250.0), (0.002, 0.006)][q % 4] if q < 4 else (1.0, 3.0)
            f = off + scale * f / 8.0
        elif kind == "front":
            f = lognormal_front_tracer(shape, rng, device)
        elif kind == "plume":
            f = plume_tracer(shape, rng, device)
        else:
            raise ValueError(kind)
        out.append(f.reshape(-1).to(dtype))
    return out


# ------------------------------------------------------------------ air mass (BC inputs)
def airmass(cfg, g, device="cpu", mh=2, seed=None):
    """Air mass at time P and M, pr_t/pr_s and the area mask on the model layout (halo mh),
    formed as get_air_mass_hlt.F90:83-98 does: (p_{k+1} - p_k) / g * area (dry air, sumq=0),
    with hybrid pressures p_k = exp(zeta_k) * (ps/pref)**b_k, b_k the normalised height.
    Returns FP32 airm1, airm0, pkps (nk, nj+2mh, ni+2mh) and the FP64 mask (nj+2mh, ni+2mh)."""
    rng = np.random.default_rng((cfg.seed if seed is None else seed) + 104729)
    nk = cfg.nk
    ny, nx = cfg.nj + 2 * mh, cfg.ni + 2 * mh
    zm = torch.tensor(g["zm"], dtype=torch.float64, device=device)
    zt = torch.tensor(g["zt"], dtype=torch.float64, device=device)
    b = ((zm - zm[0]) / (zm[nk + 1] - zm[0]))[:, None, None]
    bt = ((zt - zt[0]) / (zt[nk + 1] - zt[0]))[:, None, None]
    j = torch.arange(1 - mh, cfg.nj + mh + 1, device=device, dtype=torch.float64)
    i = torch.arange(1 - mh, cfg.ni + mh + 1, device=device, dtype=torch.float64)
    lat = g["lat0"] + (j - 1) * g["h"]
    area = (RAYT * g["h"]) ** 2 * torch.cos(lat)[:, None] * torch.ones(1, nx, device=device, dtype=torch.float64)
    out = []
    for t in range(2):
        ph = rng.uniform(0, 2 * math.pi, 2)
        ps = PREF * (1.0 + 0.03 * torch.sin(2 * math.pi * i[None, :] / nx + ph[0]) *
                     torch.cos(2 * math.pi * j[:, None] / ny + ph[1]))
        p = torch.exp(zm)[:, None, None] * (ps / PREF)[None] ** b           # levels 0..nk+1
        p[nk + 1] = ps
        airm = (p[2:nk + 2] - p[1:nk + 1]) / GRAV * area[None]               # k = 1..nk
        out.append(airm.float())
        if t == 0:
            pt = torch.exp(zt)[:, None, None] * (ps / PREF)[None] ** bt
            pkps = (pt[1:nk + 1] / ps[None]).float()
    mask = torch.ones(ny, nx, dtype=torch.float64, device=device)
    return out[1], out[0], pkps, mask
