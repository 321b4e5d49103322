// CPU ORACLE -- test infrastructure only (see oracle_interp.h).
//
// Restatement of the Fortran adz_* layer of GEM's gemdyn around the interpolation
// kernel: the SL set-up tables (adz_set.F90), wind preparation (adz_prepareWinds.F90),
// the departure-point iteration (adz_traject.F90), the derived position sets
// (adz_interp_traj.F90), the initial positions (adz_inittraj.F90), and the Bermejo-Conde
// mass fixer with its FP64 mass sums (Bermejo_Conde.F90, mass_tr_PM.F90, mass_tr_X.F90,
// almost_zero.F90).
//
// Storage policy S (template parameter):
//   S = float  : "O-ref", the reference's declared storage (Adz_uvw_d, Adz_uu_arr,
//                Adz_uvw_dep, Adz_pxyzm, Adz_pm/pmu/pmv/pt/pb are `real`, adz_mem.F90:
//                78-84) with FP64 arithmetic exactly where the Fortran source does it.
//   S = double : "O-64", every derived wind and position kept in FP64.
// Fortran expressions are evaluated left to right without FMA contraction.
#pragma once
#include <quadmath.h>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "oracle_interp.h"

namespace ora {

// Grid / configuration description (the inputs adz_set reads from GEM modules).
struct GridDesc {
  int ni, nj, nk;            // l_ni, l_nj, l_nk (= G_nk; one tile)
  int i0g, j0g;              // l_i0, l_j0
  int halox, haloy;          // Adz_halox / Adz_haloy (= Grd_maxcfl + 1)
  int west, east, south, north;
  int pil_w, pil_e, pil_s, pil_n;  // Glb_pil_* (applied where the tile touches the edge)
  int gbpil_t;               // Lam_gbpil_T
  int bc_lam_flux;           // Adz_BC_LAM_flux
  int niu, njv;              // l_niu, l_njv
  double inv_hx, inv_hy;     // geomh_inv_hx_8, geomh_inv_hy_8
  double inv_rayt;           // Dcst_inv_rayt_8
  const double* zm;          // Ver_z_8%m(0:nk+1)
  const double* zt;          // Ver_z_8%t(0:nk+1)
  const double* zx;          // Ver_z_8%x(0:nk)
  const double* lat;         // latitude (radians) for local j = lminy..lmaxy
  double zmin, zmax;         // Ver_zmin_8, Ver_zmax_8
  double wmstar_nk;          // Ver_wmstar_8(nk)
  double idz_t_nkm1;         // Ver_idz_8%t(nk-1)
};

// Everything adz_set derives (adz_set.F90:65-311, single tile, ADZ_OD_L=.false.).
struct Grid {
  GridDesc d;
  int ni, nj, nk;
  int lminx, lmaxx, lminy, lmaxy, nit, njt, nij;
  int i0, in, j0, jn, i0u, inu, j0v, jnv, i0b, inb, j0b, jnb, k0, k0t, k0m;
  double iminposx, imaxposx, iminposy, imaxposy;
  double sig;
  std::vector<double> zm, zt, zx;                          // 0-based copies of Ver_z_8
  std::vector<double> delz_m, delz_t, odelz_m, odelz_t;    // 0..nk
  double ovdzm, ovdzt;
  std::vector<int> search_m, search_t;                     // 1..pnz stored at [k-1]
  std::vector<double> vw1m, vw2m, vw3m, vw4m, vw1t, vw2t, vw3t, vw4t;  // index k (1..nk)
  double vw5, vw6;
  std::vector<double> cy;                                  // index j - lminy
  Ztab ztab_q, ztab_t;
};

inline double lag3(double zz, double z1, double z2, double z3, double z4) {
  // adz_set.F90:530-536
  return (((zz - z2) * (zz - z3) * (zz - z4)) / ((z1 - z2) * (z1 - z3) * (z1 - z4)));
}

// Level-search table, adz_set.F90:193-215 (REAL128 smallest_dz / posz as in the source).
inline void build_search(const std::vector<double>& z, const std::vector<double>& delz, int nk, double sig,
                         double* ovdz_out, std::vector<int>* table) {
  __float128 smallest = (__float128)(sig * delz[1]);
  for (int k = 2; k <= nk; k++) {
    __float128 v = (__float128)(sig * delz[k]);
    if (v < smallest) smallest = v;
  }
  double ovdz = (double)((__float128)1.0 / smallest);
  double t = sig * (z[nk + 1] - z[0]) * ovdz;
  double pn = (double)1.0f + t;
  int pnz = (int)std::lround(pn);  // NINT: round half away from zero
  ovdz = (double)((__float128)sig / smallest);
  table->assign(pnz, 0);
  int k0 = 0;
  for (int k = 1; k <= pnz; k++) {
    __float128 posz = (__float128)(sig * z[0]) + (__float128)(double)(k - 1) * smallest;
    if (posz > (__float128)(sig * z[k0 + 1])) k0 = std::min(nk + 1, k0 + 1);
    (*table)[k - 1] = k0;
  }
  *ovdz_out = ovdz;
}

inline void adz_set(const GridDesc& d, Grid* g) {
  g->d = d;
  const int ni = d.ni, nj = d.nj, nk = d.nk;
  g->ni = ni; g->nj = nj; g->nk = nk;
  g->lminx = 1 - d.halox; g->lmaxx = ni + d.halox;
  g->lminy = 1 - d.haloy; g->lmaxy = nj + d.haloy;
  g->nit = g->lmaxx - g->lminx + 1; g->njt = g->lmaxy - g->lminy + 1; g->nij = g->nit * g->njt;
  const int ext = 1;  // LAM
  const int pil_w = d.west ? d.pil_w : 0, pil_e = d.east ? d.pil_e : 0;
  const int pil_s = d.south ? d.pil_s : 0, pil_n = d.north ? d.pil_n : 0;
  g->i0 = 1 + pil_w - (ext + 1) * d.west;
  g->in = ni - pil_e + ext * d.east;
  g->j0 = 1 + pil_s - (ext + 1) * d.south;
  g->jn = nj - pil_n + ext * d.north;
  g->i0u = 1 + pil_w; g->j0v = 1 + pil_s;
  g->inu = d.niu - pil_e; g->jnv = d.njv - pil_n;
  g->k0 = d.gbpil_t + 1;
  g->k0t = std::max(1, d.gbpil_t);
  g->k0m = std::max(g->k0t - 2, 1);
  if (d.bc_lam_flux != 0) g->k0m = 1;
  const double EPS_8 = 2.0e-5;
  const int BCS_BASE = 4;
  g->iminposx = d.i0g + g->lminx + EPS_8;
  g->imaxposx = d.i0g + g->lmaxx - 2 - EPS_8;
  g->iminposy = d.j0g + g->lminy + EPS_8;
  g->imaxposy = d.j0g + g->lmaxy - 2 - EPS_8;
  if (d.west) g->iminposx = d.i0g + BCS_BASE + EPS_8;
  if (d.east) g->imaxposx = d.i0g + ni - BCS_BASE - 1 - EPS_8;
  if (d.south) g->iminposy = d.j0g + BCS_BASE + EPS_8;
  if (d.north) g->imaxposy = d.j0g + nj - BCS_BASE - 1 - EPS_8;
  g->i0b = 1 + BCS_BASE * d.west; g->inb = ni - BCS_BASE * d.east;
  g->j0b = 1 + BCS_BASE * d.south; g->jnb = nj - BCS_BASE * d.north;

  g->zm.assign(d.zm, d.zm + nk + 2);
  g->zt.assign(d.zt, d.zt + nk + 2);
  g->zx.assign(d.zx, d.zx + nk + 1);
  g->sig = (g->zm[nk] - g->zm[1]) / std::fabs(g->zm[nk] - g->zm[1]);
  g->delz_m.assign(nk + 1, 9.e33f);
  g->delz_t.assign(nk + 1, 9.e33f);
  g->odelz_m.assign(nk + 1, 0.0);
  g->odelz_t.assign(nk + 1, 0.0);
  const float verysmall = 1.17549435e-38f * 1.e5f;  // tiny(real)*1.e5 in FP32
  for (int k = 0; k <= nk; k++) {
    if (std::fabs(g->zm[k + 1] - g->zm[k]) > (double)verysmall) g->delz_m[k] = g->zm[k + 1] - g->zm[k];
    if (std::fabs(g->zt[k + 1] - g->zt[k]) > (double)verysmall) g->delz_t[k] = g->zt[k + 1] - g->zt[k];
    g->odelz_m[k] = 1.0 / g->delz_m[k];
    g->odelz_t[k] = 1.0 / g->delz_t[k];
  }
  build_search(g->zm, g->delz_m, nk, g->sig, &g->ovdzm, &g->search_m);
  build_search(g->zt, g->delz_t, nk, g->sig, &g->ovdzt, &g->search_t);

  g->cy.resize(g->njt);
  for (int j = g->lminy; j <= g->lmaxy; j++) g->cy[j - g->lminy] = 1.0 / std::cos(d.lat[j - g->lminy]);

  g->vw1m.assign(nk + 1, 0.0); g->vw2m = g->vw1m; g->vw3m = g->vw1m; g->vw4m = g->vw1m;
  g->vw1t = g->vw1m; g->vw2t = g->vw1m; g->vw3t = g->vw1m; g->vw4t = g->vw1m;
  const auto& zm = g->zm; const auto& zt = g->zt; const auto& zx = g->zx;
  for (int k = 2; k <= nk - 1; k++) {
    double rx = zm[k], ra = zx[k - 2], rb = zx[k - 1], rc = zx[k], rd = zx[k + 1];
    g->vw1m[k] = lag3(rx, ra, rb, rc, rd);
    g->vw2m[k] = lag3(rx, rb, ra, rc, rd);
    g->vw3m[k] = lag3(rx, rc, ra, rb, rd);
    g->vw4m[k] = lag3(rx, rd, ra, rb, rc);
  }
  for (int k = 2; k <= nk - 2; k++) {
    double rx = zt[k], ra = zm[k - 1], rb = zm[k], rc = zm[k + 1], rd = zm[k + 2];
    g->vw1t[k] = lag3(rx, ra, rb, rc, rd);
    g->vw2t[k] = lag3(rx, rb, ra, rc, rd);
    g->vw3t[k] = lag3(rx, rc, ra, rb, rd);
    g->vw4t[k] = lag3(rx, rd, ra, rb, rc);
  }
  g->vw5 = (zx[0] - zm[1]) / (zx[0] - zx[1]);
  g->vw6 = (zm[nk] - zx[nk]) / (zx[nk - 1] - zx[nk]);
  // Vsearch_setup_plus(Ver_z_8%m(1:G_nk), G_nk, Adz_nit, Adz_njt, 1-l_i0, 1-l_j0), :128-131
  g->ztab_q = ztab_setup(&zm[1], nk, g->nit, g->njt, 1 - d.i0g, 1 - d.j0g);
  g->ztab_t = ztab_setup(&zt[1], nk, g->nit, g->njt, 1 - d.i0g, 1 - d.j0g);
}

// ---------- array views (Fortran column-major with arbitrary lower bounds) ----------
struct Ext3 {  // (lminx:lmaxx, lminy:lmaxy, 1:nk)
  int lx, ly, nx, ny;
  long idx(int i, int j, int k) const { return (long)(i - lx) + (long)nx * ((j - ly) + (long)ny * (k - 1)); }
};

// Level search, adz_traject.F90:85-90 / zindx in adz_interp_traj.F90:194-202
inline int level_search(const std::vector<double>& z, const std::vector<int>& tab, double ovdz, double sig,
                        int nk, double pos) {
  int kk1 = (int)((pos - z[0]) * ovdz + 1.0);
  kk1 = std::min(std::max(1, kk1), (int)tab.size());
  kk1 = tab[kk1 - 1];
  if (sig * pos > sig * z[std::min(kk1 + 1, nk + 1)]) kk1 = kk1 + 1;
  if (sig * pos < sig * z[kk1]) kk1 = kk1 - 1;
  return kk1;
}

// zindx, adz_interp_traj.F90:189-204 (posz is clamped in place, as the intent(inout) does)
inline double zindx(const Grid& g, double posz) {
  posz = std::min(std::max(posz, g.d.zmin), g.d.zmax);
  int kk1 = level_search(g.zt, g.search_t, g.ovdzt, g.sig, g.nk, posz);
  kk1 = std::min(g.nk + 1, std::max(0, kk1));
  int nb = std::max(std::min(kk1, g.nk - 1), 1);
  return (posz - g.zt[nb]) * g.odelz_t[nb] + (double)nb;
}

template <class S>
struct TrajState {
  // derived winds (adz_prepareWinds)
  std::vector<S> uvw_d;     // (3, lminx:lmaxx, lminy:lmaxy, nk)
  std::vector<S> uu_arr, vv_arr, ww_arr;  // (ni, nj, nk)
  std::vector<S> uvw_dep;   // (3, ni, nj, k0m:nk)
  std::vector<S> pxyzm;     // (3, ni, nj, k0m:nk)
  std::vector<double> wpxyz;  // (-1:ni+2, -1:nj+2, nk, 3)
  std::vector<S> pm, pmu, pmv, pt, pb;
};

template <class S>
inline void alloc_state(const Grid& g, TrajState<S>* s) {
  const long nxy = (long)g.ni * g.nj;
  s->uvw_d.assign(3L * g.nij * g.nk, (S)0);
  s->uu_arr.assign(nxy * g.nk, (S)0); s->vv_arr = s->uu_arr; s->ww_arr = s->uu_arr;
  s->uvw_dep.assign(3 * nxy * (g.nk - g.k0m + 1), (S)0);
  s->pxyzm.assign(3 * nxy * (g.nk - g.k0m + 1), (S)0);
  s->wpxyz.assign((long)(g.ni + 4) * (g.nj + 4) * g.nk * 3, 0.0);
  s->pm.assign(3L * (g.in - g.i0 + 1) * (g.jn - g.j0 + 1) * (g.nk - g.k0 + 1), (S)0);
  s->pmu.assign(3L * (g.inu - g.i0u + 1) * (g.jn - g.j0 + 1) * (g.nk - g.k0 + 1), (S)0);
  s->pmv.assign(3L * (g.in - g.i0 + 1) * (g.jnv - g.j0v + 1) * (g.nk - g.k0 + 1), (S)0);
  s->pt.assign(3L * (g.in - g.i0 + 1) * (g.jn - g.j0 + 1) * (g.nk - g.k0t + 1), (S)0);
  s->pb.assign(3L * (g.inb - g.i0b + 1) * (g.jnb - g.j0b + 1) * g.nk, (S)0);
}

inline long wp_idx(const Grid& g, int i, int j, int k, int c) {  // Adz_wpxyz(i,j,k,c)
  return (long)(i + 1) + (long)(g.ni + 4) * ((j + 1) + (long)(g.nj + 4) * ((k - 1) + (long)g.nk * (c - 1)));
}
inline long p3_idx(int ni, int nj, int k0, int i, int j, int k) {  // (1:ni,1:nj,k0:..), comp 0
  return 3 * ((long)(i - 1) + (long)ni * ((j - 1) + (long)nj * (k - k0)));
}
inline long box_idx(int i0, int in, int j0, int jn, int k0, int i, int j, int k) {
  return 3 * ((long)(i - i0) + (long)(in - i0 + 1) * ((j - j0) + (long)(jn - j0 + 1) * (k - k0)));
}

// adz_inittraj.F90:16-42
template <class S>
inline void inittraj(const Grid& g, TrajState<S>* s) {
  for (int k = g.k0m; k <= g.nk; k++)
    for (int j = 1; j <= g.nj; j++)
      for (int i = 1; i <= g.ni; i++) {
        long p = p3_idx(g.ni, g.nj, g.k0m, i, j, k);
        s->pxyzm[p] = (S)(i + g.d.i0g - 1);
        s->wpxyz[wp_idx(g, i, j, k, 1)] = (double)s->pxyzm[p];
        s->pxyzm[p + 1] = (S)(j + g.d.j0g - 1);
        s->wpxyz[wp_idx(g, i, j, k, 2)] = (double)s->pxyzm[p + 1];
        s->pxyzm[p + 2] = (S)k;
        s->wpxyz[wp_idx(g, i, j, k, 3)] = (double)s->pxyzm[p + 2];
      }
}

// adz_prepareWinds.F90:16-107.  Inputs are FP32 model winds on the extended layout
// (lminx:lmaxx, lminy:lmaxy, nk): uu_ext, vv_ext, ww_ext (time t1, extended), and ut0,
// vt0 (staggered arrival winds) and zdt0 (arrival zeta-dot).
template <class S>
inline void prepare_winds(const Grid& g, const float* uu_ext, const float* vv_ext, const float* ww_ext,
                          const float* ut0, const float* vt0, const float* zdt0, TrajState<S>* s) {
  const int nk = g.nk;
  const double alpha1 = -1.0 / 16.0, alpha2 = 9.0 / 16.0;
  Ext3 e{g.lminx, g.lminy, g.nit, g.njt};
  std::vector<double> wx(nk + 1, 0.0);
  for (int k = 3; k <= nk; k++) wx[k] = g.vw1m[k];
  auto km2 = [](int k) { return std::max(1, k - 2); };
  const double ir = g.d.inv_rayt;
#pragma omp parallel for schedule(static)
  for (int k = 1; k <= nk; k++)
    for (int j = g.lminy; j <= g.lmaxy; j++)
      for (int i = g.lminx; i <= g.lmaxx; i++) {
        long q = e.idx(i, j, k);
        s->uvw_d[3 * q + 0] = (S)(ir * (double)uu_ext[q] * g.cy[j - g.lminy]);
        s->uvw_d[3 * q + 1] = (S)(ir * (double)vv_ext[q]);
      }
  const long nxy = (long)g.ni * g.nj;
#pragma omp parallel for schedule(static)
  for (int k = 1; k <= nk; k++)
    for (int j = 1; j <= g.nj; j++)
      for (int i = 1; i <= g.ni; i++) {
        long a = (i - 1) + (long)g.ni * (j - 1) + nxy * (k - 1);
        S u1 = (S)ut0[e.idx(i - 2, j, k)] + (S)ut0[e.idx(i + 1, j, k)];
        S u2 = (S)ut0[e.idx(i, j, k)] + (S)ut0[e.idx(i - 1, j, k)];
        s->uu_arr[a] = (S)((((double)u1 * alpha1 + (double)u2 * alpha2) * ir) * g.cy[j - g.lminy]);
        S v1 = (S)vt0[e.idx(i, j - 2, k)] + (S)vt0[e.idx(i, j + 1, k)];
        S v2 = (S)vt0[e.idx(i, j, k)] + (S)vt0[e.idx(i, j - 1, k)];
        s->vv_arr[a] = (S)(((double)v1 * alpha1 + (double)v2 * alpha2) * ir);
      }
#pragma omp parallel for schedule(static)
  for (int j = g.lminy; j <= g.lmaxy; j++)
    for (int i = g.lminx; i <= g.lmaxx; i++) {
      s->uvw_d[3 * e.idx(i, j, 1) + 2] = (S)(g.vw5 * (double)ww_ext[e.idx(i, j, 1)]);
      s->uvw_d[3 * e.idx(i, j, nk) + 2] = (S)(g.vw6 * (double)ww_ext[e.idx(i, j, nk - 1)]);
    }
#pragma omp parallel for schedule(static)
  for (int j = 1; j <= g.nj; j++)
    for (int i = 1; i <= g.ni; i++) {
      long a = (i - 1) + (long)g.ni * (j - 1);
      s->ww_arr[a] = (S)(g.vw5 * (double)zdt0[e.idx(i, j, 1)]);
      s->ww_arr[a + nxy * (nk - 1)] = (S)(g.vw6 * (double)zdt0[e.idx(i, j, nk - 1)]);
    }
#pragma omp parallel for schedule(static)
  for (int k = 2; k <= nk - 1; k++)
    for (int j = 1; j <= g.nj; j++)
      for (int i = 1; i <= g.ni; i++) {
        long a = (i - 1) + (long)g.ni * (j - 1) + nxy * (k - 1);
        double w = wx[k] * (double)zdt0[e.idx(i, j, km2(k))] + g.vw2m[k] * (double)zdt0[e.idx(i, j, k - 1)] +
                   g.vw3m[k] * (double)zdt0[e.idx(i, j, k)] + g.vw4m[k] * (double)zdt0[e.idx(i, j, k + 1)];
        s->ww_arr[a] = (S)w;
      }
#pragma omp parallel for schedule(static)
  for (int k = 2; k <= nk - 1; k++)
    for (int j = g.lminy; j <= g.lmaxy; j++)
      for (int i = g.lminx; i <= g.lmaxx; i++) {
        double w = wx[k] * (double)ww_ext[e.idx(i, j, km2(k))] + g.vw2m[k] * (double)ww_ext[e.idx(i, j, k - 1)] +
                   g.vw3m[k] * (double)ww_ext[e.idx(i, j, k)] + g.vw4m[k] * (double)ww_ext[e.idx(i, j, k + 1)];
        s->uvw_d[3 * e.idx(i, j, k) + 2] = (S)w;
      }
}

struct Dt { double dtA, dtzA, dtD, dtzD; };
// adz_traject.F90:38-59
inline Dt dt_split(double dt, int schm_advec, double bA_m, double bA) {
  Dt r{0, 0, 0, 0};
  if (schm_advec == 1) { r.dtA = dt * 0.5; r.dtzA = dt * 0.5; }
  if (schm_advec == 2) { r.dtA = dt * bA_m; r.dtzA = dt * bA; }
  if (schm_advec == 3) { r.dtA = (1.0 - bA_m) * dt; r.dtzA = (1.0 - bA) * dt; }
  r.dtD = dt - r.dtA; r.dtzD = dt - r.dtzA;
  if (schm_advec == 0) r = Dt{0, 0, 0, 0};
  return r;
}

// One departure-point update for point (i,j,k), adz_traject.F90:69-103.  Reads the
// current position from pxyzm, interpolates the three wind components there (zyx3) and
// writes uvw_dep, wpxyz and pxyzm.
template <class S>
inline void traj_point(const Grid& g, const Dt& t, int i, int j, int k, int order, TrajState<S>* s) {
  long p = p3_idx(g.ni, g.nj, g.k0m, i, j, k);
  double c[24]; int off;
  int zl = vcoef<S>(c, &off, s->pxyzm[p], s->pxyzm[p + 1], s->pxyzm[p + 2], g.ztab_q, order == ORD_AVX2 ? 1 : 0);  // -mfma build fuses the absolute z
  // f3 passed as Adz_uvw_d(1,1,1,1): offset of local (1,1,1) in the extended array
  long base = 3L * ((long)(1 - g.lminx) + (long)g.nit * (1 - g.lminy));
  double w3[3];
  tricub3<S>(s->uvw_d.data() + base + 3L * off, c, g.nit, g.nij, zl, order, w3);
  for (int q = 0; q < 3; q++) s->uvw_dep[p + q] = (S)w3[q];
  const long nxy = (long)g.ni * g.nj;
  long a = (i - 1) + (long)g.ni * (j - 1) + nxy * (k - 1);
  double ud = (double)s->uvw_dep[p], vd = (double)s->uvw_dep[p + 1], wd = (double)s->uvw_dep[p + 2];
  double xm = (double)(i + g.d.i0g - 1) - (t.dtD * ud + t.dtA * (double)s->uu_arr[a]) * g.d.inv_hx;
  double ym = (double)(j + g.d.j0g - 1) - (t.dtD * vd + t.dtA * (double)s->vv_arr[a]) * g.d.inv_hy;
  double pos = g.zm[k] - t.dtzD * wd - t.dtzA * (double)s->ww_arr[a];
  double zm = std::min(std::max(pos, g.d.zmin), g.d.zmax);
  int kk = level_search(g.zm, g.search_m, g.ovdzm, g.sig, g.nk, zm);
  s->wpxyz[wp_idx(g, i, j, k, 1)] = xm;
  s->pxyzm[p] = (S)std::min(std::max(xm, g.iminposx), g.imaxposx);
  s->wpxyz[wp_idx(g, i, j, k, 2)] = ym;
  s->pxyzm[p + 1] = (S)std::min(std::max(ym, g.iminposy), g.imaxposy);
  int kk1 = std::min(g.nk + 1, std::max(0, kk));
  int nb = std::max(std::min(kk1, g.nk - 1), 1);
  double pz = (zm - g.zm[nb]) * g.odelz_m[nb] + (double)nb;
  s->wpxyz[wp_idx(g, i, j, k, 3)] = pz;
  s->pxyzm[p + 2] = (S)pz;
}

inline double clampd(double v, double lo, double hi) { return std::min(std::max(v, lo), hi); }

// adz_interp_traj.F90:16-187 (Adz_slt_winds = .false.)
// klo..khi: the levels whose position sets are produced (1..nk: all); the departure points
// of levels klo-1..khi+2 must be current (the thermodynamic sets read levels k-1..k+2).
template <class S>
inline void interp_traj(const Grid& g, const Dt& t, double dt, const float* zdt0, TrajState<S>* s, int klo = 1,
                        int khi = 1 << 30) {
  (void)dt;
  auto band = [&](int k) { return k >= klo && k <= khi; };
  const int nk = g.nk;
  const double half = 0.5, aa = -0.0625, bb = 0.5625;
  const double zmin_b = 0.0, zmax_b = (double)(nk + 1);
  int k00 = std::max(g.k0t - 1, 1);
  if (g.d.bc_lam_flux != 0 && g.k0 > 1) k00 = 1;
  double ww = g.d.wmstar_nk;
  double wp = (g.zt[nk] - g.zm[nk - 1]) * g.d.idz_t_nkm1;
  double wm = 1.0 - wp;
  Ext3 e{g.lminx, g.lminy, g.nit, g.njt};
  auto W = [&](int i, int j, int k, int c) { return s->wpxyz[wp_idx(g, i, j, k, c)]; };
  auto UD3 = [&](int i, int j, int k) { return (double)s->uvw_dep[p3_idx(g.ni, g.nj, g.k0m, i, j, k) + 2]; };
  for (int k = std::max(g.k0, klo); k <= std::min(nk, khi); k++)
    for (int j = g.j0; j <= g.jn; j++)
      for (int i = g.i0u; i <= g.inu; i++) {
        double xt = aa * (W(i - 1, j, k, 1) + W(i + 2, j, k, 1)) + bb * (W(i, j, k, 1) + W(i + 1, j, k, 1)) - 0.5;
        double yt = aa * (W(i - 1, j, k, 2) + W(i + 2, j, k, 2)) + bb * (W(i, j, k, 2) + W(i + 1, j, k, 2));
        double zt = aa * (W(i - 1, j, k, 3) + W(i + 2, j, k, 3)) + bb * (W(i, j, k, 3) + W(i + 1, j, k, 3));
        long q = box_idx(g.i0u, g.inu, g.j0, g.jn, g.k0, i, j, k);
        s->pmu[q] = (S)clampd(xt, g.iminposx, g.imaxposx);
        s->pmu[q + 1] = (S)clampd(yt, g.iminposy, g.imaxposy);
        s->pmu[q + 2] = (S)clampd(zt, zmin_b, zmax_b);
      }
  for (int k = std::max(g.k0, klo); k <= std::min(nk, khi); k++)
    for (int j = g.j0v; j <= g.jnv; j++)
      for (int i = g.i0; i <= g.in; i++) {
        double xt = aa * (W(i, j - 1, k, 1) + W(i, j + 2, k, 1)) + bb * (W(i, j, k, 1) + W(i, j + 1, k, 1));
        double yt = aa * (W(i, j - 1, k, 2) + W(i, j + 2, k, 2)) + bb * (W(i, j, k, 2) + W(i, j + 1, k, 2)) - 0.5;
        double zt = aa * (W(i, j - 1, k, 3) + W(i, j + 2, k, 3)) + bb * (W(i, j, k, 3) + W(i, j + 1, k, 3));
        long q = box_idx(g.i0, g.in, g.j0v, g.jnv, g.k0, i, j, k);
        s->pmv[q] = (S)clampd(xt, g.iminposx, g.imaxposx);
        s->pmv[q + 1] = (S)clampd(yt, g.iminposy, g.imaxposy);
        s->pmv[q + 2] = (S)clampd(zt, zmin_b, zmax_b);
      }
  auto PB = [&](int i, int j, int k) { return box_idx(g.i0b, g.inb, g.j0b, g.jnb, 1, i, j, k); };
  auto Z0 = [&](int i, int j, int k) { return (double)zdt0[e.idx(i, j, k)]; };
  for (int k = std::max(std::max(k00, 2), klo); k <= std::min(nk - 2, khi); k++)
    for (int j = g.j0b; j <= g.jnb; j++)
      for (int i = g.i0b; i <= g.inb; i++) {
        double xt = g.vw1t[k] * W(i, j, k - 1, 1) + g.vw2t[k] * W(i, j, k, 1) + g.vw3t[k] * W(i, j, k + 1, 1) +
                    g.vw4t[k] * W(i, j, k + 2, 1);
        double yt = g.vw1t[k] * W(i, j, k - 1, 2) + g.vw2t[k] * W(i, j, k, 2) + g.vw3t[k] * W(i, j, k + 1, 2) +
                    g.vw4t[k] * W(i, j, k + 2, 2);
        double wdt = g.vw1t[k] * UD3(i, j, k - 1) + g.vw2t[k] * UD3(i, j, k) + g.vw3t[k] * UD3(i, j, k + 1) +
                     g.vw4t[k] * UD3(i, j, k + 2);
        double zt = g.zt[k] - t.dtzD * wdt - t.dtzA * Z0(i, j, k);
        long q = PB(i, j, k);
        s->pb[q] = (S)clampd(xt, g.iminposx, g.imaxposx);
        s->pb[q + 1] = (S)clampd(yt, g.iminposy, g.imaxposy);
        s->pb[q + 2] = (S)zindx(g, zt);
      }
  if (band(nk - 1)) {
    int k = nk - 1;
    for (int j = g.j0b; j <= g.jnb; j++)
      for (int i = g.i0b; i <= g.inb; i++) {
        double xt = (W(i, j, k, 1) + W(i, j, k + 1, 1)) * half;
        double yt = (W(i, j, k, 2) + W(i, j, k + 1, 2)) * half;
        double wdt = (UD3(i, j, k) + UD3(i, j, k + 1)) * half;
        double zt = g.zt[k] - t.dtzD * wdt - t.dtzA * Z0(i, j, k);
        long q = PB(i, j, k);
        s->pb[q] = (S)clampd(xt, g.iminposx, g.imaxposx);
        s->pb[q + 1] = (S)clampd(yt, g.iminposy, g.imaxposy);
        s->pb[q + 2] = (S)zindx(g, zt);
      }
  }
  if (band(nk)) {
    int k = nk;
    for (int j = g.j0b; j <= g.jnb; j++)
      for (int i = g.i0b; i <= g.inb; i++) {
        double xt = wp * W(i, j, k, 1) + wm * W(i, j, k - 1, 1);
        double yt = wp * W(i, j, k, 2) + wm * W(i, j, k - 1, 2);
        long q = PB(i, j, k);
        s->pb[q] = (S)clampd(xt, g.iminposx, g.imaxposx);
        s->pb[q + 1] = (S)clampd(yt, g.iminposy, g.imaxposy);
      }
    for (int j = g.j0b; j <= g.jnb; j++)
      for (int i = g.i0b; i <= g.inb; i++) {
        double wdt = (UD3(i, j, k - 1) + UD3(i, j, k)) * half;
        double zt = g.zt[k] - ww * (t.dtzD * wdt + t.dtzA * Z0(i, j, k - 1));
        s->pb[PB(i, j, k) + 2] = (S)zindx(g, zt);
      }
  }
  if (k00 == 1 && band(1)) {
    for (int j = g.j0b; j <= g.jnb; j++)
      for (int i = g.i0b; i <= g.inb; i++) {
        double xt = (W(i, j, 1, 1) + W(i, j, 2, 1)) * half;
        double yt = (W(i, j, 1, 2) + W(i, j, 2, 2)) * half;
        double wdt = (UD3(i, j, 1) + UD3(i, j, 2)) * half;
        double zt = g.zt[1] - t.dtzD * wdt - t.dtzA * Z0(i, j, 1);
        long q = PB(i, j, 1);
        s->pb[q] = (S)clampd(xt, g.iminposx, g.imaxposx);
        s->pb[q + 1] = (S)clampd(yt, g.iminposy, g.imaxposy);
        s->pb[q + 2] = (S)zindx(g, zt);
      }
  }
  for (int k = std::max(g.k0t, klo); k <= std::min(nk, khi); k++)
    for (int j = g.j0; j <= g.jn; j++)
      for (int i = g.i0; i <= g.in; i++) {
        long q = box_idx(g.i0, g.in, g.j0, g.jn, g.k0t, i, j, k), r = PB(i, j, k);
        for (int c = 0; c < 3; c++) s->pt[q + c] = s->pb[r + c];
      }
}

// adz_traject.F90:16-130 for one tile (the wpxyz halo exchange is a no-op on one tile
// of a LAM grid: the 2-point halo is never read, see DESIGN notes).
// klo..khi: only the position sets of these levels are produced, from the departure
// points of levels klo-1..khi+2 (each point's iteration reads only the winds and its own
// iterate, so a band of levels is exact); 1..nk is the whole tile.
template <class S>
inline void traject(const Grid& g, double dt, int schm_advec, double bA_m, double bA, int niter, int order,
                    const float* uu_ext, const float* vv_ext, const float* ww_ext, const float* ut0,
                    const float* vt0, const float* zdt0, TrajState<S>* s, int klo = 1, int khi = 1 << 30) {
  Dt t = dt_split(dt, schm_advec, bA_m, bA);
  prepare_winds<S>(g, uu_ext, vv_ext, ww_ext, ut0, vt0, zdt0, s);
  int k00 = g.k0m;
  if (g.k0 > 1) k00 = 1;
  const int ka = std::max(k00, klo - 1), kb = std::min(g.nk, khi + 2);
  // each arrival point only reads the static winds and its own iterate, so the points of
  // one iteration are independent (the reference shares them out per level, :67-68)
  for (int it = 1; it <= niter; it++) {
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = ka; k <= kb; k++)
      for (int j = 1; j <= g.nj; j++)
        for (int i = 1; i <= g.ni; i++) traj_point<S>(g, t, i, j, k, order, s);
  }
  for (int k = std::max(g.k0, klo); k <= std::min(g.nk, khi); k++)
    for (int j = g.j0; j <= g.jn; j++)
      for (int i = g.i0; i <= g.in; i++) {
        long q = box_idx(g.i0, g.in, g.j0, g.jn, g.k0, i, j, k), r = p3_idx(g.ni, g.nj, g.k0m, i, j, k);
        for (int c = 0; c < 3; c++) s->pm[q + c] = s->pxyzm[r + c];
      }
  interp_traj<S>(g, t, dt, zdt0, s, klo, khi);
}

// ------------------------------ Bermejo-Conde -------------------------------------

// almost_zero.F90:16-36: |bits| <= 1 after mapping negative patterns by
// two_complement - bits.  The Fortran constant int(Z'80000000', kind=8) is 2**31 (a
// 32-bit BOZ literal zero-padded to 64 bits), so the mapping wraps for negative doubles
// and no negative value (-0.0 included) is "almost zero": the function is true exactly
// for the bit patterns 0 (+0.0) and 1 (the smallest positive denormal).
inline bool almost_zero(double a) {
  uint64_t u;
  std::memcpy(&u, &a, 8);
  int64_t abit = (int64_t)u;
  if (abit < 0) abit = (int64_t)((uint64_t)0x80000000ULL - u);
  return abit >= -1 && abit <= 1;
}

// Exactly rounded sum (Shewchuk / fsum): the summation the HIP library reproduces with
// deterministic double-double partials.  The reference's sequential i-fastest order is
// available as mode 0 for comparison.
inline double fsum(const std::vector<double>& v) {
  std::vector<double> partials;
  for (double x : v) {
    size_t i = 0;
    for (double y : partials) {
      if (std::fabs(x) < std::fabs(y)) std::swap(x, y);
      double hi = x + y;
      double lo = y - (hi - x);
      if (lo != 0.0) partials[i++] = lo;
      x = hi;
    }
    partials.resize(i);
    partials.push_back(x);
  }
  // sum partials exactly rounded (Python math.fsum tail handling)
  int n = (int)partials.size();
  double hi = 0.0;
  if (n > 0) {
    hi = partials[--n];
    double lo = 0.0;
    while (n > 0) {
      double x = hi, y = partials[--n];
      hi = x + y;
      double yr = hi - x;
      lo = y - yr;
      if (lo != 0.0) break;
    }
    if (n > 0 && ((lo < 0.0 && partials[n - 1] < 0.0) || (lo > 0.0 && partials[n - 1] > 0.0))) {
      double y = lo * 2.0;
      double x = hi + y;
      double yr = x - hi;
      if (y == yr) hi = x;
    }
  }
  return hi + 0.0;  // a sum that starts from +0.0, as the Fortran accumulators do
}

// BC tracer description (type bc, adz_mem.F90:109-125); arrays on the (l_minx:l_maxx,
// l_miny:l_maxy, nk) model layout of storage type T.
template <class T>
struct BCTracer {
  const T* p; T* m; const T* lin; const T* mn; const T* mx;
  int legacy, weight, pexp_n;
  double mass_deficit;
};

struct BCDomain {
  int minx, maxx, miny, maxy, nk;   // array bounds
  int i0, in, j0, jn, k0;           // scope
  const double* mask;               // geomh_mask_8 (minx:maxx, miny:maxy)
  int min_max;                      // Adz_BC_min_max_L
  const float* pexp_list;           // Adz_BC_pexp_list (9)
  int sum_mode;                     // 0 sequential (mass_tr_*), 1 exactly rounded
  int autobar = 0;                  // Schm_autobar_L: `mask` is geomh_area_mask_8
};

template <class T, class A>
inline double mass_of(const BCDomain& D, const T* x, const A* air) {
  // mass_tr_X.F90:86-93: tracer*air_mass (Fortran real*real -> FP32 product when both are
  // real) then * geomh_mask_8.
  std::vector<double> terms;
  double acc = 0.0;
  if (D.autobar) {
    // mass_tr_X.F90:73-79 (and mass_tr_PM.F90:79-85): level 1 only, tracer (real, promoted)
    // times geomh_area_mask_8, no air mass
    for (int j = D.j0; j <= D.jn; j++)
      for (int i = D.i0; i <= D.in; i++) {
        long q2 = (long)(i - D.minx) + (long)(D.maxx - D.minx + 1) * (j - D.miny);
        double term = (double)x[q2] * D.mask[q2];
        if (D.sum_mode == 0) acc = acc + term; else terms.push_back(term);
      }
    return D.sum_mode == 0 ? acc : fsum(terms);
  }
  for (int k = D.k0; k <= D.nk; k++)
    for (int j = D.j0; j <= D.jn; j++)
      for (int i = D.i0; i <= D.in; i++) {
        long q = (long)(i - D.minx) + (long)(D.maxx - D.minx + 1) * ((j - D.miny) + (long)(D.maxy - D.miny + 1) * (k - 1));
        long q2 = (long)(i - D.minx) + (long)(D.maxx - D.minx + 1) * (j - D.miny);
        auto prod = x[q] * air[q];  // FP32*FP32 -> FP32 (reference); FP64 in O-64
        double term = (double)prod * D.mask[q2];
        if (D.sum_mode == 0) acc = acc + term; else terms.push_back(term);
      }
  return D.sum_mode == 0 ? acc : fsum(terms);
}

// Bermejo_Conde.F90:18-320 with reset_epsilon_L = .true.  flux: the masses (mass_fo,
// mass_fi) per tracer of the LAM Aranami flux (mass_tr_PM.F90:101-127), or null for zero
// fluxes.  w is computed into `w` (caller storage, same layout as m).
template <class T>
inline void bermejo_conde(const BCDomain& D, std::vector<BCTracer<T>>& tr, const T* airm1, const T* airm0,
                          const T* pkps, std::vector<std::vector<T>>& w, double* diag /* 6 per tracer or null */,
                          const double* flux = nullptr) {
  const int ntr = (int)tr.size();
  std::vector<double> mass_p(ntr), mass_m(ntr), mass_tot_p(ntr), mass_deficit(ntr), mass_wei(ntr);
  std::vector<int> cycle_tr(ntr, 0);
  const double ONE_8 = 1.0;
  auto sgn = [](double a, double b) { return std::signbit(b) ? -std::fabs(a) : std::fabs(a); };
  const long nx = D.maxx - D.minx + 1, ny = D.maxy - D.miny + 1;
  auto IDX = [&](int i, int j, int k) { return (long)(i - D.minx) + nx * ((j - D.miny) + ny * (k - 1)); };
  for (int n = 0; n < ntr; n++) {
    mass_p[n] = mass_of(D, tr[n].p, airm1);
    mass_m[n] = mass_of(D, (const T*)tr[n].m, airm0);
  }
  for (int n = 0; n < ntr; n++) {
    double p_exp = tr[n].legacy ? 1.0 : (double)D.pexp_list[tr[n].pexp_n - 1];
    double mass_bflux = flux ? flux[2 * n + 1] - flux[2 * n] : 0.0 - 0.0;  // :92
    mass_p[n] = mass_p[n] - tr[n].mass_deficit;
    mass_tot_p[n] = mass_p[n] + mass_bflux;
    double mass_tot_m = mass_m[n];
    mass_deficit[n] = mass_tot_m - mass_tot_p[n];
    w[n].assign(nx * ny * D.nk, (T)0);
    int sd = (int)sgn(ONE_8, mass_deficit[n]);
    for (int k = D.k0; k <= D.nk; k++)
      for (int j = D.j0; j <= D.jn; j++)
        for (int i = D.i0; i <= D.in; i++) {
          long q = IDX(i, j, k);
          w[n][q] = (T)0;
          T hml = tr[n].m[q] - tr[n].lin[q];  // FP32 subtraction in the reference
          double H = (double)hml;
          if (sd == (int)sgn(ONE_8, H)) {
            double v = std::pow(std::fabs(H), p_exp) * sgn(ONE_8, mass_deficit[n]);
            if (tr[n].weight == 2) v = v * std::max(0.0, (double)tr[n].m[q]);
            if (tr[n].weight == 3) v = v * (double)pkps[q];
            w[n][q] = (T)v;
          }
        }
  }
  for (int n = 0; n < ntr; n++) mass_wei[n] = mass_of(D, w[n].data(), airm0);
  for (int n = 0; n < ntr; n++) {
    if (almost_zero(mass_wei[n])) { cycle_tr[n] = 1; continue; }
    double lambda = mass_deficit[n] / mass_wei[n];
    for (int k = D.k0; k <= D.nk; k++)
      for (int j = D.j0; j <= D.jn; j++)
        for (int i = D.i0; i <= D.in; i++) {
          long q = IDX(i, j, k);
          double corr = lambda * (double)w[n][q];
          if (!D.min_max) { tr[n].m[q] = (T)((double)tr[n].m[q] - corr); continue; }
          T bc_wrk = (T)((double)tr[n].m[q] - corr);
          tr[n].m[q] = bc_wrk;
          if (corr > 0.0 && bc_wrk < tr[n].mn[q]) tr[n].m[q] = tr[n].mn[q];
          if (corr < 0.0 && bc_wrk > tr[n].mx[q]) tr[n].m[q] = tr[n].mx[q];
        }
  }
  if (D.min_max) {
    for (int n = 0; n < ntr; n++) {
      double mass_bc_r = mass_of(D, (const T*)tr[n].m, airm0);
      if (cycle_tr[n] == 1) continue;
      double proportion = mass_tot_p[n] / mass_bc_r;
      for (int k = D.k0; k <= D.nk; k++)
        for (int j = D.j0; j <= D.jn; j++)
          for (int i = D.i0; i <= D.in; i++) {
            long q = IDX(i, j, k);
            tr[n].m[q] = (T)(proportion * (double)tr[n].m[q]);
          }
    }
  }
  for (int n = 0; n < ntr; n++) {
    double mass_bc = mass_of(D, (const T*)tr[n].m, airm0);
    if (cycle_tr[n] != 1) tr[n].mass_deficit = mass_bc - mass_tot_p[n];
    if (diag) {
      diag[6 * n + 0] = mass_tot_p[n]; diag[6 * n + 1] = mass_m[n]; diag[6 * n + 2] = mass_wei[n];
      diag[6 * n + 3] = mass_bc; diag[6 * n + 4] = (double)cycle_tr[n]; diag[6 * n + 5] = tr[n].mass_deficit;
    }
  }
}

}  // namespace ora
