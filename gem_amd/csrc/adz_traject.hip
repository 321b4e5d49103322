// Departure-point trajectories on gfx950: adz_set (tables), adz_prepareWinds,
// adz_traject and adz_interp_traj of GEM's gemdyn (src/gemdyn/src/base/).
//
// The fixed-point iteration of adz_traject.F90:66-107 is independent per arrival point
// (each point reads the static wind field at its own current position), so one lane owns
// one arrival point for all iterations and keeps the iterate in registers; only the final
// iterate is stored.  The storage type S of the derived winds and positions is the
// reference's `real` (S = float, bit-exact with its trajectories) or FP64 (S = double).
// Arithmetic is FP64 everywhere, evaluated in the source's left-to-right order with no FMA
// contraction (built with -ffp-contract=off).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "adz_common.h"
#include "adz_qv.h"

using namespace adz;

namespace adz {

// Scalars and device tables a trajectory kernel needs (by value).
struct TrajParams {
  int ni, nj, nk, i0g, j0g;
  int lminx, lminy, nit, njt, nij;
  int i0, in, j0, jn, i0u, inu, j0v, jnv, i0b, inb, j0b, jnb, k0, k0t, k0m;
  int k00;   // first level of the iteration, adz_traject.F90:63-64
  int k00i;  // first level of the pb set, adz_interp_traj.F90:38-39
  double iminposx, imaxposx, iminposy, imaxposy;
  double sig, ovdzm, ovdzt, zmin, zmax, vw5, vw6, inv_hx, inv_hy, inv_rayt;
  double ww_star, wp_nk, wm_nk;  // Ver_wmstar_8(nk), wp, wm of adz_interp_traj.F90:42-44
  int pnz_m, pnz_t;
  const double* zm;      // 0..nk+1
  const double* zt;      // 0..nk+1
  const double* odelz_m; // 0..nk
  const double* odelz_t;
  const int* search_m;
  const int* search_t;
  const double* cy;      // index j - lminy
  const double* vw;      // 8 rows of nk+1: vw1m vw2m vw3m vw4m vw1t vw2t vw3t vw4t
};

__device__ __forceinline__ int d_level_search(const double* z, const int* tab, int pnz, double ovdz, double sig, int nk,
                                              double pos) {
  // adz_traject.F90:85-89; the two clamps on table indices never bind for positions
  // inside [zmin, zmax] and only keep a corrupted input from reading out of bounds
  int kk1 = (int)((pos - z[0]) * ovdz + 1.0);
  kk1 = min(max(1, kk1), pnz);
  kk1 = tab[kk1 - 1];
  if (sig * pos > sig * z[min(kk1 + 1, nk + 1)]) kk1 = kk1 + 1;
  if (sig * pos < sig * z[min(max(kk1, 0), nk + 1)]) kk1 = kk1 - 1;
  return kk1;
}

// zindx, adz_interp_traj.F90:189-204
__device__ __forceinline__ double d_zindx(const TrajParams& P, double posz) {
  posz = fmin(fmax(posz, P.zmin), P.zmax);
  int kk1 = d_level_search(P.zt, P.search_t, P.pnz_t, P.ovdzt, P.sig, P.nk, posz);
  kk1 = min(P.nk + 1, max(0, kk1));
  int nb = max(min(kk1, P.nk - 1), 1);
  return (posz - P.zt[nb]) * P.odelz_t[nb] + (double)nb;
}

__device__ __forceinline__ double clampd(double v, double lo, double hi) { return fmin(fmax(v, lo), hi); }

__device__ __forceinline__ long ext_idx(const TrajParams& P, int i, int j, int k) {
  return (long)(i - P.lminx) + (long)P.nit * ((j - P.lminy) + (long)P.njt * (k - 1));
}
__device__ __forceinline__ long wp_idx(const TrajParams& P, int i, int j, int k, int c) {
  return (long)(i + 1) + (long)(P.ni + 4) * ((j + 1) + (long)(P.nj + 4) * ((k - 1) + (long)P.nk * (c - 1)));
}
__device__ __forceinline__ long p3_idx(const TrajParams& P, int i, int j, int k) {  // (3, ni, nj, k0m:nk)
  return 3 * ((long)(i - 1) + (long)P.ni * ((j - 1) + (long)P.nj * (k - P.k0m)));
}
__device__ __forceinline__ long box_idx(int i0, int in, int j0, int jn, int k0, int i, int j, int k) {
  return 3 * ((long)(i - i0) + (long)(in - i0 + 1) * ((j - j0) + (long)(jn - j0 + 1) * (k - k0)));
}

// ---------------------------------------------------------------- adz_prepareWinds
// adz_prepareWinds.F90:37-47 and :62-101 on the extended domain.
template <class S>
__global__ void k_prep_uvwd(S* __restrict__ uvw_d, const float* __restrict__ uu, const float* __restrict__ vv,
                            const float* __restrict__ ww, TrajParams P) {
  long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  long n = (long)P.nij * P.nk;
  if (q >= n) return;
  int k = (int)(q / P.nij) + 1;
  int j = (int)((q % P.nij) / P.nit) + P.lminy;
  const double ir = P.inv_rayt;
  uvw_d[3 * q + 0] = (S)(ir * (double)uu[q] * P.cy[j - P.lminy]);
  uvw_d[3 * q + 1] = (S)(ir * (double)vv[q]);
  double w;
  const int nk = P.nk;
  if (k == 1) w = P.vw5 * (double)ww[q];
  else if (k == nk) w = P.vw6 * (double)ww[q - P.nij];
  else {
    const double* vw = P.vw;
    double wx = (k >= 3) ? vw[k] : 0.0;
    int km2 = max(1, k - 2);
    long base = q - (long)(k - 1) * P.nij;
    w = wx * (double)ww[base + (long)(km2 - 1) * P.nij] + vw[(nk + 1) + k] * (double)ww[q - P.nij] +
        vw[2 * (nk + 1) + k] * (double)ww[q] + vw[3 * (nk + 1) + k] * (double)ww[q + P.nij];
  }
  uvw_d[3 * q + 2] = (S)w;
}

// adz_prepareWinds.F90:48-61 and :69-86 on the interior: arrival winds.
template <class S>
__global__ void k_prep_arr(S* __restrict__ uarr, S* __restrict__ varr, S* __restrict__ warr,
                           const float* __restrict__ ut0, const float* __restrict__ vt0, const float* __restrict__ zdt0,
                           TrajParams P) {
  long a = (long)blockIdx.x * blockDim.x + threadIdx.x;
  long nxy = (long)P.ni * P.nj;
  if (a >= nxy * P.nk) return;
  int k = (int)(a / nxy) + 1;
  int j = (int)((a % nxy) / P.ni) + 1;
  int i = (int)(a % P.ni) + 1;
  const double alpha1 = -1.0 / 16.0, alpha2 = 9.0 / 16.0, ir = P.inv_rayt;
  S u1 = (S)ut0[ext_idx(P, i - 2, j, k)] + (S)ut0[ext_idx(P, i + 1, j, k)];
  S u2 = (S)ut0[ext_idx(P, i, j, k)] + (S)ut0[ext_idx(P, i - 1, j, k)];
  uarr[a] = (S)((((double)u1 * alpha1 + (double)u2 * alpha2) * ir) * P.cy[j - P.lminy]);
  S v1 = (S)vt0[ext_idx(P, i, j - 2, k)] + (S)vt0[ext_idx(P, i, j + 1, k)];
  S v2 = (S)vt0[ext_idx(P, i, j, k)] + (S)vt0[ext_idx(P, i, j - 1, k)];
  varr[a] = (S)(((double)v1 * alpha1 + (double)v2 * alpha2) * ir);
  const int nk = P.nk;
  double w;
  if (k == 1) w = P.vw5 * (double)zdt0[ext_idx(P, i, j, 1)];
  else if (k == nk) w = P.vw6 * (double)zdt0[ext_idx(P, i, j, nk - 1)];
  else {
    const double* vw = P.vw;
    double wx = (k >= 3) ? vw[k] : 0.0;
    int km2 = max(1, k - 2);
    w = wx * (double)zdt0[ext_idx(P, i, j, km2)] + vw[(nk + 1) + k] * (double)zdt0[ext_idx(P, i, j, k - 1)] +
        vw[2 * (nk + 1) + k] * (double)zdt0[ext_idx(P, i, j, k)] +
        vw[3 * (nk + 1) + k] * (double)zdt0[ext_idx(P, i, j, k + 1)];
  }
  warr[a] = (S)w;
}

// ---------------------------------------------------------------- adz_traject
// adz_traject.F90:66-104: niter fixed-point updates of one arrival point (i,j,k).
template <class S>
__global__ __launch_bounds__(256) void k_traject(S* __restrict__ pxyzm, double* __restrict__ wpxyz,
                                                  S* __restrict__ uvw_dep, const S* __restrict__ uvw_d,
                                                  const S* __restrict__ uarr, const S* __restrict__ varr,
                                                  const S* __restrict__ warr, ZtabDev lv, TrajParams P, int niter,
                                                  double dtA, double dtD, double dtzA, double dtzD) {
  // 32-bit point index (adz_set refuses tiles of 2^31 points or more)
  const unsigned a = blockIdx.x * blockDim.x + threadIdx.x;
  const unsigned nxy = (unsigned)P.ni * (unsigned)P.nj;
  if (a >= nxy * (unsigned)(P.nk - P.k00 + 1)) return;
  const unsigned kq = a / nxy, rem = a - kq * nxy, jq = rem / (unsigned)P.ni;
  const int k = (int)kq + P.k00;
  const int j = (int)jq + 1;
  const int i = (int)(rem - jq * (unsigned)P.ni) + 1;
  long p = p3_idx(P, i, j, k);
  long aw = (long)(i - 1) + (long)P.ni * (j - 1) + (long)nxy * (k - 1);
  S px = pxyzm[p], py = pxyzm[p + 1], pz = pxyzm[p + 2];
  S ud = (S)0, vd = (S)0, wd = (S)0;
  double xm = 0.0, ym = 0.0, pzm = 0.0;
  const double ua = (double)uarr[aw], va = (double)varr[aw], wa = (double)warr[aw];
  const S* f3 = uvw_d + 3L * ((long)(1 - P.lminx) + (long)P.nit * (1 - P.lminy));
  for (int it = 0; it < niter; it++) {
    Index x;
    double c[20], w3[3];
    if (index_step((double)px, (double)py, (double)pz, lv, &x)) {
      weights_ref(x, lv, c);
      tricub3_ref<S>(f3 + 3 * x.off, c, 3L * lv.ni, 3L * lv.nij, x.zlinear, w3);
    } else {
      w3[0] = w3[1] = w3[2] = __builtin_nan("");
    }
    ud = (S)w3[0]; vd = (S)w3[1]; wd = (S)w3[2];
    xm = (double)(i + P.i0g - 1) - (dtD * (double)ud + dtA * ua) * P.inv_hx;
    ym = (double)(j + P.j0g - 1) - (dtD * (double)vd + dtA * va) * P.inv_hy;
    double pos = P.zm[k] - dtzD * (double)wd - dtzA * wa;
    double zmv = fmin(fmax(pos, P.zmin), P.zmax);
    int kk = d_level_search(P.zm, P.search_m, P.pnz_m, P.ovdzm, P.sig, P.nk, zmv);
    int kk1 = min(P.nk + 1, max(0, kk));
    int nb = max(min(kk1, P.nk - 1), 1);
    pzm = (zmv - P.zm[nb]) * P.odelz_m[nb] + (double)nb;
    px = (S)clampd(xm, P.iminposx, P.imaxposx);
    py = (S)clampd(ym, P.iminposy, P.imaxposy);
    pz = (S)pzm;
  }
  pxyzm[p] = px; pxyzm[p + 1] = py; pxyzm[p + 2] = pz;
  uvw_dep[p] = ud; uvw_dep[p + 1] = vd; uvw_dep[p + 2] = wd;
  if (niter > 0) {
    wpxyz[wp_idx(P, i, j, k, 1)] = xm;
    wpxyz[wp_idx(P, i, j, k, 2)] = ym;
    wpxyz[wp_idx(P, i, j, k, 3)] = pzm;
  }
}

// adz_inittraj.F90:25-37
template <class S>
__global__ void k_inittraj(S* __restrict__ pxyzm, double* __restrict__ wpxyz, TrajParams P) {
  long a = (long)blockIdx.x * blockDim.x + threadIdx.x;
  long nxy = (long)P.ni * P.nj;
  if (a >= nxy * (P.nk - P.k0m + 1)) return;
  int k = (int)(a / nxy) + P.k0m;
  int j = (int)((a % nxy) / P.ni) + 1;
  int i = (int)(a % P.ni) + 1;
  long p = p3_idx(P, i, j, k);
  pxyzm[p] = (S)(i + P.i0g - 1);
  pxyzm[p + 1] = (S)(j + P.j0g - 1);
  pxyzm[p + 2] = (S)k;
  wpxyz[wp_idx(P, i, j, k, 1)] = (double)pxyzm[p];
  wpxyz[wp_idx(P, i, j, k, 2)] = (double)pxyzm[p + 1];
  wpxyz[wp_idx(P, i, j, k, 3)] = (double)pxyzm[p + 2];
}

// ---------------------------------------------------------------- adz_interp_traj
// u- and v-grid positions, adz_interp_traj.F90:45-78 (which = 0: u, 1: v)
template <class S>
__global__ void k_interp_uv(S* __restrict__ out, const double* __restrict__ W, TrajParams P, int which) {
  long a = (long)blockIdx.x * blockDim.x + threadIdx.x;
  int i0 = which == 0 ? P.i0u : P.i0, in = which == 0 ? P.inu : P.in;
  int j0 = which == 0 ? P.j0 : P.j0v, jn = which == 0 ? P.jn : P.jnv;
  long nx = in - i0 + 1, ny = jn - j0 + 1;
  if (nx <= 0 || ny <= 0 || a >= nx * ny * (P.nk - P.k0 + 1)) return;
  int k = (int)(a / (nx * ny)) + P.k0;
  int j = (int)((a % (nx * ny)) / nx) + j0;
  int i = (int)(a % nx) + i0;
  const double aa = -0.0625, bb = 0.5625;
  int di = which == 0 ? 1 : 0, dj = which == 0 ? 0 : 1;
  double r[3];
  for (int c = 1; c <= 3; c++) {
    double wm1 = W[wp_idx(P, i - di, j - dj, k, c)], w0 = W[wp_idx(P, i, j, k, c)];
    double w1 = W[wp_idx(P, i + di, j + dj, k, c)], w2 = W[wp_idx(P, i + 2 * di, j + 2 * dj, k, c)];
    r[c - 1] = aa * (wm1 + w2) + bb * (w0 + w1);
  }
  if (which == 0) r[0] = r[0] - 0.5; else r[1] = r[1] - 0.5;
  long q = box_idx(i0, in, j0, jn, P.k0, i, j, k);
  out[q] = (S)clampd(r[0], P.iminposx, P.imaxposx);
  out[q + 1] = (S)clampd(r[1], P.iminposy, P.imaxposy);
  out[q + 2] = (S)clampd(r[2], 0.0, (double)(P.nk + 1));
}

// thermodynamic-level positions Adz_pb, adz_interp_traj.F90:80-174
template <class S>
__global__ void k_interp_b(S* __restrict__ pb, const double* __restrict__ W, const S* __restrict__ uvw_dep,
                           const float* __restrict__ zdt0, TrajParams P, double dtzD, double dtzA) {
  long a = (long)blockIdx.x * blockDim.x + threadIdx.x;
  long nx = P.inb - P.i0b + 1, ny = P.jnb - P.j0b + 1;
  if (nx <= 0 || ny <= 0 || a >= nx * ny * P.nk) return;
  int k = (int)(a / (nx * ny)) + 1;
  int j = (int)((a % (nx * ny)) / nx) + P.j0b;
  int i = (int)(a % nx) + P.i0b;
  const int nk = P.nk;
  const double half = 0.5;
  int kstart = max(P.k00i, 2);
  long q = box_idx(P.i0b, P.inb, P.j0b, P.jnb, 1, i, j, k);
  auto UD3 = [&](int kk) { return (double)uvw_dep[p3_idx(P, i, j, kk) + 2]; };
  auto Z0 = [&](int kk) { return (double)zdt0[ext_idx(P, i, j, kk)]; };
  const double* vw = P.vw;
  if (k >= kstart && k <= nk - 2) {
    double v1 = vw[4 * (nk + 1) + k], v2 = vw[5 * (nk + 1) + k], v3 = vw[6 * (nk + 1) + k], v4 = vw[7 * (nk + 1) + k];
    double xt = v1 * W[wp_idx(P, i, j, k - 1, 1)] + v2 * W[wp_idx(P, i, j, k, 1)] + v3 * W[wp_idx(P, i, j, k + 1, 1)] +
                v4 * W[wp_idx(P, i, j, k + 2, 1)];
    double yt = v1 * W[wp_idx(P, i, j, k - 1, 2)] + v2 * W[wp_idx(P, i, j, k, 2)] + v3 * W[wp_idx(P, i, j, k + 1, 2)] +
                v4 * W[wp_idx(P, i, j, k + 2, 2)];
    double wdt = v1 * UD3(k - 1) + v2 * UD3(k) + v3 * UD3(k + 1) + v4 * UD3(k + 2);
    double zt = P.zt[k] - dtzD * wdt - dtzA * Z0(k);
    pb[q] = (S)clampd(xt, P.iminposx, P.imaxposx);
    pb[q + 1] = (S)clampd(yt, P.iminposy, P.imaxposy);
    pb[q + 2] = (S)d_zindx(P, zt);
  } else if (k == nk - 1) {
    double xt = (W[wp_idx(P, i, j, k, 1)] + W[wp_idx(P, i, j, k + 1, 1)]) * half;
    double yt = (W[wp_idx(P, i, j, k, 2)] + W[wp_idx(P, i, j, k + 1, 2)]) * half;
    double wdt = (UD3(k) + UD3(k + 1)) * half;
    double zt = P.zt[k] - dtzD * wdt - dtzA * Z0(k);
    pb[q] = (S)clampd(xt, P.iminposx, P.imaxposx);
    pb[q + 1] = (S)clampd(yt, P.iminposy, P.imaxposy);
    pb[q + 2] = (S)d_zindx(P, zt);
  } else if (k == nk) {
    double xt = P.wp_nk * W[wp_idx(P, i, j, k, 1)] + P.wm_nk * W[wp_idx(P, i, j, k - 1, 1)];
    double yt = P.wp_nk * W[wp_idx(P, i, j, k, 2)] + P.wm_nk * W[wp_idx(P, i, j, k - 1, 2)];
    pb[q] = (S)clampd(xt, P.iminposx, P.imaxposx);
    pb[q + 1] = (S)clampd(yt, P.iminposy, P.imaxposy);
    double wdt = (UD3(k - 1) + UD3(k)) * half;
    double zt = P.zt[k] - P.ww_star * (dtzD * wdt + dtzA * Z0(k - 1));
    pb[q + 2] = (S)d_zindx(P, zt);
  } else if (k == 1 && P.k00i == 1) {
    double xt = (W[wp_idx(P, i, j, 1, 1)] + W[wp_idx(P, i, j, 2, 1)]) * half;
    double yt = (W[wp_idx(P, i, j, 1, 2)] + W[wp_idx(P, i, j, 2, 2)]) * half;
    double wdt = (UD3(1) + UD3(2)) * half;
    double zt = P.zt[1] - dtzD * wdt - dtzA * Z0(1);
    pb[q] = (S)clampd(xt, P.iminposx, P.imaxposx);
    pb[q + 1] = (S)clampd(yt, P.iminposy, P.imaxposy);
    pb[q + 2] = (S)d_zindx(P, zt);
  }
}

// Adz_pm = Adz_pxyzm and Adz_pt = Adz_pb on [i0:in]x[j0:jn] (adz_traject.F90:112-116,
// adz_interp_traj.F90:176-181); which = 0: pm, 1: pt
template <class S>
__global__ void k_copy_box(S* __restrict__ out, const S* __restrict__ src, TrajParams P, int which) {
  long a = (long)blockIdx.x * blockDim.x + threadIdx.x;
  long nx = P.in - P.i0 + 1, ny = P.jn - P.j0 + 1;
  int kb = which == 0 ? P.k0 : P.k0t;
  if (nx <= 0 || ny <= 0 || a >= nx * ny * (P.nk - kb + 1)) return;
  int k = (int)(a / (nx * ny)) + kb;
  int j = (int)((a % (nx * ny)) / nx) + P.j0;
  int i = (int)(a % nx) + P.i0;
  long q = box_idx(P.i0, P.in, P.j0, P.jn, kb, i, j, k);
  long r = which == 0 ? p3_idx(P, i, j, k) : box_idx(P.i0b, P.inb, P.j0b, P.jnb, 1, i, j, k);
  out[q] = src[r]; out[q + 1] = src[r + 1]; out[q + 2] = src[r + 2];
}

}  // namespace adz

// ====================================================================== host side
struct adz_ctx {
  int pos_dtype, itraj, niter;
  TrajParams P;
  // host copies of the derived tables
  std::vector<double> zm, zt, zx, delz_m, delz_t, odelz_m, odelz_t, cy, vw;
  std::vector<int> search_m, search_t;
  adz_ztab* zq = nullptr;
  adz_ztab* zth = nullptr;
  void* tables = nullptr;  // device copy of the tables
  void* qvtab = nullptr;   // delz_t and the cubic / quintic denominators (BICHQV)
  void* amtab = nullptr;   // Ver_dz_8%t(k), Ver_idz_8%t(k), k = 0..nk (air mass)
  // device state
  void* uvw_d = nullptr; void* warr = nullptr; void* uvw_dep = nullptr; void* pxyzm = nullptr;
  double* wpxyz = nullptr;
  void* pm = nullptr; void* pmu = nullptr; void* pmv = nullptr; void* pt = nullptr; void* pb = nullptr;
  long n_uvw_d, n_arr, n_p3, n_wp, n_pm, n_pmu, n_pmv, n_pt, n_pb;
};

namespace {

// Level-search table: adz_set.F90:193-215 (REAL128 smallest_dz and bin starts as written)
void search_table(const std::vector<double>& z, const std::vector<double>& delz, int nk, double sig, double* ovdz,
                  std::vector<int>* tab) {
  __float128 smallest = (__float128)(sig * delz[1]);
  for (int k = 2; k <= nk; k++) {
    __float128 v = (__float128)(sig * delz[k]);
    if (v < smallest) smallest = v;
  }
  double o = (double)((__float128)1.0 / smallest);
  double pn = (double)1.0f + sig * (z[nk + 1] - z[0]) * o;
  int pnz = (int)std::lround(pn);
  *ovdz = (double)((__float128)sig / smallest);
  tab->assign(pnz, 0);
  int k0 = 0;
  for (int k = 1; k <= pnz; k++) {
    __float128 posz = (__float128)(sig * z[0]) + (__float128)(double)(k - 1) * smallest;
    if (posz > (__float128)(sig * z[k0 + 1])) k0 = std::min(nk + 1, k0 + 1);
    (*tab)[k - 1] = k0;
  }
}

double lag3(double zz, double z1, double z2, double z3, double z4) {
  return (((zz - z2) * (zz - z3) * (zz - z4)) / ((z1 - z2) * (z1 - z3) * (z1 - z4)));
}

unsigned nb(long n) { return (unsigned)((n + 255) / 256); }

void free_ctx(adz_ctx* c) {
  void* bufs[] = {c->tables, c->qvtab, c->amtab, c->uvw_d, c->warr, c->uvw_dep, c->pxyzm, c->wpxyz,
                  c->pm, c->pmu, c->pmv, c->pt, c->pb};
  for (void* b : bufs)
    if (b) (void)hipFree(b);
  if (c->zq) adz_vsearch_free(c->zq);
  if (c->zth) adz_vsearch_free(c->zth);
  delete c;
}

template <class S>
int run_traject(adz_ctx* c, double dt, int advec, double bA_m, double bA, const float* uu, const float* vv,
                const float* ww, const float* ut0, const float* vt0, const float* zdt0, hipStream_t st) {
  TrajParams& P = c->P;
  double dtA = 0, dtzA = 0;
  if (advec == 1) { dtA = dt * 0.5; dtzA = dt * 0.5; }
  if (advec == 2) { dtA = dt * bA_m; dtzA = dt * bA; }
  if (advec == 3) { dtA = (1.0 - bA_m) * dt; dtzA = (1.0 - bA) * dt; }
  double dtD = dt - dtA, dtzD = dt - dtzA;
  if (advec == 0) dtA = dtD = dtzA = dtzD = 0.0;
  long next = (long)P.nij * P.nk, nint = (long)P.ni * P.nj * P.nk;
  hipLaunchKernelGGL(k_prep_uvwd<S>, dim3(nb(next)), dim3(256), 0, st, (S*)c->uvw_d, uu, vv, ww, P);
  S* arr = (S*)c->warr;
  hipLaunchKernelGGL(k_prep_arr<S>, dim3(nb(nint)), dim3(256), 0, st, arr, arr + nint, arr + 2 * nint, ut0, vt0, zdt0, P);
  long npts = (long)P.ni * P.nj * (P.nk - P.k00 + 1);
  hipLaunchKernelGGL(k_traject<S>, dim3(nb(npts)), dim3(256), 0, st, (S*)c->pxyzm, c->wpxyz, (S*)c->uvw_dep,
                     (const S*)c->uvw_d, arr, arr + nint, arr + 2 * nint, c->zq->dev, P, c->niter, dtA, dtD, dtzA,
                     dtzD);
  hipLaunchKernelGGL(k_copy_box<S>, dim3(nb(c->n_pm / 3)), dim3(256), 0, st, (S*)c->pm, (const S*)c->pxyzm, P, 0);
  c->niter = c->itraj;  // adz_traject.F90:120
  hipLaunchKernelGGL(k_interp_uv<S>, dim3(nb(c->n_pmu / 3)), dim3(256), 0, st, (S*)c->pmu, c->wpxyz, P, 0);
  hipLaunchKernelGGL(k_interp_uv<S>, dim3(nb(c->n_pmv / 3)), dim3(256), 0, st, (S*)c->pmv, c->wpxyz, P, 1);
  hipLaunchKernelGGL(k_interp_b<S>, dim3(nb(c->n_pb / 3)), dim3(256), 0, st, (S*)c->pb, c->wpxyz, (const S*)c->uvw_dep,
                     zdt0, P, dtzD, dtzA);
  hipLaunchKernelGGL(k_copy_box<S>, dim3(nb(c->n_pt / 3)), dim3(256), 0, st, (S*)c->pt, (const S*)c->pb, P, 1);
  ADZ_HIP_CHECK(hipGetLastError());
  return ADZ_OK;
}

}  // namespace

extern "C" {

int adz_set(const adz_grid* g, int pos_dtype, int itraj, adz_ctx** out) {
  if (!g || !out || !g->zm || !g->zt || !g->zx || !g->lat) return fail(ADZ_EINVAL, "adz_set: null argument");
  if (pos_dtype != ADZ_F32 && pos_dtype != ADZ_F64) return fail(ADZ_EINVAL, "adz_set: pos_dtype must be F32 or F64");
  if (g->nk < 4 || g->ni < 8 || g->nj < 8 || g->halox < 2 || g->haloy < 2 || itraj < 1)
    return fail(ADZ_EINVAL, "adz_set: need nk>=4, ni,nj>=8, halo>=2, itraj>=1");
  if ((long)g->ni * g->nj * (g->nk + 1) >= (1L << 31))
    return fail(ADZ_EINVAL, "adz_set: a tile of 2^31 points or more (split the domain)");
  adz_ctx* c = new adz_ctx();
  c->pos_dtype = pos_dtype; c->itraj = itraj; c->niter = itraj;
  TrajParams& P = c->P;
  std::memset(&P, 0, sizeof(P));
  const int ni = g->ni, nj = g->nj, nk = g->nk;
  P.ni = ni; P.nj = nj; P.nk = nk; P.i0g = g->i0g; P.j0g = g->j0g;
  // adz_set.F90:65-172
  int lmaxx = ni + g->halox, lmaxy = nj + g->haloy;
  P.lminx = 1 - g->halox; P.lminy = 1 - g->haloy;
  P.nit = lmaxx - P.lminx + 1; P.njt = lmaxy - P.lminy + 1; P.nij = P.nit * P.njt;
  const int ext = 1;
  int pil_w = g->west ? g->pil_w : 0, pil_e = g->east ? g->pil_e : 0;
  int pil_s = g->south ? g->pil_s : 0, pil_n = g->north ? g->pil_n : 0;
  P.i0 = 1 + pil_w - (ext + 1) * g->west; P.in = ni - pil_e + ext * g->east;
  P.j0 = 1 + pil_s - (ext + 1) * g->south; P.jn = nj - pil_n + ext * g->north;
  P.i0u = 1 + pil_w; P.j0v = 1 + pil_s; P.inu = g->niu - pil_e; P.jnv = g->njv - pil_n;
  P.k0 = g->gbpil_t + 1; P.k0t = std::max(1, g->gbpil_t); P.k0m = std::max(P.k0t - 2, 1);
  if (g->bc_lam_flux != 0) P.k0m = 1;
  P.k00 = (P.k0 > 1) ? 1 : P.k0m;  // adz_traject.F90:63-64
  P.k00i = std::max(P.k0t - 1, 1);  // adz_interp_traj.F90:38-39
  if (g->bc_lam_flux != 0 && P.k0 > 1) P.k00i = 1;
  const double EPS_8 = 2.0e-5;
  const int BCS = 4;
  P.iminposx = g->i0g + P.lminx + EPS_8; P.imaxposx = g->i0g + lmaxx - 2 - EPS_8;
  P.iminposy = g->j0g + P.lminy + EPS_8; P.imaxposy = g->j0g + lmaxy - 2 - EPS_8;
  if (g->west) P.iminposx = g->i0g + BCS + EPS_8;
  if (g->east) P.imaxposx = g->i0g + ni - BCS - 1 - EPS_8;
  if (g->south) P.iminposy = g->j0g + BCS + EPS_8;
  if (g->north) P.imaxposy = g->j0g + nj - BCS - 1 - EPS_8;
  P.i0b = 1 + BCS * g->west; P.inb = ni - BCS * g->east; P.j0b = 1 + BCS * g->south; P.jnb = nj - BCS * g->north;
  if (P.i0 < 1 || P.in > ni || P.j0 < 1 || P.jn > nj || P.i0 < P.i0b || P.in > P.inb || P.j0 < P.j0b || P.jn > P.jnb) {
    delete c;
    return fail(ADZ_EINVAL, "adz_set: pilot widths too small for the interpolation boxes (need pil >= 6 on edges)");
  }
  if (P.in < P.i0 || P.jn < P.j0 || P.inu < P.i0u || P.jnv < P.j0v || P.inb < P.i0b || P.jnb < P.j0b ||
      P.k0 > nk || P.k0t > nk) {
    delete c;
    return fail(ADZ_EINVAL, "adz_set: tile too small for its pilot zones (an interpolation box is empty)");
  }
  c->zm.assign(g->zm, g->zm + nk + 2); c->zt.assign(g->zt, g->zt + nk + 2); c->zx.assign(g->zx, g->zx + nk + 1);
  const auto& zm = c->zm; const auto& zt = c->zt; const auto& zx = c->zx;
  P.sig = (zm[nk] - zm[1]) / std::fabs(zm[nk] - zm[1]);
  c->delz_m.assign(nk + 1, (double)9.e33f); c->delz_t.assign(nk + 1, (double)9.e33f);
  c->odelz_m.assign(nk + 1, 0.0); c->odelz_t.assign(nk + 1, 0.0);
  const float verysmall = 1.17549435e-38f * 1.e5f;
  for (int k = 0; k <= nk; k++) {
    if (std::fabs(zm[k + 1] - zm[k]) > (double)verysmall) c->delz_m[k] = zm[k + 1] - zm[k];
    if (std::fabs(zt[k + 1] - zt[k]) > (double)verysmall) c->delz_t[k] = zt[k + 1] - zt[k];
    c->odelz_m[k] = 1.0 / c->delz_m[k];
    c->odelz_t[k] = 1.0 / c->delz_t[k];
  }
  search_table(zm, c->delz_m, nk, P.sig, &P.ovdzm, &c->search_m);
  search_table(zt, c->delz_t, nk, P.sig, &P.ovdzt, &c->search_t);
  P.pnz_m = (int)c->search_m.size(); P.pnz_t = (int)c->search_t.size();
  c->cy.resize(P.njt);
  for (int j = 0; j < P.njt; j++) c->cy[j] = 1.0 / std::cos(g->lat[j]);
  c->vw.assign(8 * (nk + 1), 0.0);
  double* vw = c->vw.data();
  for (int k = 2; k <= nk - 1; k++) {
    double rx = zm[k], ra = zx[k - 2], rb = zx[k - 1], rc = zx[k], rd = zx[k + 1];
    vw[k] = lag3(rx, ra, rb, rc, rd);
    vw[(nk + 1) + k] = lag3(rx, rb, ra, rc, rd);
    vw[2 * (nk + 1) + k] = lag3(rx, rc, ra, rb, rd);
    vw[3 * (nk + 1) + k] = lag3(rx, rd, ra, rb, rc);
  }
  for (int k = 2; k <= nk - 2; k++) {
    double rx = zt[k], ra = zm[k - 1], rb = zm[k], rc = zm[k + 1], rd = zm[k + 2];
    vw[4 * (nk + 1) + k] = lag3(rx, ra, rb, rc, rd);
    vw[5 * (nk + 1) + k] = lag3(rx, rb, ra, rc, rd);
    vw[6 * (nk + 1) + k] = lag3(rx, rc, ra, rb, rd);
    vw[7 * (nk + 1) + k] = lag3(rx, rd, ra, rb, rc);
  }
  P.vw5 = (zx[0] - zm[1]) / (zx[0] - zx[1]);
  P.vw6 = (zm[nk] - zx[nk]) / (zx[nk - 1] - zx[nk]);
  P.zmin = g->zmin; P.zmax = g->zmax; P.inv_hx = g->inv_hx; P.inv_hy = g->inv_hy; P.inv_rayt = g->inv_rayt;
  P.ww_star = g->wmstar_nk;
  P.wp_nk = (zt[nk] - zm[nk - 1]) * g->idz_t_nkm1;
  P.wm_nk = 1.0 - P.wp_nk;

  // device tables
  size_t nd = (size_t)(nk + 2) * 2 + (size_t)(nk + 1) * 2 + P.njt + 8 * (nk + 1);
  size_t bytes = nd * 8 + (c->search_m.size() + c->search_t.size()) * 4;
  if (hipMalloc(&c->tables, bytes) != hipSuccess) { free_ctx(c); return fail(ADZ_ENOMEM, "adz_set: tables"); }
  std::vector<char> h(bytes);
  char* p = h.data();
  auto put = [&](const void* src, size_t b) { std::memcpy(p, src, b); p += b; };
  put(zm.data(), 8 * (nk + 2)); put(zt.data(), 8 * (nk + 2)); put(c->odelz_m.data(), 8 * (nk + 1));
  put(c->odelz_t.data(), 8 * (nk + 1)); put(c->cy.data(), 8 * P.njt); put(vw, 8 * 8 * (nk + 1));
  put(c->search_m.data(), 4 * c->search_m.size()); put(c->search_t.data(), 4 * c->search_t.size());
  if (hipMemcpy(c->tables, h.data(), bytes, hipMemcpyHostToDevice) != hipSuccess) {
    free_ctx(c);
    return fail(ADZ_EHIP, "adz_set: table upload");
  }
  char* d = (char*)c->tables;
  P.zm = (const double*)d; d += 8 * (nk + 2);
  P.zt = (const double*)d; d += 8 * (nk + 2);
  P.odelz_m = (const double*)d; d += 8 * (nk + 1);
  P.odelz_t = (const double*)d; d += 8 * (nk + 1);
  P.cy = (const double*)d; d += 8 * P.njt;
  P.vw = (const double*)d; d += 8 * 8 * (nk + 1);
  P.search_m = (const int*)d; d += 4 * c->search_m.size();
  P.search_t = (const int*)d;

  // BICHQV tables (adz_set.F90:286-307): delz_t, then the inverse cubic and quintic
  // Lagrange denominators of the thermodynamic levels, each indexed by k
  {
    std::vector<double> q(11 * (size_t)(nk + 1), 0.0);
    std::memcpy(q.data(), c->delz_t.data(), 8 * (nk + 1));
    double* c4 = q.data() + (nk + 1);
    double* c6 = c4 + 4 * (nk + 1);
    auto tri = [](double za, double zb, double zc, double zd) { return ((za - zb) * (za - zc) * (za - zd)); };
    auto qui = [](double zx, double za, double zb, double zc, double zd, double ze) {
      return ((zx - za) * (zx - zb) * (zx - zc) * (zx - zd) * (zx - ze));
    };
    for (int k = 2; k <= nk - 2; k++) {
      double ra = zt[k - 1], rb = zt[k], rc = zt[k + 1], rd = zt[k + 2];
      c4[k] = 1.0 / tri(ra, rb, rc, rd);
      c4[(nk + 1) + k] = 1.0 / tri(rb, ra, rc, rd);
      c4[2 * (nk + 1) + k] = 1.0 / tri(rc, ra, rb, rd);
      c4[3 * (nk + 1) + k] = 1.0 / tri(rd, ra, rb, rc);
    }
    for (int k = 3; k <= nk - 3; k++) {
      double rx = zt[k - 2], ra = zt[k - 1], rb = zt[k], rc = zt[k + 1], rd = zt[k + 2], re = zt[k + 3];
      c6[k] = 1.0 / qui(rx, ra, rb, rc, rd, re);
      c6[(nk + 1) + k] = 1.0 / qui(ra, rx, rb, rc, rd, re);
      c6[2 * (nk + 1) + k] = 1.0 / qui(rb, rx, ra, rc, rd, re);
      c6[3 * (nk + 1) + k] = 1.0 / qui(rc, rx, ra, rb, rd, re);
      c6[4 * (nk + 1) + k] = 1.0 / qui(rd, rx, ra, rb, rc, re);
      c6[5 * (nk + 1) + k] = 1.0 / qui(re, rx, ra, rb, rc, rd);
    }
    if (hipMalloc(&c->qvtab, 8 * q.size()) != hipSuccess) { free_ctx(c); return fail(ADZ_ENOMEM, "adz_set: tables"); }
    if (hipMemcpy(c->qvtab, q.data(), 8 * q.size(), hipMemcpyHostToDevice) != hipSuccess) {
      free_ctx(c);
      return fail(ADZ_EHIP, "adz_set: table upload");
    }
  }

  // layer thicknesses of the air mass (set_zeta.F90:271-275): Ver_dz_8%t(k) =
  // Ver_z_8%m(k+1) - Ver_z_8%m(k) and its inverse, k = 1..nk (entry 0 unused)
  {
    std::vector<double> a(2 * (size_t)(nk + 1), 0.0);
    for (int k = 1; k <= nk; k++) {
      a[k] = zm[k + 1] - zm[k];
      a[(nk + 1) + k] = 1.0 / a[k];
    }
    if (hipMalloc(&c->amtab, 8 * a.size()) != hipSuccess) { free_ctx(c); return fail(ADZ_ENOMEM, "adz_set: tables"); }
    if (hipMemcpy(c->amtab, a.data(), 8 * a.size(), hipMemcpyHostToDevice) != hipSuccess) {
      free_ctx(c);
      return fail(ADZ_EHIP, "adz_set: table upload");
    }
  }

  // vertical descriptors (adz_set.F90:128-131) over the extended arrays
  int st = adz_vsearch_setup_plus(&zm[1], nk, P.nit, P.njt, 1 - g->i0g, 1 - g->j0g, &c->zq);
  if (st == ADZ_OK) st = adz_vsearch_setup_plus(&zt[1], nk, P.nit, P.njt, 1 - g->i0g, 1 - g->j0g, &c->zth);
  if (st != ADZ_OK) { free_ctx(c); return st; }
  adz_vsearch_set_bounds(c->zq, P.lminx, lmaxx, P.lminy, lmaxy);
  adz_vsearch_set_bounds(c->zth, P.lminx, lmaxx, P.lminy, lmaxy);

  // device state (adz_set.F90:313-341, 346-360)
  const size_t sz = pos_dtype == ADZ_F32 ? 4 : 8;
  long nxy = (long)ni * nj;
  c->n_uvw_d = 3L * P.nij * nk; c->n_arr = 3 * nxy * nk; c->n_p3 = 3 * nxy * (nk - P.k0m + 1);
  c->n_wp = (long)(ni + 4) * (nj + 4) * nk * 3;
  c->n_pm = 3L * (P.in - P.i0 + 1) * (P.jn - P.j0 + 1) * (nk - P.k0 + 1);
  c->n_pmu = 3L * (P.inu - P.i0u + 1) * (P.jn - P.j0 + 1) * (nk - P.k0 + 1);
  c->n_pmv = 3L * (P.in - P.i0 + 1) * (P.jnv - P.j0v + 1) * (nk - P.k0 + 1);
  c->n_pt = 3L * (P.in - P.i0 + 1) * (P.jn - P.j0 + 1) * (nk - P.k0t + 1);
  c->n_pb = 3L * (P.inb - P.i0b + 1) * (P.jnb - P.j0b + 1) * nk;
  struct { void** p; size_t b; } al[] = {
      {&c->uvw_d, c->n_uvw_d * sz}, {&c->warr, c->n_arr * sz}, {&c->uvw_dep, c->n_p3 * sz},
      {&c->pxyzm, c->n_p3 * sz},    {(void**)&c->wpxyz, (size_t)c->n_wp * 8}, {&c->pm, c->n_pm * sz},
      {&c->pmu, c->n_pmu * sz},     {&c->pmv, c->n_pmv * sz}, {&c->pt, c->n_pt * sz}, {&c->pb, c->n_pb * sz}};
  for (auto& a : al) {
    if (hipMalloc(a.p, a.b) != hipSuccess) { free_ctx(c); return fail(ADZ_ENOMEM, "adz_set: state allocation"); }
    if (hipMemset(*a.p, 0, a.b) != hipSuccess) { free_ctx(c); return fail(ADZ_EHIP, "adz_set: state clear"); }
  }
  *out = c;
  return ADZ_OK;
}

int adz_ctx_free(adz_ctx* c) {
  if (c) free_ctx(c);
  return ADZ_OK;
}

int adz_ctx_info(const adz_ctx* c, int* iv, double* dv) {
  if (!c || !iv || !dv) return fail(ADZ_EINVAL, "adz_ctx_info: null argument");
  const TrajParams& P = c->P;
  int a[] = {P.lminx, P.ni + (P.nit - P.ni) / 2, P.lminy, P.nj + (P.njt - P.nj) / 2, P.nit, P.njt, P.nij,
             P.i0, P.in, P.j0, P.jn, P.i0u, P.inu, P.j0v, P.jnv, P.i0b, P.inb, P.j0b, P.jnb, P.k0, P.k0t, P.k0m,
             P.pnz_m, P.pnz_t};
  std::memcpy(iv, a, sizeof(a));
  double b[] = {P.iminposx, P.imaxposx, P.iminposy, P.imaxposy, P.sig, P.ovdzm, P.ovdzt, P.vw5, P.vw6};
  std::memcpy(dv, b, sizeof(b));
  return ADZ_OK;
}

int adz_ctx_table(const adz_ctx* c, int which, void* out) {
  if (!c || !out || which < 0 || which > 12) return fail(ADZ_EINVAL, "adz_ctx_table: bad argument");
  int nk = c->P.nk;
  if (which == 0) std::memcpy(out, c->search_m.data(), 4 * c->search_m.size());
  else if (which == 1) std::memcpy(out, c->search_t.data(), 4 * c->search_t.size());
  else if (which <= 9) std::memcpy(out, c->vw.data() + (which - 2) * (nk + 1), 8 * (nk + 1));
  else if (which == 10) std::memcpy(out, c->cy.data(), 8 * c->cy.size());
  else if (which == 11) std::memcpy(out, c->odelz_m.data(), 8 * (nk + 1));
  else std::memcpy(out, c->odelz_t.data(), 8 * (nk + 1));
  return ADZ_OK;
}

int adz_ctx_array(adz_ctx* c, int which, void** dptr, long* nelem) {
  if (!c || !dptr) return fail(ADZ_EINVAL, "adz_ctx_array: null argument");
  void* p[] = {c->pxyzm, c->wpxyz, c->uvw_dep, c->pm, c->pmu, c->pmv, c->pt, c->pb, c->uvw_d, c->warr};
  long n[] = {c->n_p3, c->n_wp, c->n_p3, c->n_pm, c->n_pmu, c->n_pmv, c->n_pt, c->n_pb, c->n_uvw_d, c->n_arr};
  if (which < 0 || which > 9) return fail(ADZ_EINVAL, "adz_ctx_array: unknown array");
  *dptr = p[which];
  if (nelem) *nelem = n[which];
  return ADZ_OK;
}

const adz_ztab* adz_ctx_ztab(const adz_ctx* c, int thermo) { return c ? (thermo ? c->zth : c->zq) : nullptr; }

}  // extern "C"

bool adz::ctx_qv(const adz_ctx* c, QvDev* q) {
  if (!c || !c->qvtab) return false;
  const int nk = c->P.nk;
  q->posz = c->P.zt;
  q->odz = c->P.odelz_t;
  q->dz = (const double*)c->qvtab;
  q->c4 = q->dz + (nk + 1);
  q->c6 = q->c4 + 4 * (nk + 1);
  q->nk = nk;
  q->offi = 1 - c->P.i0g;
  q->offj = 1 - c->P.j0g;
  q->lminx = c->P.lminx; q->lmaxx = c->P.lminx + c->P.nit - 1;
  q->lminy = c->P.lminy; q->lmaxy = c->P.lminy + c->P.njt - 1;
  q->pos_dtype = c->pos_dtype;
  return true;
}

bool adz::ctx_airmass(const adz_ctx* c, AirmassTables* t) {
  if (!c || !c->amtab) return false;
  const int nk = c->P.nk;
  t->dz = (const double*)c->amtab;
  t->idz = t->dz + (nk + 1);
  t->nk = nk; t->ni = c->P.ni; t->nj = c->P.nj; t->k0t = c->P.k0t;
  return true;
}

extern "C" {

// ---------------------------------------------------------------- restart of the SL state
// What the reference checkpoints for the semi-Lagrangian scheme: ADZ_PXYZM and ADZ_WPXYZ,
// registered with GMM_FLAG_RSTR (adz_set.F90:343-360).  The blob is a fixed header, then
// Adz_pxyzm in the context's position storage, then Adz_wpxyz in FP64.  The iteration
// count of the next adz_traject is kept too (the reference restarts at a step boundary,
// where it is Adz_itraj: adz_traject.F90:120, adz_set.F90:424).
namespace {
struct StateHeader {
  char magic[8];
  int32_t pos_dtype, ni, nj, nk, k0m, niter;
  int64_t n_p3, n_wp;
};
constexpr char kStateMagic[8] = {'A', 'D', 'Z', 'S', 'T', 'A', 'T', '1'};

size_t state_bytes(const adz_ctx* c) {
  const size_t sz = c->pos_dtype == ADZ_F32 ? 4 : 8;
  return sizeof(StateHeader) + (size_t)c->n_p3 * sz + (size_t)c->n_wp * 8;
}
}  // namespace

int adz_ctx_state_bytes(const adz_ctx* c, size_t* bytes) {
  if (!c || !bytes) return fail(ADZ_EINVAL, "adz_ctx_state_bytes: null argument");
  *bytes = state_bytes(c);
  return ADZ_OK;
}

int adz_ctx_save_state(const adz_ctx* c, void* host, size_t bytes, adz_stream s) {
  if (!c || !host) return fail(ADZ_EINVAL, "adz_ctx_save_state: null argument");
  if (bytes < state_bytes(c)) return fail(ADZ_EINVAL, "adz_ctx_save_state: buffer too small");
  StateHeader h;
  std::memcpy(h.magic, kStateMagic, 8);
  h.pos_dtype = c->pos_dtype; h.ni = c->P.ni; h.nj = c->P.nj; h.nk = c->P.nk; h.k0m = c->P.k0m;
  h.niter = c->niter; h.n_p3 = c->n_p3; h.n_wp = c->n_wp;
  std::memcpy(host, &h, sizeof(h));
  char* p = (char*)host + sizeof(h);
  const size_t b3 = (size_t)c->n_p3 * (c->pos_dtype == ADZ_F32 ? 4 : 8);
  ADZ_HIP_CHECK(hipMemcpyAsync(p, c->pxyzm, b3, hipMemcpyDeviceToHost, (hipStream_t)s));
  ADZ_HIP_CHECK(hipMemcpyAsync(p + b3, c->wpxyz, (size_t)c->n_wp * 8, hipMemcpyDeviceToHost, (hipStream_t)s));
  ADZ_HIP_CHECK(hipStreamSynchronize((hipStream_t)s));
  return ADZ_OK;
}

int adz_ctx_load_state(adz_ctx* c, const void* host, size_t bytes, adz_stream s) {
  if (!c || !host) return fail(ADZ_EINVAL, "adz_ctx_load_state: null argument");
  if (bytes < sizeof(StateHeader)) return fail(ADZ_EINVAL, "adz_ctx_load_state: buffer too small");
  StateHeader h;
  std::memcpy(&h, host, sizeof(h));
  if (std::memcmp(h.magic, kStateMagic, 8) != 0) return fail(ADZ_EINVAL, "adz_ctx_load_state: not an SL state");
  if (h.pos_dtype != c->pos_dtype || h.ni != c->P.ni || h.nj != c->P.nj || h.nk != c->P.nk || h.k0m != c->P.k0m ||
      h.n_p3 != c->n_p3 || h.n_wp != c->n_wp)
    return fail(ADZ_EINVAL, "adz_ctx_load_state: the state belongs to another grid or position storage");
  if (bytes < state_bytes(c)) return fail(ADZ_EINVAL, "adz_ctx_load_state: buffer too small");
  if (h.niter < 1) return fail(ADZ_EINVAL, "adz_ctx_load_state: bad iteration count");
  const char* p = (const char*)host + sizeof(h);
  const size_t b3 = (size_t)c->n_p3 * (c->pos_dtype == ADZ_F32 ? 4 : 8);
  ADZ_HIP_CHECK(hipMemcpyAsync(c->pxyzm, p, b3, hipMemcpyHostToDevice, (hipStream_t)s));
  ADZ_HIP_CHECK(hipMemcpyAsync(c->wpxyz, p + b3, (size_t)c->n_wp * 8, hipMemcpyHostToDevice, (hipStream_t)s));
  ADZ_HIP_CHECK(hipStreamSynchronize((hipStream_t)s));
  c->niter = h.niter;
  return ADZ_OK;
}

int adz_inittraj(adz_ctx* c, adz_stream s) {
  if (!c) return fail(ADZ_EINVAL, "adz_inittraj: null context");
  long n = (long)c->P.ni * c->P.nj * (c->P.nk - c->P.k0m + 1);
  if (c->pos_dtype == ADZ_F32)
    hipLaunchKernelGGL(k_inittraj<float>, dim3(nb(n)), dim3(256), 0, (hipStream_t)s, (float*)c->pxyzm, c->wpxyz, c->P);
  else
    hipLaunchKernelGGL(k_inittraj<double>, dim3(nb(n)), dim3(256), 0, (hipStream_t)s, (double*)c->pxyzm, c->wpxyz, c->P);
  ADZ_HIP_CHECK(hipGetLastError());
  c->niter = std::max(6, c->itraj);
  return ADZ_OK;
}

int adz_traject(adz_ctx* c, double dt, int schm_advec, double bA_m, double bA, const float* uu_ext,
                const float* vv_ext, const float* ww_ext, const float* ut0, const float* vt0, const float* zdt0,
                adz_stream s) {
  if (!c || !uu_ext || !vv_ext || !ww_ext || !ut0 || !vt0 || !zdt0) return fail(ADZ_EINVAL, "adz_traject: null argument");
  if (schm_advec < 0 || schm_advec > 3) return fail(ADZ_EINVAL, "adz_traject: Schm_advec must be 0..3");
  if (c->pos_dtype == ADZ_F32)
    return run_traject<float>(c, dt, schm_advec, bA_m, bA, uu_ext, vv_ext, ww_ext, ut0, vt0, zdt0, (hipStream_t)s);
  return run_traject<double>(c, dt, schm_advec, bA_m, bA, uu_ext, vv_ext, ww_ext, ut0, vt0, zdt0, (hipStream_t)s);
}

}  // extern "C"
