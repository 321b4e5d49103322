// Air masses of the mass fixer on gfx950: get_air_mass_hlt (get_air_mass_hlt.F90:18-103) at
// time P (F_time = 1) and time M (F_time = 0), and pr_t/pr_s at time M, as set_post_tr_hlt
// computes them at the start of every tracer step (set_post_tr_hlt.F90:44-70).
//
// One thread per (i, j) column of the array layout walks the levels: pr_m(k+1) read at
// level k is pr_m(k) of level k+1, and the column's area and p0 are read once, so every
// input element is read once and every output element written once (HBM streaming).
// Per interior element and time: pr_m, the humidity and the hydrometeors when they count,
// out: the air mass; pkps reads pr_t.  Columns outside 1..l_ni x 1..l_nj and levels below
// k0 are set to zero, as the reference does (:84-91).
//
// Arithmetic in the reference's order: sumq accumulates in FP64 (sumhydro_hlt.F90:35-60
// starts from 0 and adds each hydrometeor in tracer order, then the humidity,
// get_air_mass_hlt.F90:67-79); the pressure difference is taken in the storage type (the
// reference's `real` pressures subtract in FP32); dens = diff * (1 - sumq) * idz * (1/g),
// air mass = dens * area * dz, left to right in FP64, rounded to the storage type.  No
// multiply-add is contracted (-ffp-contract=off); none is written either.
#include <algorithm>
#include <cstring>

#include "adz_common.h"

namespace adz {

constexpr int kMaxHyd = 8;  // hydrometeor tracers summed into sumq (per time)

struct AmArgs {
  int minx, miny, nxa, nya, nk, ni, nj, k0;
  long nij;                         // plane pitch of the 3-D arrays
  const void* pm[2];                // pr_m at P, M (nk+1 levels)
  const void* hu[2];                // humidity at P, M (or null)
  const void* hyd[2][kMaxHyd];      // hydrometeors at P, M
  int nhyd;
  const void* pt;                   // pr_t at M
  const void* p0;                   // p0 at M (2-D)
  const double* area;               // geomh_area_8 (2-D)
  const double* dz;                 // Ver_dz_8%t(k), k = 0..nk
  const double* idz;                // Ver_idz_8%t(k)
  double inv_grav;
  void* out[2];                     // airm1, airm0
  void* pkps;
};

template <class T> __device__ __forceinline__ T ratio(T a, T b);
// FP32: the quotient formed in FP64 and rounded once is the correctly rounded FP32
// quotient (53 >= 2*24 + 2 bits), i.e. the IEEE division the reference's `real` does
template <> __device__ __forceinline__ float ratio<float>(float a, float b) { return (float)((double)a / (double)b); }
template <> __device__ __forceinline__ double ratio<double>(double a, double b) { return a / b; }

template <class T>
__global__ __launch_bounds__(256) void k_airmass(AmArgs A) {
  const long col = (long)blockIdx.x * 256 + threadIdx.x;
  if (col >= (long)A.nxa * A.nya) return;
  const int i = A.minx + (int)(col % A.nxa), j = A.miny + (int)(col / A.nxa);
  const long o = (long)(i - 1) + (long)A.nxa * (j - 1);  // element (i, j, 1) from (1, 1, 1)
  typedef ADZ_GLOBAL T* GT;
  typedef const ADZ_GLOBAL T* CGT;
  GT out0 = (GT)A.out[0], out1 = (GT)A.out[1];
  const bool inner = i >= 1 && i <= A.ni && j >= 1 && j <= A.nj;
  if (!inner) {  // get_air_mass_hlt.F90:84-91: zero outside the tile
    for (int k = 1; k <= A.nk; k++) {
      const long q = o + (long)(k - 1) * A.nij;
      out0[q] = (T)0;
      out1[q] = (T)0;
    }
    return;
  }
  for (int k = 1; k < A.k0; k++) {
    const long q = o + (long)(k - 1) * A.nij;
    out0[q] = (T)0;
    out1[q] = (T)0;
  }
  const double area = ((const ADZ_GLOBAL double*)A.area)[o];
  const ADZ_GLOBAL double* dz = (const ADZ_GLOBAL double*)A.dz;
  const ADZ_GLOBAL double* idz = (const ADZ_GLOBAL double*)A.idz;
  T prev[2];
#pragma unroll
  for (int t = 0; t < 2; t++) prev[t] = ((CGT)A.pm[t])[o + (long)(A.k0 - 1) * A.nij];
  for (int k = A.k0; k <= A.nk; k++) {
    const long q = o + (long)(k - 1) * A.nij;
#pragma unroll
    for (int t = 0; t < 2; t++) {
      const T next = ((CGT)A.pm[t])[q + A.nij];  // pr_m(i,j,k+1)
      double sumq = 0.0;
      for (int n = 0; n < A.nhyd; n++) sumq = sumq + (double)((CGT)A.hyd[t][n])[q];
      if (A.hu[t]) sumq = sumq + (double)((CGT)A.hu[t])[q];
      const T diff = next - prev[t];
      const double dens = (double)diff * (1.0 - sumq) * idz[k] * A.inv_grav;
      ((GT)A.out[t])[q] = (T)(dens * area * dz[k]);
      prev[t] = next;
    }
  }
  // set_post_tr_hlt.F90:62-70: pkps = pr_t / p0 at time M on the tile, every level
  if (A.pkps) {
    const T p0 = ((CGT)A.p0)[o];
    for (int k = 1; k <= A.nk; k++) {
      const long q = o + (long)(k - 1) * A.nij;
      ((GT)A.pkps)[q] = ratio<T>(((CGT)A.pt)[q], p0);
    }
  }
}

}  // namespace adz

using namespace adz;

extern "C" {

int adz_set_post_tr(const adz_ctx* c, const adz_airmass_fields* f, double grav, void* airm1, void* airm0, void* pkps,
                    int dtype, adz_stream stream) {
  if (!c || !f || !airm1 || !airm0) return fail(ADZ_EINVAL, "adz_set_post_tr: null argument");
  if (dtype != ADZ_F32 && dtype != ADZ_F64) return fail(ADZ_EINVAL, "adz_set_post_tr: dtype must be F32 or F64");
  AirmassTables tb;
  if (!ctx_airmass(c, &tb)) return fail(ADZ_EINVAL, "adz_set_post_tr: context without air-mass tables");
  if (!f->pm_p || !f->pm_m || (pkps && (!f->pt_m || !f->p0_m)) || !f->area)
    return fail(ADZ_EINVAL, "adz_set_post_tr: null input array");
  if (f->maxx - f->minx + 1 <= 0 || f->maxy - f->miny + 1 <= 0 || f->minx > 1 || f->miny > 1 || f->maxx < tb.ni ||
      f->maxy < tb.nj)
    return fail(ADZ_EINVAL, "adz_set_post_tr: layout must hold the tile 1..l_ni x 1..l_nj");
  if (!(grav > 0.0)) return fail(ADZ_EINVAL, "adz_set_post_tr: gravity must be positive");
  const bool dry = f->dry_mixing_ratio != 0, hyd = dry && f->wload != 0;  // get_air_mass_hlt.F90:67-68
  if (dry && (!f->hu_p || !f->hu_m)) return fail(ADZ_EINVAL, "adz_set_post_tr: dry mixing ratio needs the humidity");
  if (hyd && (f->nhyd < 0 || f->nhyd > kMaxHyd || (f->nhyd > 0 && (!f->hyd_p || !f->hyd_m))))
    return fail(ADZ_EINVAL, "adz_set_post_tr: 0..8 hydrometeor tracers per time");
  AmArgs a;
  std::memset(&a, 0, sizeof(a));
  a.minx = f->minx; a.miny = f->miny;
  a.nxa = f->maxx - f->minx + 1; a.nya = f->maxy - f->miny + 1;
  a.nk = tb.nk; a.ni = tb.ni; a.nj = tb.nj; a.k0 = tb.k0t;
  a.nij = (long)a.nxa * a.nya;
  a.pm[0] = f->pm_p; a.pm[1] = f->pm_m;
  if (dry) { a.hu[0] = f->hu_p; a.hu[1] = f->hu_m; }
  if (hyd) {
    a.nhyd = f->nhyd;
    for (int n = 0; n < f->nhyd; n++) {
      if (!f->hyd_p[n] || !f->hyd_m[n]) return fail(ADZ_EINVAL, "adz_set_post_tr: null hydrometeor array");
      a.hyd[0][n] = f->hyd_p[n];
      a.hyd[1][n] = f->hyd_m[n];
    }
  }
  a.pt = f->pt_m; a.p0 = f->p0_m; a.area = f->area;
  a.dz = tb.dz; a.idz = tb.idz;
  a.inv_grav = 1.0 / grav;  // get_air_mass_hlt.F90:83
  a.out[0] = airm1; a.out[1] = airm0; a.pkps = pkps;
  const long cols = (long)a.nxa * a.nya;
  dim3 g((unsigned)((cols + 255) / 256)), b(256);
  if (dtype == ADZ_F32) hipLaunchKernelGGL((k_airmass<float>), g, b, 0, (hipStream_t)stream, a);
  else hipLaunchKernelGGL((k_airmass<double>), g, b, 0, (hipStream_t)stream, a);
  ADZ_HIP_CHECK(hipGetLastError());
  return ADZ_OK;
}

}  // extern "C"
