// Bicubic-horizontal / quintic-vertical interpolation with the linear result, min and max
// of the inner box (the BICHQV tracer groups) on gfx950: adz_tricub_hlt with F_QV_L
// (adz_interp_hlt.F90:59-66,113-115) through bicubHQV_lin_min_max_ntr
// (bicubHQV_lin_mono.F90:20-172) and bicubHQV_lin_min_max_zyx
// (bicubHQV_lin_min_max_zyx.F90:20-99).
//
// One lane per target.  The weights are formed in FP64 in the reference's expression order
// (compiled without contraction), the six plane values of each stencil column are summed
// left to right, the four row results are rounded to FP32 (the reference stores them in a
// `real` array) before the x sum, and the same for the two rows of the linear result.
// Vertical modes: quintic for 3 <= kk1 <= nk-3 (6 planes), cubic for kk1 = 2 or nk-2
// (4 planes), linear otherwise with kk1 clamped to [1, nk-1] (2 planes).
#include <cstring>

#include "adz_common.h"
#include "adz_qv.h"

namespace adz {

constexpr int kQvBlock = 256;
constexpr int kQvFields = 16;

struct QvArgs {
  QvDev q;
  const void* xyz;
  int i0, j0, k0, nx, ny;
  long npts;
  int dni, dnj;            // destination layout (row / plane pitch)
  int sni;                 // source row pitch
  long snij;               // source plane pitch
  int ilo, ihi, jlo, jhi;  // local index range of the source arrays
  int nf;
  const float* src[kQvFields];
  float* dst[kQvFields];
  float* lin[kQvFields];
  float* mn[kQvFields];
  float* mx[kQvFields];
};

__device__ __forceinline__ double qv_tri(double za, double zb, double zc, double zd) {
  return ((za - zb) * (za - zc) * (za - zd));
}
__device__ __forceinline__ double qv_qui(double zx, double za, double zb, double zc, double zd, double ze) {
  return ((zx - za) * (zx - zb) * (zx - zc) * (zx - zd) * (zx - ze));
}
__device__ __forceinline__ float qv_max(float a, float b) { return (b > a) ? b : a; }
__device__ __forceinline__ float qv_min(float a, float b) { return (b < a) ? b : a; }

// cubic weights of the reference's bicubic form (bicubHQV_lin_mono.F90:67-75)
__device__ __forceinline__ void qv_cubic(double del, double* c) {
  const double ov6 = 1.0 / 6.0, ov2 = 1.0 / 2.0;
  const double del12 = (del - 1) * (del - 2);
  const double del10 = (del + 1) * del;
  c[0] = -(del * del12 * ov6);
  c[1] = (del + 1) * del12 * ov2;
  c[2] = -(del10 * (del - 2) * ov2);
  c[3] = del10 * (del - 1) * ov6;
}

template <class TP, int POST>
__global__ __launch_bounds__(kQvBlock) void k_qv(QvArgs A) {
  const long p = (long)blockIdx.x * kQvBlock + threadIdx.x;
  if (p >= A.npts) return;
  const int ii = (int)(p % A.nx);
  const long r = p / A.nx;
  const int i = A.i0 + ii, j = A.j0 + (int)(r % A.ny), k = A.k0 + (int)(r / A.ny);
  const long o = (long)(i - 1) + (long)A.dni * ((j - 1) + (long)A.dnj * (k - 1));
  double px, py, pz;
  load_pos<TP>(A.xyz, p, &px, &py, &pz);
  const QvDev& Q = A.q;
  bool ok = fabs(px) < 1.0e9 && fabs(py) < 1.0e9 && fabs(pz) < 1.0e9;
  double cx[4], cy[4], qz[6], lx[2], ly[2], lz[2];
  int ix = 0, iy = 0, pb0 = 0, zq = 0, zc = 0;
  if (ok) {
    const int ii1 = (int)px, jj1 = (int)py;
    const double dx = px - (double)ii1, dy = py - (double)jj1;
    qv_cubic(dx, cx);
    lx[0] = 1.0 - dx; lx[1] = dx;
    qv_cubic(dy, cy);
    ly[0] = 1.0 - dy; ly[1] = dy;
    ix = ii1 + Q.offi;  // local index of the target cell: columns ix-1..ix+2
    iy = jj1 + Q.offj;
    ok = ix - 1 >= A.ilo && ix + 2 <= A.ihi && iy - 1 >= A.jlo && iy + 2 <= A.jhi;
    int kk1 = (int)pz;
    const int nk = Q.nk;
    zq = kk1 >= 3 && kk1 <= nk - 3;
    zc = kk1 == 2 || kk1 == nk - 2;
    if (!zq && !zc) kk1 = min(nk - 1, max(1, kk1));
    const double* z = Q.posz;
    const double zt = z[kk1] + (pz - (double)kk1) * Q.dz[kk1];
    lz[1] = (zt - z[kk1]) * Q.odz[kk1];
    lz[0] = 1.0 - lz[1];
    const long K = nk + 1;
    if (zq) {
      pb0 = kk1 - 3;
      const double rx = z[kk1 - 2], ra = z[kk1 - 1], rb = z[kk1], rc = z[kk1 + 1], rd = z[kk1 + 2], re = z[kk1 + 3];
      qz[0] = qv_qui(zt, ra, rb, rc, rd, re) * Q.c6[kk1];
      qz[1] = qv_qui(zt, rx, rb, rc, rd, re) * Q.c6[K + kk1];
      qz[2] = qv_qui(zt, rx, ra, rc, rd, re) * Q.c6[2 * K + kk1];
      qz[3] = qv_qui(zt, rx, ra, rb, rd, re) * Q.c6[3 * K + kk1];
      qz[4] = qv_qui(zt, rx, ra, rb, rc, re) * Q.c6[4 * K + kk1];
      qz[5] = qv_qui(zt, rx, ra, rb, rc, rd) * Q.c6[5 * K + kk1];
    } else if (zc) {
      pb0 = kk1 - 2;
      const double ra = z[kk1 - 1], rb = z[kk1], rc = z[kk1 + 1], rd = z[kk1 + 2];
      qz[0] = 0.;
      qz[1] = qv_tri(zt, rb, rc, rd) * Q.c4[kk1];
      qz[2] = qv_tri(zt, ra, rc, rd) * Q.c4[K + kk1];
      qz[3] = qv_tri(zt, ra, rb, rd) * Q.c4[2 * K + kk1];
      qz[4] = qv_tri(zt, ra, rb, rc) * Q.c4[3 * K + kk1];
      qz[5] = 0.;
    } else {
      pb0 = kk1 - 1;
      qz[0] = 0.;
      qz[1] = 0.;
      qz[3] = (zt - z[kk1]) * Q.odz[kk1];
      qz[2] = 1.0 - qz[3];
      qz[4] = 0.;
      qz[5] = 0.;
    }
  }
  // plane offsets of the six terms (bicubHQV_lin_min_max_zyx.F90:41-50)
  const long ninj = A.snij;
  const long ninjc = zq ? ninj : 0, ninjl = (zq || zc) ? ninj : 0;
  const long n1 = ninjc, n2 = n1 + ninjl, n3 = n2 + ninj, n4 = n3 + ninjl, n5 = n4 + ninjc;
  const long ni = A.sni, ni2 = ni + ni, ni3 = ni2 + ni;
  const long base = (long)(ix - 2) + (long)(iy - 2) * A.sni + (long)pb0 * A.snij;
  for (int f = 0; f < A.nf; f++) {
    float v, li, mi, ma;
    if (ok) {
      const ADZ_GLOBAL float* s = (const ADZ_GLOBAL float*)A.src[f] + base;
      auto S = [&](long q) { return (double)s[q]; };
      auto col = [&](long b) {
        return S(b) * qz[0] + S(b + n1) * qz[1] + S(b + n2) * qz[2] + S(b + n3) * qz[3] + S(b + n4) * qz[4] +
               S(b + n5) * qz[5];
      };
      float dst[4];
#pragma unroll
      for (int c = 0; c < 4; c++) {
        const double va = col(c), vb = col(c + ni), vc = col(c + ni2), vd = col(c + ni3);
        dst[c] = (float)(va * cy[0] + vb * cy[1] + vc * cy[2] + vd * cy[3]);
      }
      v = (float)((double)dst[0] * cx[0] + (double)dst[1] * cx[1] + (double)dst[2] * cx[2] + (double)dst[3] * cx[3]);
      float dsl[2];
#pragma unroll
      for (int c = 1; c <= 2; c++) {
        const double vb = S(c + ni + n2) * lz[0] + S(c + ni + n3) * lz[1];
        const double vc = S(c + ni2 + n2) * lz[0] + S(c + ni2 + n3) * lz[1];
        dsl[c - 1] = (float)(vb * ly[0] + vc * ly[1]);
      }
      li = (float)((double)dsl[0] * lx[0] + (double)dsl[1] * lx[1]);
      float hi = s[1 + ni + n2], lo = hi;
#pragma unroll
      for (int c = 1; c <= 2; c++) {
        hi = qv_max(hi, s[c + ni + n2]); hi = qv_max(hi, s[c + ni + n3]);
        lo = qv_min(lo, s[c + ni + n2]); lo = qv_min(lo, s[c + ni + n3]);
        hi = qv_max(hi, s[c + ni2 + n2]); hi = qv_max(hi, s[c + ni2 + n3]);
        lo = qv_min(lo, s[c + ni2 + n2]); lo = qv_min(lo, s[c + ni2 + n3]);
      }
      mi = lo;
      ma = hi;
    } else {
      v = li = mi = ma = __builtin_nanf("");
    }
    // adz_interp_hlt.F90:146-163: the cubic result (clipped for mono = 1), lin clipped to
    // [min, max], min and max
    if (POST & ADZ_POST_CLIP) v = (float)fmax((double)mi, fmin((double)ma, (double)v));
    ((ADZ_GLOBAL float*)A.dst[f])[o] = v;
    if (POST & ADZ_POST_STORE) {
      ((ADZ_GLOBAL float*)A.lin[f])[o] = (float)fmax((double)mi, fmin((double)ma, (double)li));
      ((ADZ_GLOBAL float*)A.mn[f])[o] = mi;
      ((ADZ_GLOBAL float*)A.mx[f])[o] = ma;
    }
  }
}

template <class TP>
void* pick_qv(int post) {
  switch (post) {
    case 0: return (void*)&k_qv<TP, 0>;
    case 1: return (void*)&k_qv<TP, 1>;
    case 2: return (void*)&k_qv<TP, 2>;
    default: return (void*)&k_qv<TP, 3>;
  }
}

}  // namespace adz

using namespace adz;

extern "C" {

int adz_tricub_hlt_qv(adz_ctx* ctx, const adz_field* fld, int nptr, const void* xyz, int i0, int in, int j0, int jn,
                      int k0, int dminx, int dmaxx, int dminy, int dmaxy, int ext, int post, adz_stream stream) {
  QvDev q;
  if (!ctx_qv(ctx, &q)) return fail(ADZ_EINVAL, "adz_tricub_hlt_qv: null context");
  if (!fld || nptr < 0 || !xyz) return fail(ADZ_EINVAL, "adz_tricub_hlt_qv: bad argument");
  if (nptr == 0 || in < i0 || jn < j0) return ADZ_OK;  // nij <= 0: nothing to do (:61)
  if (post < 0 || post > 3) return fail(ADZ_EINVAL, "adz_tricub_hlt_qv: post must be 0..3");
  if (k0 < 1 || k0 > q.nk || i0 < dminx || in > dmaxx || j0 < dminy || jn > dmaxy)
    return fail(ADZ_EINVAL, "adz_tricub_hlt_qv: target box outside the destination");
  for (int f = 0; f < nptr; f++) {
    if (!fld[f].src || !fld[f].dst) return fail(ADZ_EINVAL, "adz_tricub_hlt_qv: null src/dst");
    if ((post & ADZ_POST_STORE) && (!fld[f].lin || !fld[f].mn || !fld[f].mx))
      return fail(ADZ_EINVAL, "adz_tricub_hlt_qv: post arrays required");
  }
  QvArgs a;
  std::memset(&a, 0, sizeof(a));
  a.q = q;
  a.xyz = xyz;
  a.i0 = i0; a.j0 = j0; a.k0 = k0; a.nx = in - i0 + 1; a.ny = jn - j0 + 1;
  a.npts = (long)a.nx * a.ny * (q.nk - k0 + 1);
  a.dni = dmaxx - dminx + 1; a.dnj = dmaxy - dminy + 1;
  if (ext) {
    a.ilo = q.lminx; a.ihi = q.lmaxx; a.jlo = q.lminy; a.jhi = q.lmaxy;
  } else {  // the source has the destination's (model) layout (adz_interp_hlt.F90:70-89 skipped)
    a.ilo = dminx; a.ihi = dmaxx; a.jlo = dminy; a.jhi = dmaxy;
  }
  a.sni = a.ihi - a.ilo + 1;
  a.snij = (long)a.sni * (a.jhi - a.jlo + 1);
  void* kern = q.pos_dtype == ADZ_F32 ? pick_qv<float>(post) : pick_qv<double>(post);
  for (int f0 = 0; f0 < nptr; f0 += kQvFields) {
    a.nf = std::min(kQvFields, nptr - f0);
    for (int f = 0; f < a.nf; f++) {
      const adz_field& e = fld[f0 + f];
      a.src[f] = (const float*)e.src; a.dst[f] = (float*)e.dst;
      a.lin[f] = (float*)e.lin; a.mn[f] = (float*)e.mn; a.mx[f] = (float*)e.mx;
    }
    void* args[] = {&a};
    ADZ_HIP_CHECK(hipLaunchKernel(kern, dim3((unsigned)((a.npts + kQvBlock - 1) / kQvBlock)), dim3(kQvBlock), args, 0,
                                  (hipStream_t)stream));
  }
  return ADZ_OK;
}

}  // extern "C"
