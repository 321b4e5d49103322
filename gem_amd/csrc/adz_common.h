// Shared device/host helpers of libadz_hip (gfx950).  Built with -ffp-contract=off: every
// fused multiply-add in device code is an explicit fma()/fmaf(), so the operation order
// is exactly the one written.
#pragma once
#include <hip/hip_bf16.h>
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "../../include/adz_hip.h"

namespace adz {

// ------------------------------------------------------------------ error reporting
void set_error(const std::string& msg);
int fail(int code, const std::string& msg);
#define ADZ_HIP_CHECK(call)                                                                       \
  do {                                                                                            \
    hipError_t e_ = (call);                                                                       \
    if (e_ != hipSuccess) return ::adz::fail(ADZ_EHIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

// ------------------------------------------------------------------ context tables
// Air-mass tables of a context (adz_set): Ver_dz_8%t(k) and Ver_idz_8%t(k) indexed by k
// (device), the tile size and Adz_k0t.  False for a null context.
struct AirmassTables {
  const double* dz;
  const double* idz;
  int nk, ni, nj, k0t;
};
bool ctx_airmass(const adz_ctx* c, AirmassTables* t);

// ------------------------------------------------------------------ storage types
typedef __hip_bfloat16 bf16;

// Pointers read from device tables are generic; casting them into the global address
// space lets the compiler emit global_load/global_store instead of flat accesses (which
// also count against the LDS counter and force coarser waits).
#define ADZ_GLOBAL __attribute__((address_space(1)))

template <class T> __device__ __forceinline__ float to_f32(T v);
template <> __device__ __forceinline__ float to_f32<float>(float v) { return v; }
template <> __device__ __forceinline__ float to_f32<double>(double v) { return (float)v; }
template <> __device__ __forceinline__ float to_f32<bf16>(bf16 v) {
  uint16_t b = __builtin_bit_cast(uint16_t, v);
  return __builtin_bit_cast(float, (uint32_t)b << 16);
}
template <class T> __device__ __forceinline__ double to_f64(T v) { return (double)to_f32<T>(v); }
template <> __device__ __forceinline__ double to_f64<double>(double v) { return v; }
// Global-address-space pointer types of a storage type.  bf16 is a class whose members
// expect generic objects, so it keeps generic pointers.
template <class TF> struct GPtr { using type = ADZ_GLOBAL TF*; using ctype = const ADZ_GLOBAL TF*; };
template <> struct GPtr<bf16> { using type = bf16*; using ctype = const bf16*; };

// round-to-nearest-even FP32 -> BF16 (NaN kept quiet)
__device__ __forceinline__ bf16 f32_to_bf16(float f) {
  uint32_t u = __builtin_bit_cast(uint32_t, f);
  uint16_t r;
  if ((u & 0x7fffffffu) > 0x7f800000u) r = (uint16_t)((u >> 16) | 0x40u);
  else r = (uint16_t)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
  return __builtin_bit_cast(bf16, r);
}
template <class T> __device__ __forceinline__ T from_f64(double v);
template <> __device__ __forceinline__ double from_f64<double>(double v) { return v; }
template <> __device__ __forceinline__ float from_f64<float>(double v) { return (float)v; }
template <> __device__ __forceinline__ bf16 from_f64<bf16>(double v) { return f32_to_bf16((float)v); }
template <class T> __device__ __forceinline__ T from_f32(float v);
template <> __device__ __forceinline__ double from_f32<double>(float v) { return (double)v; }
template <> __device__ __forceinline__ float from_f32<float>(float v) { return v; }
template <> __device__ __forceinline__ bf16 from_f32<bf16>(float v) { return f32_to_bf16(v); }

// ------------------------------------------------------------------ vertical descriptor
// Device view of a vertical descriptor (the reference's ztab, turbo_tricublin.c:164-173,
// plus the FP32 spacing tables of the mixed-precision form and the source bounds).
struct ZtabDev {
  int ni, nj, nk, nij, offi, offj;
  int ilo, ihi, jlo, jhi;      // allowed LOCAL index range of the source arrays
  const double* z;             // nk levels
  const double* ocz;           // 4*nk inverse Lagrange denominators (FP64)
  const float* hz;             // 4*nk: z[iz]-z[iz-1], z[iz+1]-z[iz], z[iz+2]-z[iz+1], 0
  const float* oczf;           // 4*nk inverse denominators rounded to FP32
};

// Index step shared by every form: truncation of the (FP32 or FP64) position widened to
// FP64, exactly Vcoef_pxyz4_inline (turbo_tricublin.c:395-431).  Returns false when the
// 4x4 stencil would leave the allowed source range (or the position is not finite).
struct Index {
  long off;      // element offset of the stencil's lower-left corner from f(1,1,1)
  int zlinear;   // nonzero: linear along z (4x4x2 stencil)
  int iz0;       // iz - 1 (0-based level of the interval's lower level)
  int ix, iy;    // local (1-based) index of the target cell; stencil columns ix-1..ix+2
  int pz;        // 0-based first plane of the stencil (iz0 - 1 cubic, iz0 linear)
  double dx, dy, dz;
};

__device__ __forceinline__ bool index_step(double px, double py, double pz, const ZtabDev& lv, Index* r) {
  // reject NaN / huge values before the integer conversion
  if (!(fabs(px) < 1.0e9 && fabs(py) < 1.0e9 && fabs(pz) < 1.0e9)) return false;
  int ix = (int)px;
  double dx = px - (double)ix;
  ix = ix + lv.offi;
  int iy = (int)py;
  double dy = py - (double)iy;
  iy = iy + lv.offj;
  if (ix - 1 < lv.ilo || ix + 2 > lv.ihi || iy - 1 < lv.jlo || iy + 2 > lv.jhi) return false;
  int iz = (int)pz;
  if (iz < 1) iz = 1;
  if (iz > lv.nk - 1) iz = lv.nk - 1;
  double dz = pz - (double)iz;
  long ijk = (long)(ix - 2) + (long)(iy - 2) * lv.ni + (long)(iz - 1) * lv.nij;
  iz--;
  int zlinear = ((iz - 1) | (lv.nk - 3 - iz)) >> 31;
  if (!zlinear) ijk -= lv.nij;
  r->off = ijk; r->zlinear = zlinear; r->iz0 = iz; r->dx = dx; r->dy = dy; r->dz = dz;
  r->ix = ix; r->iy = iy; r->pz = zlinear ? iz : iz - 1;
  return true;
}

// FP64 weights in the reference's formulas and order (turbo_tricublin.c:401-509, with the
// absolute z formed as fma(dz, z1, (1-dz)*z0) as the reference's -mfma build does).
// c[0..3] x, c[4..7] y, c[8..11] z, c[12..15] linear x, c[16..17] linear y, c[18..19] z.
__device__ __forceinline__ void weights_ref(const Index& x, const ZtabDev& lv, double* c) {
  const double one = 1.0, two = 2.0, cp167 = 1.0 / 6.0, cm167 = -1.0 / 6.0, cp5 = 0.5, cm5 = -0.5;
  double px = x.dx, py = x.dy, dz = x.dz;
  c[12] = 0.0; c[13] = 1.0 - px; c[14] = px; c[15] = 0.0;
  c[16] = 1.0 - py; c[17] = py;
  c[18] = 1.0 - dz; c[19] = dz;
  if (x.zlinear) {
    c[8] = 0.0; c[9] = 1.0 - dz; c[10] = dz; c[11] = 0.0;
  } else {
    const ADZ_GLOBAL double* base = (const ADZ_GLOBAL double*)lv.ocz + 4 * x.iz0;
    const ADZ_GLOBAL double* pos = (const ADZ_GLOBAL double*)lv.z + x.iz0;
    double za = fma(dz, pos[1], (1.0 - dz) * pos[0]);
    double zza = za - pos[-1], zzb = za - pos[0], zzc = za - pos[1], zzd = za - pos[2];
    double zzab = zza * zzb, zzcd = zzc * zzd;
    c[8] = zzb * zzcd * base[0];
    c[9] = zza * zzcd * base[1];
    c[10] = zzd * zzab * base[2];
    c[11] = zzc * zzab * base[3];
  }
  c[0] = cm167 * px * (px - one) * (px - two);
  c[1] = cp5 * (px + one) * (px - one) * (px - two);
  c[2] = cm5 * px * (px + one) * (px - two);
  c[3] = cp167 * px * (px + one) * (px - one);
  c[4] = cm167 * py * (py - one) * (py - two);
  c[5] = cp5 * (py + one) * (py - one) * (py - two);
  c[6] = cm5 * py * (py + one) * (py - two);
  c[7] = cp167 * py * (py + one) * (py - one);
}

// FP32 weights of the mixed-precision form.  Fractions come from the FP64 index step; the
// vertical Lagrange weights use distances rebuilt from FP32 level spacings, so no
// cancellation of absolute levels happens in FP32.  Same slot layout as weights_ref.
__device__ __forceinline__ void weights_f32(const Index& x, const ZtabDev& lv, float* c) {
  float px = (float)x.dx, py = (float)x.dy, dz = (float)x.dz;
  const float s6 = 1.0f / 6.0f;
  c[12] = 0.0f; c[13] = 1.0f - px; c[14] = px; c[15] = 0.0f;
  c[16] = 1.0f - py; c[17] = py;
  c[18] = 1.0f - dz; c[19] = dz;
  if (x.zlinear) {
    c[8] = 0.0f; c[9] = 1.0f - dz; c[10] = dz; c[11] = 0.0f;
  } else {
    const ADZ_GLOBAL float* h = (const ADZ_GLOBAL float*)lv.hz + 4 * x.iz0;
    const ADZ_GLOBAL float* o = (const ADZ_GLOBAL float*)lv.oczf + 4 * x.iz0;
    float zzb = dz * h[1];
    float zza = zzb + h[0];
    float zzc = zzb - h[1];
    float zzd = zzc - h[2];
    float zzab = zza * zzb, zzcd = zzc * zzd;
    c[8] = zzb * zzcd * o[0];
    c[9] = zza * zzcd * o[1];
    c[10] = zzd * zzab * o[2];
    c[11] = zzc * zzab * o[3];
  }
  float pm1 = px - 1.0f, pm2 = px - 2.0f, pp1 = px + 1.0f;
  c[0] = -s6 * px * pm1 * pm2;
  c[1] = 0.5f * pp1 * pm1 * pm2;
  c[2] = -0.5f * px * pp1 * pm2;
  c[3] = s6 * px * pp1 * pm1;
  float qm1 = py - 1.0f, qm2 = py - 2.0f, qp1 = py + 1.0f;
  c[4] = -s6 * py * qm1 * qm2;
  c[5] = 0.5f * qp1 * qm1 * qm2;
  c[6] = -0.5f * py * qp1 * qm2;
  c[7] = s6 * py * qp1 * qm1;
}

// Three interleaved components f(3,ni,nj,nk) in reference arithmetic: z by mul + 3 fma,
// rows folded by fma, and the x stage accumulated column by column with fma
// (Tricublin_zyxf3_inline AVX2 path, turbo_tricublin.c:1142-1257).  f points at the
// stencil corner; ni3/nij3 are the element strides of a row / plane (3*ni, 3*nij).
template <class TF>
__device__ __forceinline__ void tricub3_ref(const TF* __restrict__ f, const double* c, long ni3, long nij3,
                                            int zlinear, double* out) {
  const long o1 = zlinear ? 0 : nij3;
  const long o2 = o1 + nij3, o3 = o2 + o1;
  // rows outer, the 12 interleaved values of a row inner: only one row's values and the
  // 12 running sums are live at a time (each element keeps the same operation order)
  double dst[12];
#pragma unroll 1
  for (int r = 0; r < 4; r++) {
    const TF* s = f + r * ni3;
    double za[12];
#pragma unroll
    for (int e = 0; e < 12; e++) za[e] = to_f64(s[e]) * c[8];
#pragma unroll
    for (int e = 0; e < 12; e++) za[e] = fma(to_f64(s[e + o1]), c[9], za[e]);
#pragma unroll
    for (int e = 0; e < 12; e++) za[e] = fma(to_f64(s[e + o2]), c[10], za[e]);
#pragma unroll
    for (int e = 0; e < 12; e++) za[e] = fma(to_f64(s[e + o3]), c[11], za[e]);
    const double cyr = r == 0 ? c[4] : (r == 1 ? c[5] : (r == 2 ? c[6] : c[7]));  // no dynamic index into c
#pragma unroll
    for (int e = 0; e < 12; e++) dst[e] = (r == 0) ? za[e] * cyr : fma(za[e], cyr, dst[e]);
  }
#pragma unroll
  for (int q = 0; q < 3; q++) {
    double x0 = dst[q] * c[0];
    x0 = fma(dst[q + 3], c[1], x0);
    x0 = fma(dst[q + 6], c[2], x0);
    x0 = fma(dst[q + 9], c[3], x0);
    out[q] = x0;
  }
}

}  // namespace adz

// Host handle of a vertical descriptor (shared by every translation unit).
struct adz_ztab {
  adz::ZtabDev dev;
  void* mem;
};

namespace adz {

template <class TP>
__device__ __forceinline__ void load_pos(const void* pxyz, long p, double* x, double* y, double* z) {
  const ADZ_GLOBAL TP* q = (const ADZ_GLOBAL TP*)pxyz + 3 * p;  // positions live in global memory
  *x = (double)q[0]; *y = (double)q[1]; *z = (double)q[2];
}

}  // namespace adz
