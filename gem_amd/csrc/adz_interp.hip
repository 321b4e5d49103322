// Tricubic Lagrange interpolation at departure points on gfx950: the device form of
// GEM's turbo_tricublin.c interpolators (src/gemdyn/src/base/turbo_tricublin.c).
//
// One lane per target point; consecutive lanes take consecutive targets, i.e. consecutive
// arrival points along i, so the 16 stencil rows a wavefront reads for one field are 16
// nearly contiguous segments of ~67 elements: each row load coalesces into 2-3 cache
// lines.  The index and weights are computed once per target and reused for every field
// of the call (the reference does the same, :803-814).
#include <cstdio>
#include <mutex>
#include <string>
#include <vector>

#include "adz_common.h"
#include "adz_eval.h"

namespace adz {

static thread_local std::string g_last_error;
void set_error(const std::string& msg) { g_last_error = msg; }
int fail(int code, const std::string& msg) {
  set_error(msg);
  return code;
}

constexpr int kMaxFieldsPerLaunch = 64;
struct FieldPtrs {
  const void* f[kMaxFieldsPerLaunch];
};

// ------------------------------------------------------------------ kernels
enum { K_ZYX1 = 0, K_MONO = 1 };

template <class TF, class TP, int MATH, int KIND>
__global__ __launch_bounds__(256) void k_tricub_m_n(TF* __restrict__ d, TF* __restrict__ l, TF* __restrict__ mi,
                                                     TF* __restrict__ ma, FieldPtrs fs, const void* __restrict__ pxyz,
                                                     ZtabDev lv, long n, int m, long dstride) {
  long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  double px, py, pz;
  load_pos<TP>(pxyz, p, &px, &py, &pz);
  Index x;
  const bool ok = index_step(px, py, pz, lv, &x);
  const long ni = lv.ni;
  const long o1 = x.zlinear ? 0 : (long)lv.nij;
  const long o2 = o1 + lv.nij, o3 = o2 + o1;
  if (MATH == ADZ_MATH_REF) {
    double c[20];
    if (ok) weights_ref(x, lv, c);
    for (int f = 0; f < m; f++) {
      long o = (long)f * dstride + p;
      if (!ok) {
        d[o] = from_f64<TF>(__builtin_nan(""));
        if (KIND == K_MONO) { l[o] = d[o]; mi[o] = d[o]; ma[o] = d[o]; }
        continue;
      }
      const TF* f1 = (const TF*)fs.f[f] + x.off;
      d[o] = from_f64<TF>(tricub_ref<TF>(f1, c, ni, o1, o2, o3));
      if (KIND == K_MONO) {
        double a, b, e;
        mono_extra<TF, double>(f1, c, ni, o1, o2, &a, &b, &e);
        l[o] = from_f64<TF>(a); mi[o] = from_f64<TF>(b); ma[o] = from_f64<TF>(e);
      }
    }
  } else {
    float c[20];
    if (ok) weights_f32(x, lv, c);
    for (int f = 0; f < m; f++) {
      long o = (long)f * dstride + p;
      if (!ok) {
        d[o] = from_f32<TF>(__builtin_nanf(""));
        if (KIND == K_MONO) { l[o] = d[o]; mi[o] = d[o]; ma[o] = d[o]; }
        continue;
      }
      const TF* f1 = (const TF*)fs.f[f] + x.off;
      float v = tricub_f32<TF>(f1, c, ni, o1, o2, o3);
      if (KIND == K_MONO) {
        float a, b, e;
        mono_extra<TF, float>(f1, c, ni, o1, o2, &a, &b, &e);
        v = snap_to_range(v, b, e);
        l[o] = from_f32<TF>(a); mi[o] = from_f32<TF>(b); ma[o] = from_f32<TF>(e);
      }
      d[o] = from_f32<TF>(v);
    }
  }
}

// Tricublin_zyxf3_inline (turbo_tricublin.c:1122-1287): three interleaved components;
// AVX2 x stage accumulates columns by fma (:1243-1257).
template <class TF, class TP, int MATH>
__global__ __launch_bounds__(256) void k_tricub3(TF* __restrict__ d, const TF* __restrict__ f3,
                                                  const void* __restrict__ pxyz, ZtabDev lv, long n) {
  long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  double px, py, pz;
  load_pos<TP>(pxyz, p, &px, &py, &pz);
  Index x;
  if (!index_step(px, py, pz, lv, &x)) {
    for (int q = 0; q < 3; q++) d[3 * p + q] = from_f64<TF>(__builtin_nan(""));
    return;
  }
  const long ni = 3L * lv.ni, nij = 3L * lv.nij;
  const long o1 = x.zlinear ? 0 : nij;
  const long o2 = o1 + nij, o3 = o2 + o1;
  const TF* f = f3 + 3 * x.off;
  if (MATH == ADZ_MATH_REF) {
    double c[20], o[3];
    weights_ref(x, lv, c);
    tricub3_ref<TF>(f, c, ni, nij, x.zlinear, o);
#pragma unroll
    for (int q = 0; q < 3; q++) d[3 * p + q] = from_f64<TF>(o[q]);
  } else {
    float c[20];
    weights_f32(x, lv, c);
    float dst[12];
#pragma unroll
    for (int e = 0; e < 12; e++) {
      float ya = 0.0f;
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const TF* s = f + r * ni;
        float za = to_f32(s[e]) * c[8];
        za = fmaf(to_f32(s[e + o1]), c[9], za);
        za = fmaf(to_f32(s[e + o2]), c[10], za);
        za = fmaf(to_f32(s[e + o3]), c[11], za);
        ya = (r == 0) ? za * c[4] : fmaf(za, c[4 + r], ya);
      }
      dst[e] = ya;
    }
#pragma unroll
    for (int q = 0; q < 3; q++) {
      float x0 = dst[q] * c[0];
      x0 = fmaf(dst[q + 3], c[1], x0);
      x0 = fmaf(dst[q + 6], c[2], x0);
      x0 = fmaf(dst[q + 9], c[3], x0);
      d[3 * p + q] = from_f32<TF>(x0);
    }
  }
}

template <class TP>
__global__ __launch_bounds__(256) void k_vcoef(int* __restrict__ ijk, int* __restrict__ zlin, double* __restrict__ coef,
                                                const void* __restrict__ pxyz, ZtabDev lv, long n) {
  long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  double px, py, pz;
  load_pos<TP>(pxyz, p, &px, &py, &pz);
  Index x;
  if (!index_step(px, py, pz, lv, &x)) {
    ijk[p] = INT32_MIN; zlin[p] = 0;
    if (coef) for (int q = 0; q < 20; q++) coef[20 * p + q] = __builtin_nan("");
    return;
  }
  ijk[p] = (int)x.off;
  zlin[p] = x.zlinear;
  if (coef) {
    double c[20];
    weights_ref(x, lv, c);
    for (int q = 0; q < 20; q++) coef[20 * p + q] = c[q];
  }
}

}  // namespace adz

// ====================================================================== C ABI
using namespace adz;

namespace {
inline unsigned nblocks(long n) { return (unsigned)((n + 255) / 256); }
bool dtype_ok(int t) { return t == ADZ_F64 || t == ADZ_F32 || t == ADZ_BF16; }
bool pos_ok(int t) { return t == ADZ_F64 || t == ADZ_F32; }
bool math_ok(int m) { return m == ADZ_MATH_REF || m == ADZ_MATH_F32; }

template <class TF, class TP, int MATH, int KIND>
void launch_m_n(void* d, void* l, void* mi, void* ma, const void* const* fs, const void* pxyz, const ZtabDev& lv,
                long n, int m, hipStream_t st) {
  for (int f0 = 0; f0 < m; f0 += kMaxFieldsPerLaunch) {
    int mm = m - f0 < kMaxFieldsPerLaunch ? m - f0 : kMaxFieldsPerLaunch;
    FieldPtrs P;
    for (int f = 0; f < mm; f++) P.f[f] = fs[f0 + f];
    long o = (long)f0 * n;
    hipLaunchKernelGGL((k_tricub_m_n<TF, TP, MATH, KIND>), dim3(nblocks(n)), dim3(256), 0, st, (TF*)d + o,
                       l ? (TF*)l + o : nullptr, mi ? (TF*)mi + o : nullptr, ma ? (TF*)ma + o : nullptr, P, pxyz,
                       lv, n, mm, n);
  }
}

template <int KIND>
int dispatch_m_n(void* d, void* l, void* mi, void* ma, const void* const* fs, const void* pxyz, const adz_ztab* lv,
                 int n, int m, int fdt, int pdt, int math, adz_stream s) {
  if (!d || !fs || !pxyz || !lv || n < 0 || m < 0) return fail(ADZ_EINVAL, "tricublin: null pointer or negative size");
  if (KIND == K_MONO && (!l || !mi || !ma)) return fail(ADZ_EINVAL, "tricublin_mono: null output");
  if (!dtype_ok(fdt) || !pos_ok(pdt) || !math_ok(math)) return fail(ADZ_EINVAL, "tricublin: bad dtype/math selector");
  for (int f = 0; f < m; f++)
    if (!fs[f]) return fail(ADZ_EINVAL, "tricublin: null source field pointer");
  if (n == 0 || m == 0) return ADZ_OK;
  hipStream_t st = (hipStream_t)s;
  const ZtabDev& z = lv->dev;
#define ADZ_CASE(TF_, TP_, M_)                                                                \
  launch_m_n<TF_, TP_, M_, KIND>(d, l, mi, ma, fs, pxyz, z, n, m, st);                      \
  break;
  int key = fdt * 100 + pdt * 10 + math;
  switch (key) {
    case ADZ_F64 * 100 + ADZ_F64 * 10 + ADZ_MATH_REF: ADZ_CASE(double, double, ADZ_MATH_REF)
    case ADZ_F64 * 100 + ADZ_F32 * 10 + ADZ_MATH_REF: ADZ_CASE(double, float, ADZ_MATH_REF)
    case ADZ_F32 * 100 + ADZ_F64 * 10 + ADZ_MATH_REF: ADZ_CASE(float, double, ADZ_MATH_REF)
    case ADZ_F32 * 100 + ADZ_F32 * 10 + ADZ_MATH_REF: ADZ_CASE(float, float, ADZ_MATH_REF)
    case ADZ_F32 * 100 + ADZ_F64 * 10 + ADZ_MATH_F32: ADZ_CASE(float, double, ADZ_MATH_F32)
    case ADZ_F32 * 100 + ADZ_F32 * 10 + ADZ_MATH_F32: ADZ_CASE(float, float, ADZ_MATH_F32)
    case ADZ_BF16 * 100 + ADZ_F64 * 10 + ADZ_MATH_F32: ADZ_CASE(bf16, double, ADZ_MATH_F32)
    case ADZ_BF16 * 100 + ADZ_F32 * 10 + ADZ_MATH_F32: ADZ_CASE(bf16, float, ADZ_MATH_F32)
    case ADZ_F64 * 100 + ADZ_F64 * 10 + ADZ_MATH_F32: ADZ_CASE(double, double, ADZ_MATH_F32)
    case ADZ_F64 * 100 + ADZ_F32 * 10 + ADZ_MATH_F32: ADZ_CASE(double, float, ADZ_MATH_F32)
    default: return fail(ADZ_EINVAL, "tricublin: unsupported (field dtype, position dtype, math) combination");
  }
#undef ADZ_CASE
  ADZ_HIP_CHECK(hipGetLastError());
  return ADZ_OK;
}
}  // namespace

extern "C" {

const char* adz_last_error(void) { return g_last_error.c_str(); }
int adz_version(void) { return 100; }

int adz_vsearch_setup_plus(const double* levels, int nk, int ni, int nj, int offseti, int offsetj, adz_ztab** out) {
  if (!levels || !out || nk < 4 || ni < 4 || nj < 4) return fail(ADZ_EINVAL, "vsearch_setup: need nk>=4, ni,nj>=4");
  // turbo_tricublin.c:320-363: z copy and inverse denominators; entries 0, nk-2 and
  // nk-1 reuse the first / last complete stencils.
  std::vector<double> z(levels, levels + nk), ocz(4 * nk);
  auto den = [](double* r, double a, double b, double c, double d) {
    r[0] = 1.0 / ((a - b) * (a - c) * (a - d));
    r[1] = 1.0 / ((b - a) * (b - c) * (b - d));
    r[2] = 1.0 / ((c - a) * (c - b) * (c - d));
    r[3] = 1.0 / ((d - a) * (d - b) * (d - c));
  };
  den(&ocz[0], z[0], z[1], z[2], z[3]);
  for (int i = 1; i < nk - 2; i++) den(&ocz[4 * i], z[i - 1], z[i], z[i + 1], z[i + 2]);
  den(&ocz[4 * (nk - 2)], z[nk - 4], z[nk - 3], z[nk - 2], z[nk - 1]);
  den(&ocz[4 * (nk - 1)], z[nk - 4], z[nk - 3], z[nk - 2], z[nk - 1]);
  std::vector<float> hz(4 * nk, 0.0f), oczf(4 * nk);
  for (int i = 0; i < nk; i++) {
    if (i >= 1 && i + 2 < nk) {
      hz[4 * i + 0] = (float)(z[i] - z[i - 1]);
      hz[4 * i + 1] = (float)(z[i + 1] - z[i]);
      hz[4 * i + 2] = (float)(z[i + 2] - z[i + 1]);
    }
    for (int q = 0; q < 4; q++) oczf[4 * i + q] = (float)ocz[4 * i + q];
  }
  size_t bytes = (size_t)nk * (8 + 32 + 16 + 16);
  void* mem = nullptr;
  if (hipMalloc(&mem, bytes) != hipSuccess) return fail(ADZ_ENOMEM, "vsearch_setup: hipMalloc failed");
  char* b = (char*)mem;
  double* dz = (double*)b;
  double* docz = (double*)(b + 8 * nk);
  float* dhz = (float*)(b + 40 * nk);
  float* doczf = (float*)(b + 56 * nk);
  if (hipMemcpy(dz, z.data(), 8 * nk, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(docz, ocz.data(), 32 * nk, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(dhz, hz.data(), 16 * nk, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(doczf, oczf.data(), 16 * nk, hipMemcpyHostToDevice) != hipSuccess) {
    (void)hipFree(mem);
    return fail(ADZ_EHIP, "vsearch_setup: hipMemcpy failed");
  }
  adz_ztab* t = new adz_ztab;
  t->mem = mem;
  ZtabDev& v = t->dev;
  v.ni = ni; v.nj = nj; v.nk = nk; v.nij = ni * nj; v.offi = offseti; v.offj = offsetj;
  v.ilo = 1; v.ihi = ni; v.jlo = 1; v.jhi = nj;
  v.z = dz; v.ocz = docz; v.hz = dhz; v.oczf = doczf;
  *out = t;
  return ADZ_OK;
}

int adz_vsearch_set_bounds(adz_ztab* lv, int ilo, int ihi, int jlo, int jhi) {
  if (!lv || ihi - ilo + 1 > lv->dev.ni || jhi - jlo + 1 > lv->dev.nj || ihi < ilo || jhi < jlo)
    return fail(ADZ_EINVAL, "vsearch_set_bounds: range larger than the described array");
  lv->dev.ilo = ilo; lv->dev.ihi = ihi; lv->dev.jlo = jlo; lv->dev.jhi = jhi;
  return ADZ_OK;
}

int adz_vsearch_free(adz_ztab* lv) {
  if (!lv) return ADZ_OK;
  hipError_t e = hipFree(lv->mem);
  delete lv;
  if (e != hipSuccess) return fail(ADZ_EHIP, "vsearch_free: hipFree failed");
  return ADZ_OK;
}

int adz_vcoef(int* ijk, int* zlin, double* coef, const void* pxyz, int pdtype, const adz_ztab* lv, int n,
              adz_stream s) {
  if (!ijk || !zlin || !pxyz || !lv || n < 0 || !pos_ok(pdtype)) return fail(ADZ_EINVAL, "vcoef: bad argument");
  if (n == 0) return ADZ_OK;
  if (pdtype == ADZ_F32)
    hipLaunchKernelGGL(k_vcoef<float>, dim3(nblocks(n)), dim3(256), 0, (hipStream_t)s, ijk, zlin, coef, pxyz, lv->dev, (long)n);
  else
    hipLaunchKernelGGL(k_vcoef<double>, dim3(nblocks(n)), dim3(256), 0, (hipStream_t)s, ijk, zlin, coef, pxyz, lv->dev, (long)n);
  ADZ_HIP_CHECK(hipGetLastError());
  return ADZ_OK;
}

int adz_tricublin_zyx1_m_n(void* d, const void* const* fs, const void* pxyz, const adz_ztab* lv, int n, int m,
                           int fdtype, int pdtype, int math, adz_stream stream) {
  return dispatch_m_n<K_ZYX1>(d, nullptr, nullptr, nullptr, fs, pxyz, lv, n, m, fdtype, pdtype, math, stream);
}

int adz_tricublin_mono_zyx_m_n(void* d, void* l, void* mi, void* ma, const void* const* fs, const void* pxyz,
                               const adz_ztab* lv, int n, int m, int fdtype, int pdtype, int math,
                               adz_stream stream) {
  return dispatch_m_n<K_MONO>(d, l, mi, ma, fs, pxyz, lv, n, m, fdtype, pdtype, math, stream);
}

int adz_tricublin_zyx3_n(void* d, const void* f3, const void* pxyz, const adz_ztab* lv, int n, int fdtype,
                         int pdtype, int math, adz_stream stream) {
  if (!d || !f3 || !pxyz || !lv || n < 0) return fail(ADZ_EINVAL, "tricublin_zyx3: bad argument");
  if (!pos_ok(pdtype) || !math_ok(math)) return fail(ADZ_EINVAL, "tricublin_zyx3: bad selector");
  if (n == 0) return ADZ_OK;
  hipStream_t st = (hipStream_t)stream;
  dim3 g(nblocks(n)), b(256);
  long nn = n;
  if (fdtype == ADZ_F32 && pdtype == ADZ_F32 && math == ADZ_MATH_REF)
    hipLaunchKernelGGL((k_tricub3<float, float, ADZ_MATH_REF>), g, b, 0, st, (float*)d, (const float*)f3, pxyz, lv->dev, nn);
  else if (fdtype == ADZ_F32 && pdtype == ADZ_F64 && math == ADZ_MATH_REF)
    hipLaunchKernelGGL((k_tricub3<float, double, ADZ_MATH_REF>), g, b, 0, st, (float*)d, (const float*)f3, pxyz, lv->dev, nn);
  else if (fdtype == ADZ_F64 && pdtype == ADZ_F64 && math == ADZ_MATH_REF)
    hipLaunchKernelGGL((k_tricub3<double, double, ADZ_MATH_REF>), g, b, 0, st, (double*)d, (const double*)f3, pxyz, lv->dev, nn);
  else if (fdtype == ADZ_F64 && pdtype == ADZ_F32 && math == ADZ_MATH_REF)
    hipLaunchKernelGGL((k_tricub3<double, float, ADZ_MATH_REF>), g, b, 0, st, (double*)d, (const double*)f3, pxyz, lv->dev, nn);
  else if (fdtype == ADZ_F32 && pdtype == ADZ_F64 && math == ADZ_MATH_F32)
    hipLaunchKernelGGL((k_tricub3<float, double, ADZ_MATH_F32>), g, b, 0, st, (float*)d, (const float*)f3, pxyz, lv->dev, nn);
  else if (fdtype == ADZ_F32 && pdtype == ADZ_F32 && math == ADZ_MATH_F32)
    hipLaunchKernelGGL((k_tricub3<float, float, ADZ_MATH_F32>), g, b, 0, st, (float*)d, (const float*)f3, pxyz, lv->dev, nn);
  else
    return fail(ADZ_EINVAL, "tricublin_zyx3: unsupported (field dtype, position dtype, math) combination");
  ADZ_HIP_CHECK(hipGetLastError());
  return ADZ_OK;
}

}  // extern "C"
