// Per-target field evaluators shared by the interpolation kernels (gfx950).
#pragma once
#include <type_traits>

#include "adz_common.h"

namespace adz {

// ------------------------------------------------------------------ field evaluators
// Reference arithmetic, AVX2 order (turbo_tricublin.c:553-615): per row, z by one mul and
// three fma, rows folded into y by fma, x by (y0c0+y2c2)+(y1c1+y3c3).
template <class TF, class P>
__device__ __forceinline__ double tricub_ref(P f1, const double* c, long ni, long o1, long o2, long o3) {
  double ya[4];
#pragma unroll
  for (int r = 0; r < 4; r++) {
    P s = f1 + r * ni;
#pragma unroll
    for (int i = 0; i < 4; i++) {
      double za = to_f64<TF>(s[i]) * c[8];
      za = fma(to_f64<TF>(s[i + o1]), c[9], za);
      za = fma(to_f64<TF>(s[i + o2]), c[10], za);
      za = fma(to_f64<TF>(s[i + o3]), c[11], za);
      ya[i] = (r == 0) ? za * c[4] : fma(za, c[4 + r], ya[i]);
    }
  }
  double y0 = ya[0] * c[0], y1 = ya[1] * c[1], y2 = ya[2] * c[2], y3 = ya[3] * c[3];
  return (y0 + y2) + (y1 + y3);
}

// FP32 arithmetic of the mixed form, same contraction structure.
template <class TF, class P>
__device__ __forceinline__ float tricub_f32(P f1, const float* c, long ni, long o1, long o2, long o3) {
  float ya[4];
#pragma unroll
  for (int r = 0; r < 4; r++) {
    P s = f1 + r * ni;
#pragma unroll
    for (int i = 0; i < 4; i++) {
      float za = to_f32<TF>(s[i]) * c[8];
      za = fmaf(to_f32<TF>(s[i + o1]), c[9], za);
      za = fmaf(to_f32<TF>(s[i + o2]), c[10], za);
      za = fmaf(to_f32<TF>(s[i + o3]), c[11], za);
      ya[i] = (r == 0) ? za * c[4] : fmaf(za, c[4 + r], ya[i]);
    }
  }
  return fmaf(ya[3], c[3], fmaf(ya[2], c[2], fmaf(ya[1], c[1], ya[0] * c[0])));
}

// min / max / linear of the monotone interpolator (turbo_tricublin.c:883-989): min/max
// over x columns 1..2, y rows 0..3, z planes 1..2 with the MINPD/MAXPD operand rule; the
// linear result from rows 1..2 and planes 1..2.
// The FP32 form (TC = float) takes min / max with v_min3/v_max3 (IEEE minNum/maxNum):
// the same values except where the operand rule matters (NaN, or +0 against -0).
template <class TF, class TC, class P>
__device__ __forceinline__ void mono_extra(P f1, const TC* c, long ni, long o1, long o2, TC* lin, TC* mi, TC* ma) {
  if constexpr (std::is_same<TC, float>::value) {
    float lo = to_f32<TF>(f1[1 + o1]), hi = lo;
#pragma unroll
    for (int r = 0; r < 4; r++) {
      P s = f1 + r * ni;
      float a1 = to_f32<TF>(s[1 + o1]), a2 = to_f32<TF>(s[2 + o1]);
      float b1 = to_f32<TF>(s[1 + o2]), b2 = to_f32<TF>(s[2 + o2]);
      lo = fminf(fminf(lo, a1), fminf(a2, fminf(b1, b2)));
      hi = fmaxf(fmaxf(hi, a1), fmaxf(a2, fmaxf(b1, b2)));
    }
    *mi = lo;
    *ma = hi;
  } else {
    TC vmi[4], vma[4];
#pragma unroll
    for (int r = 0; r < 4; r++) {
      P s = f1 + r * ni;
#pragma unroll
      for (int i = 1; i < 3; i++) {
        TC a = (TC)to_f64<TF>(s[i + o1]), b = (TC)to_f64<TF>(s[i + o2]);
        if (r == 0) { vmi[i] = a; vma[i] = a; }
        else { vmi[i] = (vmi[i] < a) ? vmi[i] : a; vma[i] = (vma[i] > a) ? vma[i] : a; }
        vmi[i] = (vmi[i] < b) ? vmi[i] : b;
        vma[i] = (vma[i] > b) ? vma[i] : b;
      }
    }
    *mi = (vmi[1] < vmi[2]) ? vmi[1] : vmi[2];
    *ma = (vma[1] > vma[2]) ? vma[1] : vma[2];
  }
  TC vyl[4];
#pragma unroll
  for (int i = 0; i < 4; i++) {
    P s1 = f1 + ni;
    P s2 = f1 + 2 * ni;
    TC vz1 = (TC)to_f64<TF>(s1[i + o1]) * c[18];
    vz1 = fma(((TC)to_f64<TF>(s1[i + o2])), c[19], vz1);
    TC vy = vz1 * c[16];
    TC vz2 = (TC)to_f64<TF>(s2[i + o1]) * c[18];
    vz2 = fma((TC)to_f64<TF>(s2[i + o2]), c[19], vz2);
    vyl[i] = fma(vz2, c[17], vy);
  }
  TC v0 = vyl[0] * c[12], v1 = vyl[1] * c[13], v2 = vyl[2] * c[14], v3 = vyl[3] * c[15];
  *lin = (v0 + v2) + (v1 + v3);
}

}  // namespace adz
