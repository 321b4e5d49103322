// Per-target field evaluators shared by the interpolation kernels (gfx950).
#pragma once
#include <type_traits>

#include "adz_common.h"

namespace adz {

// ------------------------------------------------------------------ field evaluators
// Each evaluator reads the stencil through an accessor at(r, z, i): row r (y, 0..3),
// plane z (0..3) and column i (x, 0..3) of the 4x4x4 source block.  Pointer-based callers
// wrap a source pointer (global memory or an LDS box); the buffer-load path wraps the
// 16 row vectors it holds in registers.  All of them run the same operation order.

// Reference arithmetic, AVX2 order (turbo_tricublin.c:553-615): per row, z by one mul and
// three fma, rows folded into y by fma, x by (y0c0+y2c2)+(y1c1+y3c3).
template <class TF, class Acc>
__device__ __forceinline__ double tricub_ref_a(Acc at, const double* c) {
  double ya[4];
#pragma unroll
  for (int r = 0; r < 4; r++) {
#pragma unroll
    for (int i = 0; i < 4; i++) {
      double za = to_f64<TF>(at(r, 0, i)) * c[8];
      za = fma(to_f64<TF>(at(r, 1, i)), c[9], za);
      za = fma(to_f64<TF>(at(r, 2, i)), c[10], za);
      za = fma(to_f64<TF>(at(r, 3, i)), c[11], za);
      ya[i] = (r == 0) ? za * c[4] : fma(za, c[4 + r], ya[i]);
    }
  }
  double y0 = ya[0] * c[0], y1 = ya[1] * c[1], y2 = ya[2] * c[2], y3 = ya[3] * c[3];
  return (y0 + y2) + (y1 + y3);
}

// FP32 arithmetic of the mixed form, same contraction structure.
template <class TF, class Acc>
__device__ __forceinline__ float tricub_f32_a(Acc at, const float* c) {
  float ya[4];
#pragma unroll
  for (int r = 0; r < 4; r++) {
#pragma unroll
    for (int i = 0; i < 4; i++) {
      float za = to_f32<TF>(at(r, 0, i)) * c[8];
      za = fmaf(to_f32<TF>(at(r, 1, i)), c[9], za);
      za = fmaf(to_f32<TF>(at(r, 2, i)), c[10], za);
      za = fmaf(to_f32<TF>(at(r, 3, i)), c[11], za);
      ya[i] = (r == 0) ? za * c[4] : fmaf(za, c[4 + r], ya[i]);
    }
  }
  return fmaf(ya[3], c[3], fmaf(ya[2], c[2], fmaf(ya[1], c[1], ya[0] * c[0])));
}

// min / max / linear of the monotone interpolator (turbo_tricublin.c:883-989): min/max
// over x columns 1..2, y rows 0..3, z planes 1..2 with the MINPD/MAXPD operand rule; the
// linear result from rows 1..2 and planes 1..2.  The FP32 form (TC = float) takes min /
// max with v_min3/v_max3 (IEEE minNum/maxNum): the same values except where the operand
// rule matters (NaN, or +0 against -0).
template <class TF, class TC, class Acc>
__device__ __forceinline__ void mono_extra_a(Acc at, const TC* c, TC* lin, TC* mi, TC* ma) {
  if constexpr (std::is_same<TC, float>::value) {
    float lo = to_f32<TF>(at(0, 1, 1)), hi = lo;
#pragma unroll
    for (int r = 0; r < 4; r++) {
      float a1 = to_f32<TF>(at(r, 1, 1)), a2 = to_f32<TF>(at(r, 1, 2));
      float b1 = to_f32<TF>(at(r, 2, 1)), b2 = to_f32<TF>(at(r, 2, 2));
      lo = fminf(fminf(lo, a1), fminf(a2, fminf(b1, b2)));
      hi = fmaxf(fmaxf(hi, a1), fmaxf(a2, fmaxf(b1, b2)));
    }
    *mi = lo;
    *ma = hi;
  } else {
    TC vmi[4], vma[4];
#pragma unroll
    for (int r = 0; r < 4; r++) {
#pragma unroll
      for (int i = 1; i < 3; i++) {
        TC a = (TC)to_f64<TF>(at(r, 1, i)), b = (TC)to_f64<TF>(at(r, 2, i));
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
    TC vz1 = (TC)to_f64<TF>(at(1, 1, i)) * c[18];
    vz1 = fma(((TC)to_f64<TF>(at(1, 2, i))), c[19], vz1);
    TC vy = vz1 * c[16];
    TC vz2 = (TC)to_f64<TF>(at(2, 1, i)) * c[18];
    vz2 = fma((TC)to_f64<TF>(at(2, 2, i)), c[19], vz2);
    vyl[i] = fma(vz2, c[17], vy);
  }
  TC v0 = vyl[0] * c[12], v1 = vyl[1] * c[13], v2 = vyl[2] * c[14], v3 = vyl[3] * c[15];
  *lin = (v0 + v2) + (v1 + v3);
}

// Pointer-based forms: f1 addresses column 0, row 0, plane 0 of the stencil; rows are ni
// apart and planes 0, o1, o2, o3 from it.
template <class P>
struct PtrStencil {
  P f1;
  long ni, o1, o2, o3;
  __device__ __forceinline__ auto operator()(int r, int z, int i) const {
    const long zo = z == 0 ? 0 : (z == 1 ? o1 : (z == 2 ? o2 : o3));
    return f1[r * ni + zo + i];
  }
};

template <class TF, class P>
__device__ __forceinline__ double tricub_ref(P f1, const double* c, long ni, long o1, long o2, long o3) {
  return tricub_ref_a<TF>(PtrStencil<P>{f1, ni, o1, o2, o3}, c);
}
template <class TF, class P>
__device__ __forceinline__ float tricub_f32(P f1, const float* c, long ni, long o1, long o2, long o3) {
  return tricub_f32_a<TF>(PtrStencil<P>{f1, ni, o1, o2, o3}, c);
}
// The FP32 form's cubic value put on its stencil's inner range [mi, ma] when it lies
// outside by at most 2^-20 of the range's larger magnitude (8 to 16 ulp).  FP32 weights do
// not sum to exactly 1, so a constant neighbourhood comes out a few ulp off the constant,
// outside [mi, ma] = [c, c], where the reference's FP64 arithmetic narrowed to FP32 gives c
// itself; a shape-preserving pass (clipping, ILMC, the fixer's min/max reset) would then
// treat every such cell as a violation.  Values farther out (genuine overshoots) and NaN
// are left as they are.  The reference-arithmetic form is not touched.
__device__ __forceinline__ float snap_to_range(float v, float mi, float ma) {
  const float tol = fmaxf(fabsf(mi), fabsf(ma)) * 0x1p-20f;
  if (v > ma && v - ma <= tol) return ma;
  if (v < mi && mi - v <= tol) return mi;
  return v;
}

template <class TF, class TC, class P>
__device__ __forceinline__ void mono_extra(P f1, const TC* c, long ni, long o1, long o2, TC* lin, TC* mi, TC* ma) {
  mono_extra_a<TF, TC>(PtrStencil<P>{f1, ni, o1, o2, o1 + o2}, c, lin, mi, ma);
}

}  // namespace adz
