// CPU ORACLE -- test infrastructure only. Nothing in the product (gem_amd/, the HIP
// library, bench timing legs other than cpu_baseline) may link or call this code.
//
// Restatement of the index/weight step and the tricubic Lagrange interpolators of
// GEM's turbo_tricublin.c (src/gemdyn/src/base/turbo_tricublin.c), written from the
// reference's documented behaviour, templated on
//   TF : field storage type (float = reference; double = plain FP64 form)
//   TP : departure-position type (float = reference pxpypz; double = FP64 positions)
// Arithmetic is always FP64 (as in the reference, :561-627). Two operation orders:
//   ORD_AVX2   -- the reference's __AVX2__ path (mul + 3 fma per z column, y by fma,
//                 x by (y0c0+y2c2)+(y1c1+y3c3)), :553-615, :874-989, :1142-1257
//   ORD_SCALAR -- the reference's portable path, left-to-right sums, :617-627, :990-1026
// Compile with -ffp-contract=off: every fused multiply-add below is an explicit fma().
#pragma once
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <vector>

namespace ora {

enum { ORD_AVX2 = 0, ORD_SCALAR = 1 };

// Vertical descriptor; restates the reference's `ztab` (turbo_tricublin.c:164-173) and
// Vsearch_setup / Vsearch_setup_plus (turbo_tricublin.c:320-380).
struct Ztab {
  std::vector<double> z;    // nk levels
  std::vector<double> ocz;  // 4*max(nk,4) inverse Lagrange denominators
  int ni, nj, nk, nij, offi, offj;
};

// TRIPRD / denominators(): turbo_tricublin.c:250-257
inline double triprd(double x, double a, double b, double c) { return (x - a) * (x - b) * (x - c); }
inline void denominators(double* r, double a, double b, double c, double d) {
  r[0] = 1.0 / triprd(a, b, c, d);
  r[1] = 1.0 / triprd(b, a, c, d);
  r[2] = 1.0 / triprd(c, a, b, d);
  r[3] = 1.0 / triprd(d, a, b, c);
}

// turbo_tricublin.c:320-363 (+ :368-380 for the offsets)
inline Ztab ztab_setup(const double* levels, int nk, int ni, int nj, int offi, int offj) {
  Ztab t;
  int nkl = nk > 4 ? nk : 4;
  t.z.assign(levels, levels + nk);
  t.ocz.assign(4 * nkl, 1.0);
  if (nk >= 4) {
    denominators(&t.ocz[0], levels[0], levels[1], levels[2], levels[3]);
    for (int i = 1; i < nk - 2; i++) denominators(&t.ocz[4 * i], t.z[i - 1], t.z[i], t.z[i + 1], t.z[i + 2]);
    denominators(&t.ocz[4 * (nk - 2)], levels[nk - 4], levels[nk - 3], levels[nk - 2], levels[nk - 1]);
    denominators(&t.ocz[4 * (nk - 1)], levels[nk - 4], levels[nk - 3], levels[nk - 2], levels[nk - 1]);
  }
  t.ni = ni; t.nj = nj; t.nk = nk; t.nij = ni * nj; t.offi = offi; t.offj = offj;
  return t;
}

// Index + weights for one target point; restates Vcoef_pxyz4_inline (non -DSIMD branch),
// turbo_tricublin.c:382-512.  c[0..3] cubic x, c[4..7] cubic y, c[8..11] z (cubic or
// linear), c[12..15] linear x (0,1-dx,dx,0), c[16..17] linear y, c[18..19] linear z.
// Returns the zlinear flag; *off receives the element offset of the stencil's lower
// left corner relative to f(1,1,1).
// `zabs_fma` selects how the absolute z of the target is formed:
//   0 : dz*z1 + (1-dz)*z0 with separate roundings (the source expression, :438)
//   1 : fma(dz, z1, (1-dz)*z0)   2 : fma(1-dz, z0, dz*z1)
// (which one a C compiler emits when contracting is fixed by the oracle/_ref checks.)
template <class TP>
inline int vcoef(double* c, int* off, TP px8, TP py8, TP pz8, const Ztab& lv, int zabs_fma) {
  const double one = 1.0, two = 2.0, cp167 = 1.0 / 6.0, cm167 = -1.0 / 6.0, cp5 = 0.5, cm5 = -0.5;
  double px = (double)px8, py = (double)py8, pz = (double)pz8;
  int ix = (int)px;                 // truncation (C conversion), == floor on clamped domain
  px = px - ix;
  ix = ix + lv.offi;
  int ijk = ix - 2;
  c[12] = 0.0; c[13] = 1.0 - px; c[14] = px; c[15] = 0.0;
  int iy = (int)py;
  py = py - iy;
  iy = iy + lv.offj;
  ijk = ijk + (iy - 2) * lv.ni;
  c[16] = 1.0 - py; c[17] = py;
  int iz = (int)pz;
  if (iz < 1) iz = 1;
  if (iz > lv.nk - 1) iz = lv.nk - 1;
  double dz = pz - iz;
  ijk = ijk + (iz - 1) * lv.nij;
  c[18] = 1.0 - dz; c[19] = dz;
  iz--;
  int zlinear = (iz - 1) | (lv.nk - 3 - iz);
  zlinear >>= 31;
  if (!zlinear) ijk = ijk - lv.nij;
  *off = ijk;
  if (zlinear) {
    c[8] = 0.0; c[9] = 1.0 - dz; c[10] = dz; c[11] = 0.0;
  } else {
    const double* base = &lv.ocz[4 * iz];
    const double* pos = &lv.z[iz];
    double za;
    if (zabs_fma == 1) za = std::fma(dz, pos[1], (1.0 - dz) * pos[0]);
    else if (zabs_fma == 2) za = std::fma(1.0 - dz, pos[0], dz * pos[1]);
    else za = dz * pos[1] + (1.0 - dz) * pos[0];
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
  return zlinear;
}

// Tricublin_zyxf1_inline, turbo_tricublin.c:525-629.  Returns the FP64 result; the
// reference narrows it to float on store.
template <class TF>
inline double tricub1(const TF* f1, const double* c, int ni, int ninj, int zlinear, int order) {
  const double *px = c, *py = c + 4, *pz = c + 8;
  int ninjl = zlinear ? 0 : ninj;
  const long o1 = ninjl, o2 = (long)ninjl + ninj, o3 = (long)ninjl + ninj + ninjl;
  if (order == ORD_AVX2) {
    double ya[4];
    for (int r = 0; r < 4; r++) {
      const TF* s = f1 + (long)r * ni;
      for (int i = 0; i < 4; i++) {
        double za = (double)s[i] * pz[0];
        za = std::fma((double)s[i + o1], pz[1], za);
        za = std::fma((double)s[i + o2], pz[2], za);
        za = std::fma((double)s[i + o3], pz[3], za);
        ya[i] = (r == 0) ? za * py[0] : std::fma(za, py[r], ya[i]);
      }
    }
    double y0 = ya[0] * px[0], y1 = ya[1] * px[1], y2 = ya[2] * px[2], y3 = ya[3] * px[3];
    return (y0 + y2) + (y1 + y3);
  }
  double dst[4];
  for (int i = 0; i < 4; i++) {
    double v[4];
    for (int r = 0; r < 4; r++) {
      const TF* s = f1 + (long)r * ni;
      v[r] = (double)s[i] * pz[0] + (double)s[i + o1] * pz[1] + (double)s[i + o2] * pz[2] + (double)s[i + o3] * pz[3];
    }
    dst[i] = v[0] * py[0] + v[1] * py[1] + v[2] * py[2] + v[3] * py[3];
  }
  return dst[0] * px[0] + dst[1] * px[1] + dst[2] * px[2] + dst[3] * px[3];
}

// min/max with the operand rule of both reference paths: MINPD/MAXPD return the second
// operand unless the first is strictly smaller/larger (:899-980), and the portable MIN/MAX
// macros (:111-112) are the same ternaries.  Keeps the sign of zero bit-exact.
inline double omin(double a, double b) { return (a < b) ? a : b; }
inline double omax(double a, double b) { return (a > b) ? a : b; }

// Tricublin_zyx_mm_d_inline, turbo_tricublin.c:838-1027: cubic result plus the linear
// interpolation on the inner cells and the min/max.  NOTE the two reference paths differ
// in the min/max footprint: the AVX2 path folds rows 0..3 (x columns 1..2, planes 1..2,
// 16 values, :883-980) while the portable path uses rows 1..2 only (8 values, :1017-1025).
template <class TF>
inline void tricub_mono(const TF* f1, const double* c, int ni, int ninj, int zlinear, int order,
                        double* d, double* lin, double* mi, double* ma) {
  *d = tricub1(f1, c, ni, ninj, zlinear, order);
  int ninjl = zlinear ? 0 : ninj;
  const long o1 = ninjl, o2 = (long)ninjl + ninj;
  if (order == ORD_AVX2) {
    // min / max: 4-wide vectors over i = 0..3, folded to lanes 1 and 2 only
    double vmi[4], vma[4];
    for (int r = 0; r < 4; r++) {
      const TF* s = f1 + (long)r * ni;
      for (int i = 0; i < 4; i++) {
        double a = (double)s[i + o1], b = (double)s[i + o2];
        if (r == 0) { vmi[i] = a; vma[i] = a; }
        else { vmi[i] = omin(vmi[i], a); vma[i] = omax(vma[i], a); }
        vmi[i] = omin(vmi[i], b); vma[i] = omax(vma[i], b);
      }
    }
    *mi = omin(vmi[1], vmi[2]);
    *ma = omax(vma[1], vma[2]);
    // linear: rows 1 and 2, planes 1 and 2, then x with (0,1-dx,dx,0)
    double vyl[4];
    for (int i = 0; i < 4; i++) {
      const TF* s1 = f1 + ni;
      const TF* s2 = f1 + 2L * ni;
      double vz1 = (double)s1[i + o1] * c[18];
      vz1 = std::fma((double)s1[i + o2], c[19], vz1);
      double vy = vz1 * c[16];
      double vz2 = (double)s2[i + o1] * c[18];
      vz2 = std::fma((double)s2[i + o2], c[19], vz2);
      vyl[i] = std::fma(vz2, c[17], vy);
    }
    double v0 = vyl[0] * c[12], v1 = vyl[1] * c[13], v2 = vyl[2] * c[14], v3 = vyl[3] * c[15];
    *lin = (v0 + v2) + (v1 + v3);
    return;
  }
  double dsl[3];
  for (int i = 1; i < 3; i++) {
    double vb = (double)f1[i + ni + o1] * c[18] + (double)f1[i + ni + o2] * c[19];
    double vc = (double)f1[i + 2L * ni + o1] * c[18] + (double)f1[i + 2L * ni + o2] * c[19];
    dsl[i] = vb * c[16] + vc * c[17];
  }
  *lin = dsl[1] * c[13] + dsl[2] * c[14];
  double mx = (double)f1[1 + ni + o1], mn = mx;
  for (int i = 1; i < 3; i++) {
    mx = omax(mx, (double)f1[i + ni + o1]); mx = omax(mx, (double)f1[i + ni + o2]);
    mn = omin(mn, (double)f1[i + ni + o1]); mn = omin(mn, (double)f1[i + ni + o2]);
    mx = omax(mx, (double)f1[i + 2L * ni + o1]); mx = omax(mx, (double)f1[i + 2L * ni + o2]);
    mn = omin(mn, (double)f1[i + 2L * ni + o1]); mn = omin(mn, (double)f1[i + 2L * ni + o2]);
  }
  *mi = mn; *ma = mx;
}

// Tricublin_zyxf3_inline, turbo_tricublin.c:1122-1287: 3 interleaved components
// f(3,ni,nj,nk).  The AVX2 x stage accumulates column by column with fma (:1243-1257),
// unlike the single-field fold.
template <class TF>
inline void tricub3(const TF* f, const double* c, int NI, int NINJ, int zlinear, int order, double* out) {
  const double *px = c, *py = c + 4, *pz = c + 8;
  const long ni = 3L * NI, ninj = 3L * NINJ;
  long ninjl = zlinear ? 0 : ninj;
  const long o1 = ninjl, o2 = ninjl + ninj, o3 = ninjl + ninj + ninjl;
  double dst[12];
  if (order == ORD_AVX2) {
    for (int e = 0; e < 12; e++) {
      double ya = 0.0;
      for (int r = 0; r < 4; r++) {
        const TF* s = f + r * ni;
        double za = (double)s[e] * pz[0];
        za = std::fma((double)s[e + o1], pz[1], za);
        za = std::fma((double)s[e + o2], pz[2], za);
        za = std::fma((double)s[e + o3], pz[3], za);
        ya = (r == 0) ? za * py[0] : std::fma(za, py[r], ya);
      }
      dst[e] = ya;
    }
    for (int q = 0; q < 3; q++) {
      double x0 = dst[q] * px[0];
      x0 = std::fma(dst[q + 3], px[1], x0);
      x0 = std::fma(dst[q + 6], px[2], x0);
      x0 = std::fma(dst[q + 9], px[3], x0);
      out[q] = x0;
    }
    return;
  }
  for (int e = 0; e < 12; e++) {
    double v[4];
    for (int r = 0; r < 4; r++) {
      const TF* s = f + r * ni;
      v[r] = (double)s[e] * pz[0] + (double)s[e + o1] * pz[1] + (double)s[e + o2] * pz[2] + (double)s[e + o3] * pz[3];
    }
    dst[e] = v[0] * py[0] + v[1] * py[1] + v[2] * py[2] + v[3] * py[3];
  }
  for (int q = 0; q < 3; q++) out[q] = dst[q] * px[0] + dst[q + 3] * px[1] + dst[q + 6] * px[2] + dst[q + 9] * px[3];
}

}  // namespace ora
