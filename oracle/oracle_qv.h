// CPU ORACLE -- test infrastructure only (tests/, smoke() and bench.py's CPU baseline).
//
// Bicubic horizontal / quintic vertical interpolation with the linear result, min and max
// of the inner 2x2x2 box (the BICHQV tracer groups):
//   - tables of adz_set.F90:181-190 (delz, odelz) and :286-307 (inverse Lagrange
//     denominators, cubic and quintic), restated in qv_setup;
//   - bicubHQV_lin_min_max_ntr (bicubHQV_lin_mono.F90:20-172): weights and stencil base of
//     each target, restated in qv_coef;
//   - bicubHQV_lin_min_max_zyx (bicubHQV_lin_min_max_zyx.F90:20-99): the evaluation,
//     restated in qv_zyx.
// Fortran evaluates each expression left to right in REAL64 (real operands widened), with
// the intermediate dst(0:3) and dsl(1:2) stored as `real`; no contraction into FMA.
#pragma once
#include <cmath>
#include <vector>

namespace ora {

struct QvTab {
  int nk = 0;
  std::vector<double> posz;   // Ver_z_8%t(0..nk+1)
  std::vector<double> dz;     // Adz_delz_t(0..nk)
  std::vector<double> odz;    // Adz_odelz_t(0..nk)
  std::vector<double> c4[4];  // Adz_zabcd_8%t .. Adz_zdabc_8%t, index k (2..nk-2 set)
  std::vector<double> c6[6];  // Adz_zxabcde_8%t .. Adz_zexabcd_8%t, index k (3..nk-3 set)
};

inline double qv_triprod(double za, double zb, double zc, double zd) { return ((za - zb) * (za - zc) * (za - zd)); }
inline double qv_quiprod(double zx, double za, double zb, double zc, double zd, double ze) {
  return ((zx - za) * (zx - zb) * (zx - zc) * (zx - zd) * (zx - ze));
}

// zt: Ver_z_8%t(0:nk+1)
inline QvTab qv_setup(const double* zt, int nk) {
  QvTab T;
  T.nk = nk;
  T.posz.assign(zt, zt + nk + 2);
  T.dz.assign(nk + 1, (double)9.e33f);  // Adz_delz_t = 9.e33 (a default real constant)
  T.odz.assign(nk + 1, 0.0);
  const float verysmall = 1.17549435e-38f * 1.e5f;  // tiny(real)*1.e5
  for (int k = 0; k <= nk; k++) {
    if (std::fabs(zt[k + 1] - zt[k]) > (double)verysmall) T.dz[k] = zt[k + 1] - zt[k];
    T.odz[k] = 1.0 / T.dz[k];
  }
  for (auto& v : T.c4) v.assign(nk + 1, 0.0);
  for (auto& v : T.c6) v.assign(nk + 1, 0.0);
  for (int k = 2; k <= nk - 2; k++) {
    double ra = zt[k - 1], rb = zt[k], rc = zt[k + 1], rd = zt[k + 2];
    T.c4[0][k] = 1.0 / qv_triprod(ra, rb, rc, rd);
    T.c4[1][k] = 1.0 / qv_triprod(rb, ra, rc, rd);
    T.c4[2][k] = 1.0 / qv_triprod(rc, ra, rb, rd);
    T.c4[3][k] = 1.0 / qv_triprod(rd, ra, rb, rc);
  }
  for (int k = 3; k <= nk - 3; k++) {
    double rx = zt[k - 2], ra = zt[k - 1], rb = zt[k], rc = zt[k + 1], rd = zt[k + 2], re = zt[k + 3];
    T.c6[0][k] = 1.0 / qv_quiprod(rx, ra, rb, rc, rd, re);
    T.c6[1][k] = 1.0 / qv_quiprod(ra, rx, rb, rc, rd, re);
    T.c6[2][k] = 1.0 / qv_quiprod(rb, rx, ra, rc, rd, re);
    T.c6[3][k] = 1.0 / qv_quiprod(rc, rx, ra, rb, rd, re);
    T.c6[4][k] = 1.0 / qv_quiprod(rd, rx, ra, rb, rc, re);
    T.c6[5][k] = 1.0 / qv_quiprod(re, rx, ra, rb, rc, rd);
  }
  return T;
}

// Weights of one target (bicubHQV_lin_mono.F90:67-149).  Returns the 0-based element
// offset of the stencil's first value (row iy-1, column ix-1, first plane) in an array of
// row pitch nit and plane pitch nij that starts at (minx, miny, 1); ioff / joff are
// Adz_ioff / Adz_joff (l_i0 - 2, l_j0 - 2).
template <class TP>
inline long qv_coef(TP px_, TP py_, TP pz_, const QvTab& T, int minx, int miny, long nit, long nij, int ioff,
                    int joff, double* cx, double* cy, double* qz, double* lx, double* ly, double* lz, bool* zq,
                    bool* zc) {
  const double ov6 = 1.0 / 6.0, ov2 = 1.0 / 2.0;
  const int nk = T.nk;
  int ii1 = (int)px_;
  double del = (double)px_ - (double)ii1;
  double del12 = (del - 1) * (del - 2);
  double del10 = (del + 1) * del;
  cx[0] = -(del * del12 * ov6);
  cx[1] = (del + 1) * del12 * ov2;
  cx[2] = -(del10 * (del - 2) * ov2);
  cx[3] = del10 * (del - 1) * ov6;
  lx[0] = 1.0 - del; lx[1] = del;
  int jj1 = (int)py_;
  del = (double)py_ - (double)jj1;
  del12 = (del - 1) * (del - 2);
  del10 = (del + 1) * del;
  cy[0] = -(del * del12 * ov6);
  cy[1] = (del + 1) * del12 * ov2;
  cy[2] = -(del10 * (del - 2) * ov2);
  cy[3] = del10 * (del - 1) * ov6;
  ly[0] = 1.0 - del; ly[1] = del;
  ii1 = ii1 - minx - 1;
  jj1 = jj1 - miny - 1;
  long indxij = (long)(jj1 - joff) * nit + (ii1 - ioff) - nit;
  int kk1 = (int)pz_;
  *zq = (kk1 >= 3) && (kk1 <= nk - 3);
  *zc = (kk1 == 2) || (kk1 == nk - 2);
  if (!*zq && !*zc) kk1 = std::min(nk - 1, std::max(1, kk1));
  const double* posz = T.posz.data();
  double zt = posz[kk1] + ((double)pz_ - (double)kk1) * T.dz[kk1];
  lz[1] = (zt - posz[kk1]) * T.odz[kk1];
  lz[0] = 1.0 - lz[1];
  long o1;
  if (*zq) {
    o1 = (long)(kk1 - 3) * nij + indxij;
    double rx = posz[kk1 - 2], ra = posz[kk1 - 1], rb = posz[kk1], rc = posz[kk1 + 1], rd = posz[kk1 + 2],
           re = posz[kk1 + 3];
    qz[0] = qv_quiprod(zt, ra, rb, rc, rd, re) * T.c6[0][kk1];
    qz[1] = qv_quiprod(zt, rx, rb, rc, rd, re) * T.c6[1][kk1];
    qz[2] = qv_quiprod(zt, rx, ra, rc, rd, re) * T.c6[2][kk1];
    qz[3] = qv_quiprod(zt, rx, ra, rb, rd, re) * T.c6[3][kk1];
    qz[4] = qv_quiprod(zt, rx, ra, rb, rc, re) * T.c6[4][kk1];
    qz[5] = qv_quiprod(zt, rx, ra, rb, rc, rd) * T.c6[5][kk1];
  } else if (*zc) {
    o1 = (long)(kk1 - 2) * nij + indxij;
    double ra = posz[kk1 - 1], rb = posz[kk1], rc = posz[kk1 + 1], rd = posz[kk1 + 2];
    qz[0] = 0.;
    qz[1] = qv_triprod(zt, rb, rc, rd) * T.c4[0][kk1];
    qz[2] = qv_triprod(zt, ra, rc, rd) * T.c4[1][kk1];
    qz[3] = qv_triprod(zt, ra, rb, rd) * T.c4[2][kk1];
    qz[4] = qv_triprod(zt, ra, rb, rc) * T.c4[3][kk1];
    qz[5] = 0.;
  } else {
    o1 = (long)(kk1 - 1) * nij + indxij;
    qz[0] = 0.;
    qz[1] = 0.;
    qz[3] = (zt - posz[kk1]) * T.odz[kk1];
    qz[2] = 1.0 - qz[3];
    qz[4] = 0.;
    qz[5] = 0.;
  }
  return o1 - 1;  // F_in(o1) is element o1 (1-based)
}

inline float qv_max(float a, float b) { return (b > a) ? b : a; }
inline float qv_min(float a, float b) { return (b < a) ? b : a; }

// bicubHQV_lin_min_max_zyx.F90:20-99; src points at the stencil's first value.
inline void qv_zyx(float* xyz, float* li, float* mi, float* ma, const float* src, const double* cx,
                   const double* cy, const double* qz, const double* lx, const double* ly, const double* lz, long ni,
                   long ninj, bool zq, bool zc) {
  const long ninjc = zq ? ninj : 0;
  const long ninjl = (zq || zc) ? ninj : 0;
  const long ninj1 = 0 + ninjc, ninj2 = ninj1 + ninjl, ninj3 = ninj2 + ninj, ninj4 = ninj3 + ninjl,
             ninj5 = ninj4 + ninjc;
  const long ni2 = ni + ni, ni3 = ni2 + ni;
  auto S = [&](long q) { return (double)src[q]; };
  auto col = [&](long b) {
    return S(b) * qz[0] + S(b + ninj1) * qz[1] + S(b + ninj2) * qz[2] + S(b + ninj3) * qz[3] + S(b + ninj4) * qz[4] +
           S(b + ninj5) * qz[5];
  };
  float dst[4], dsl[3];
  for (int i = 0; i <= 3; i++) {
    double va4 = col(i), vb4 = col(i + ni), vc4 = col(i + ni2), vd4 = col(i + ni3);
    dst[i] = (float)(va4 * cy[0] + vb4 * cy[1] + vc4 * cy[2] + vd4 * cy[3]);
  }
  *xyz = (float)((double)dst[0] * cx[0] + (double)dst[1] * cx[1] + (double)dst[2] * cx[2] + (double)dst[3] * cx[3]);
  for (int i = 1; i <= 2; i++) {
    double vb4 = S(i + ni + ninj2) * lz[0] + S(i + ni + ninj3) * lz[1];
    double vc4 = S(i + ni2 + ninj2) * lz[0] + S(i + ni2 + ninj3) * lz[1];
    dsl[i] = (float)(vb4 * ly[0] + vc4 * ly[1]);
  }
  *li = (float)((double)dsl[1] * lx[0] + (double)dsl[2] * lx[1]);
  float ma_ = src[1 + ni + ninj2], mi_ = ma_;
  for (int i = 1; i <= 2; i++) {
    ma_ = qv_max(ma_, src[i + ni + ninj2]); ma_ = qv_max(ma_, src[i + ni + ninj3]);
    mi_ = qv_min(mi_, src[i + ni + ninj2]); mi_ = qv_min(mi_, src[i + ni + ninj3]);
    ma_ = qv_max(ma_, src[i + ni2 + ninj2]); ma_ = qv_max(ma_, src[i + ni2 + ninj3]);
    mi_ = qv_min(mi_, src[i + ni2 + ninj2]); mi_ = qv_min(mi_, src[i + ni2 + ninj3]);
  }
  *ma = ma_;
  *mi = mi_;
}

}  // namespace ora
