// Device tables of the bicubic-horizontal / quintic-vertical interpolation that adz_set
// builds with the context (adz_set.F90:181-190,286-307).
#pragma once
#include "adz_common.h"

namespace adz {

struct QvDev {
  const double* posz;  // Ver_z_8%t(0..nk+1)
  const double* dz;    // Adz_delz_t(0..nk)
  const double* odz;   // Adz_odelz_t(0..nk)
  const double* c4;    // [4][nk+1]: Adz_zabcd_8%t .. Adz_zdabc_8%t (k = 2..nk-2)
  const double* c6;    // [6][nk+1]: Adz_zxabcde_8%t .. Adz_zexabcd_8%t (k = 3..nk-3)
  int nk, offi, offj;  // offi = 1 - l_i0: global index to local
  int lminx, lmaxx, lminy, lmaxy;  // the extended layout
  int pos_dtype;
};

// The context's tables (false for a null context).
bool ctx_qv(const adz_ctx* c, QvDev* q);

}  // namespace adz
