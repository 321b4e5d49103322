// CPU ORACLE -- test infrastructure only. Only tests/, __graft_entry__.smoke() and
// bench.py's cpu_baseline leg may load liboracle.so, and only as the checker or the
// timed CPU baseline; the product (gem_amd/) never links it.
//
// extern "C" surface over oracle_interp.h / oracle_adz.h for ctypes.
#include <omp.h>

#include <cstdio>
#include <cstring>

#include "oracle_adz.h"

using namespace ora;

namespace {
template <class TP>
const TP* P3(const void* p, long i) { return (const TP*)p + 3 * i; }
}  // namespace

// threads of the OpenMP loops (the CPU baselines time all cores and one core)
extern "C" void ora_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }

// ---------------------------- vertical descriptor -----------------------------------
extern "C" void* ora_ztab_new(const double* levels, int nk, int ni, int nj, int offi, int offj) {
  return new Ztab(ztab_setup(levels, nk, ni, nj, offi, offj));
}
extern "C" void ora_ztab_free(void* h) { delete (Ztab*)h; }
extern "C" void ora_ztab_get(const void* h, double* z, double* ocz) {
  const Ztab* t = (const Ztab*)h;
  std::memcpy(z, t->z.data(), t->z.size() * 8);
  std::memcpy(ocz, t->ocz.data(), t->ocz.size() * 8);
}

// ------------------------------- index / weights ------------------------------------
// pdt: 1 = float positions (reference pxpypz), 0 = double positions.
extern "C" void ora_vcoef(const void* lv, const void* pos, int pdt, int n, int zabs_fma, int* ijk, int* zlin, double* coef) {
  const Ztab& t = *(const Ztab*)lv;
  for (long p = 0; p < n; p++) {
    if (pdt == 1) {
      const float* q = P3<float>(pos, p);
      zlin[p] = vcoef<float>(coef + 20 * p, ijk + p, q[0], q[1], q[2], t, zabs_fma);
    } else {
      const double* q = P3<double>(pos, p);
      zlin[p] = vcoef<double>(coef + 20 * p, ijk + p, q[0], q[1], q[2], t, zabs_fma);
    }
  }
}

// --------------------------------- interpolation ------------------------------------
// Tricublin_zyx1_m_n semantics (turbo_tricublin.c:793-815): d[f*n + p].  fdt: 1 = float
// fields (results narrowed to float, exactly the reference), 0 = double fields (double
// results).  fs[f] points at f(1,1,1) of field f.
template <class TF, class TP>
static void zyx1_m_n(TF* d, const TF* const* fs, const void* pos, const Ztab& t, long n, int m, int order, int zf) {
#pragma omp parallel for schedule(static)
  for (long p = 0; p < n; p++) {
    const TP* q = P3<TP>(pos, p);
    double c[24]; int off;
    int zl = vcoef<TP>(c, &off, q[0], q[1], q[2], t, zf);
    for (int f = 0; f < m; f++) d[(long)f * n + p] = (TF)tricub1<TF>(fs[f] + off, c, t.ni, t.nij, zl, order);
  }
}
extern "C" void ora_zyx1_m_n(void* d, const void* const* fs, int fdt, const void* pos, int pdt, const void* lv, long n, int m,
                  int order, int zabs_fma) {
  const Ztab& t = *(const Ztab*)lv;
  if (fdt == 1 && pdt == 1) zyx1_m_n<float, float>((float*)d, (const float* const*)fs, pos, t, n, m, order, zabs_fma);
  if (fdt == 1 && pdt == 0) zyx1_m_n<float, double>((float*)d, (const float* const*)fs, pos, t, n, m, order, zabs_fma);
  if (fdt == 0 && pdt == 1) zyx1_m_n<double, float>((double*)d, (const double* const*)fs, pos, t, n, m, order, zabs_fma);
  if (fdt == 0 && pdt == 0) zyx1_m_n<double, double>((double*)d, (const double* const*)fs, pos, t, n, m, order, zabs_fma);
}

// Tricublin_mono_zyx_m_n semantics (turbo_tricublin.c:1080-1109).
template <class TF, class TP>
static void mono_m_n(TF* d, TF* l, TF* mi, TF* ma, const TF* const* fs, const void* pos, const Ztab& t, long n, int m,
                     int order, int zf) {
#pragma omp parallel for schedule(static)
  for (long p = 0; p < n; p++) {
    const TP* q = P3<TP>(pos, p);
    double c[24]; int off;
    int zl = vcoef<TP>(c, &off, q[0], q[1], q[2], t, zf);
    for (int f = 0; f < m; f++) {
      double a, b, e, g;
      tricub_mono<TF>(fs[f] + off, c, t.ni, t.nij, zl, order, &a, &b, &e, &g);
      long o = (long)f * n + p;
      d[o] = (TF)a; l[o] = (TF)b; mi[o] = (TF)e; ma[o] = (TF)g;
    }
  }
}
extern "C" void ora_mono_m_n(void* d, void* l, void* mi, void* ma, const void* const* fs, int fdt, const void* pos, int pdt,
                  const void* lv, long n, int m, int order, int zabs_fma) {
  const Ztab& t = *(const Ztab*)lv;
  if (fdt == 1 && pdt == 1)
    mono_m_n<float, float>((float*)d, (float*)l, (float*)mi, (float*)ma, (const float* const*)fs, pos, t, n, m, order, zabs_fma);
  if (fdt == 1 && pdt == 0)
    mono_m_n<float, double>((float*)d, (float*)l, (float*)mi, (float*)ma, (const float* const*)fs, pos, t, n, m, order, zabs_fma);
  if (fdt == 0 && pdt == 1)
    mono_m_n<double, float>((double*)d, (double*)l, (double*)mi, (double*)ma, (const double* const*)fs, pos, t, n, m, order, zabs_fma);
  if (fdt == 0 && pdt == 0)
    mono_m_n<double, double>((double*)d, (double*)l, (double*)mi, (double*)ma, (const double* const*)fs, pos, t, n, m, order, zabs_fma);
}

// Tricublin_zyx3_n semantics (turbo_tricublin.c:1298-1312): d(3,n) from f3(3,ni,nj,nk).
template <class TF, class TP>
static void zyx3_n(TF* d, const TF* f3, const void* pos, const Ztab& t, long n, int order, int zf) {
#pragma omp parallel for schedule(static)
  for (long p = 0; p < n; p++) {
    const TP* q = P3<TP>(pos, p);
    double c[24]; int off;
    int zl = vcoef<TP>(c, &off, q[0], q[1], q[2], t, zf);
    double o[3];
    tricub3<TF>(f3 + 3L * off, c, t.ni, t.nij, zl, order, o);
    for (int k = 0; k < 3; k++) d[3 * p + k] = (TF)o[k];
  }
}
extern "C" void ora_zyx3_n(void* d, const void* f3, int fdt, const void* pos, int pdt, const void* lv, long n, int order,
                int zabs_fma) {
  const Ztab& t = *(const Ztab*)lv;
  if (fdt == 1 && pdt == 1) zyx3_n<float, float>((float*)d, (const float*)f3, pos, t, n, order, zabs_fma);
  if (fdt == 1 && pdt == 0) zyx3_n<float, double>((float*)d, (const float*)f3, pos, t, n, order, zabs_fma);
  if (fdt == 0 && pdt == 1) zyx3_n<double, float>((double*)d, (const double*)f3, pos, t, n, order, zabs_fma);
  if (fdt == 0 && pdt == 0) zyx3_n<double, double>((double*)d, (const double*)f3, pos, t, n, order, zabs_fma);
}

// ------------------------------------ grid ------------------------------------------
extern "C" void* ora_grid_new(const GridDesc* d) {
  Grid* g = new Grid();
  adz_set(*d, g);
  return g;
}
extern "C" void ora_grid_free(void* h) { delete (Grid*)h; }
// ints: lminx lmaxx lminy lmaxy nit njt nij i0 in j0 jn i0u inu j0v jnv i0b inb j0b jnb k0 k0t k0m
//       pnz_m pnz_t
// dbls: iminposx imaxposx iminposy imaxposy sig ovdzm ovdzt vw5 vw6
extern "C" void ora_grid_info(const void* h, int* iv, double* dv) {
  const Grid& g = *(const Grid*)h;
  int a[] = {g.lminx, g.lmaxx, g.lminy, g.lmaxy, g.nit, g.njt, g.nij, g.i0, g.in, g.j0, g.jn, g.i0u,
             g.inu, g.j0v, g.jnv, g.i0b, g.inb, g.j0b, g.jnb, g.k0, g.k0t, g.k0m,
             (int)g.search_m.size(), (int)g.search_t.size()};
  std::memcpy(iv, a, sizeof(a));
  double b[] = {g.iminposx, g.imaxposx, g.iminposy, g.imaxposy, g.sig, g.ovdzm, g.ovdzt, g.vw5, g.vw6};
  std::memcpy(dv, b, sizeof(b));
}
// which: 0 search_m, 1 search_t (int); 2..9 vw1m..vw4m, vw1t..vw4t (nk+1 doubles, index k);
// 10 cy (njt); 11 odelz_m, 12 odelz_t (nk+1)
extern "C" void ora_grid_table(const void* h, int which, void* out) {
  const Grid& g = *(const Grid*)h;
  const std::vector<double>* v[] = {nullptr, nullptr, &g.vw1m, &g.vw2m, &g.vw3m, &g.vw4m, &g.vw1t, &g.vw2t,
                                    &g.vw3t, &g.vw4t, &g.cy, &g.odelz_m, &g.odelz_t};
  if (which == 0) std::memcpy(out, g.search_m.data(), g.search_m.size() * 4);
  else if (which == 1) std::memcpy(out, g.search_t.data(), g.search_t.size() * 4);
  else std::memcpy(out, v[which]->data(), v[which]->size() * 8);
}
extern "C" void ora_level_search(const void* h, int which, long n, const double* pos, int* kk) {
  const Grid& g = *(const Grid*)h;
  for (long p = 0; p < n; p++)
    kk[p] = which == 0 ? level_search(g.zm, g.search_m, g.ovdzm, g.sig, g.nk, pos[p])
                       : level_search(g.zt, g.search_t, g.ovdzt, g.sig, g.nk, pos[p]);
}
extern "C" void ora_zindx(const void* h, long n, const double* pos, double* out) {
  const Grid& g = *(const Grid*)h;
  for (long p = 0; p < n; p++) out[p] = zindx(g, pos[p]);
}

// --------------------------------- trajectory ---------------------------------------
// smode: 1 = reference storage (float), 0 = FP64 storage.  pxyzm / wpxyz are in/out
// history (caller owns); if init != 0 adz_inittraj seeds them first.  Outputs are copied
// to caller buffers of the matching storage type (null pointers are skipped).
template <class S>
static void traj_run(const Grid& g, int init, double dt, int advec, double bA_m, double bA, int niter, int order,
                     const float* uu, const float* vv, const float* ww, const float* ut0, const float* vt0,
                     const float* zdt0, S* pxyzm, double* wpxyz, S* uvw_dep, S* pm, S* pmu, S* pmv, S* pt, S* pb,
                     S* uvw_d, S* warr, int klo, int khi) {
  TrajState<S> s;
  alloc_state<S>(g, &s);
  if (init) inittraj<S>(g, &s);
  else {
    std::memcpy(s.pxyzm.data(), pxyzm, s.pxyzm.size() * sizeof(S));
    std::memcpy(s.wpxyz.data(), wpxyz, s.wpxyz.size() * 8);
  }
  traject<S>(g, dt, advec, bA_m, bA, niter, order, uu, vv, ww, ut0, vt0, zdt0, &s, klo, khi);
  auto out = [](auto* dst, const auto& v) { if (dst) std::memcpy(dst, v.data(), v.size() * sizeof(v[0])); };
  out(pxyzm, s.pxyzm); out(wpxyz, s.wpxyz); out(uvw_dep, s.uvw_dep); out(pm, s.pm); out(pmu, s.pmu);
  out(pmv, s.pmv); out(pt, s.pt); out(pb, s.pb); out(uvw_d, s.uvw_d);
  if (warr) {
    size_t n = s.uu_arr.size();
    std::memcpy(warr, s.uu_arr.data(), n * sizeof(S));
    std::memcpy(warr + n, s.vv_arr.data(), n * sizeof(S));
    std::memcpy(warr + 2 * n, s.ww_arr.data(), n * sizeof(S));
  }
}
extern "C" void ora_traject(const void* h, int smode, int init, double dt, int advec, double bA_m, double bA, int niter,
                 int order, const float* uu, const float* vv, const float* ww, const float* ut0, const float* vt0,
                 const float* zdt0, void* pxyzm, double* wpxyz, void* uvw_dep, void* pm, void* pmu, void* pmv,
                 void* pt, void* pb, void* uvw_d, void* warr, int klo, int khi) {
  const Grid& g = *(const Grid*)h;
  if (smode == 1)
    traj_run<float>(g, init, dt, advec, bA_m, bA, niter, order, uu, vv, ww, ut0, vt0, zdt0, (float*)pxyzm, wpxyz,
                    (float*)uvw_dep, (float*)pm, (float*)pmu, (float*)pmv, (float*)pt, (float*)pb, (float*)uvw_d,
                    (float*)warr, klo, khi);
  else
    traj_run<double>(g, init, dt, advec, bA_m, bA, niter, order, uu, vv, ww, ut0, vt0, zdt0, (double*)pxyzm, wpxyz,
                     (double*)uvw_dep, (double*)pm, (double*)pmu, (double*)pmv, (double*)pt, (double*)pb,
                     (double*)uvw_d, (double*)warr, klo, khi);
}

// -------------------------------- Bermejo-Conde -------------------------------------
extern "C" int ora_almost_zero(double a) { return almost_zero(a) ? 1 : 0; }
extern "C" double ora_fsum(const double* v, long n) { return fsum(std::vector<double>(v, v + n)); }

// dom: minx maxx miny maxy nk i0 in j0 jn k0 min_max sum_mode
template <class T>
static void bc_run(const int* dom, int ntr, const void* const* p, void* const* m, const void* const* lin,
                   const void* const* mn, const void* const* mx, const int* legacy, const int* weight,
                   const int* pexp_n, double* deficit, const void* airm1, const void* airm0, const void* pkps,
                   const double* mask, const float* pexp_list, double* diag, const double* flux) {
  BCDomain D{dom[0], dom[1], dom[2], dom[3], dom[4], dom[5], dom[6], dom[7], dom[8], dom[9], mask, dom[10],
             pexp_list, dom[11], dom[12]};
  std::vector<BCTracer<T>> tr(ntr);
  for (int n = 0; n < ntr; n++)
    tr[n] = BCTracer<T>{(const T*)p[n], (T*)m[n], (const T*)lin[n], (const T*)mn[n], (const T*)mx[n],
                        legacy[n], weight[n], pexp_n[n], deficit[n]};
  std::vector<std::vector<T>> w(ntr);
  bermejo_conde<T>(D, tr, (const T*)airm1, (const T*)airm0, (const T*)pkps, w, diag, flux);
  for (int n = 0; n < ntr; n++) deficit[n] = tr[n].mass_deficit;
}
extern "C" void ora_bermejo_conde(int tdt, const int* dom, int ntr, const void* const* p, void* const* m,
                       const void* const* lin, const void* const* mn, const void* const* mx, const int* legacy,
                       const int* weight, const int* pexp_n, double* deficit, const void* airm1, const void* airm0,
                       const void* pkps, const double* mask, const float* pexp_list, double* diag,
                       const double* flux) {
  if (tdt == 1)
    bc_run<float>(dom, ntr, p, m, lin, mn, mx, legacy, weight, pexp_n, deficit, airm1, airm0, pkps, mask, pexp_list,
                  diag, flux);
  else
    bc_run<double>(dom, ntr, p, m, lin, mn, mx, legacy, weight, pexp_n, deficit, airm1, airm0, pkps, mask, pexp_list,
                   diag, flux);
}
// dom: minx maxx miny maxy nk i0 in j0 jn k0 min_max sum_mode autobar
extern "C" double ora_mass(int tdt, const int* dom, const void* x, const void* air, const double* mask) {
  BCDomain D{dom[0], dom[1], dom[2], dom[3], dom[4], dom[5], dom[6], dom[7], dom[8], dom[9], mask, 0, nullptr, dom[11],
             dom[12]};
  if (tdt == 1) return mass_of(D, (const float*)x, (const float*)air);
  return mass_of(D, (const double*)x, (const double*)air);
}


// -------------------------------- bicubic H / quintic V (BICHQV) --------------------------
#include "oracle_qv.h"

extern "C" void* ora_qv_new(const double* zt, int nk) { return new QvTab(qv_setup(zt, nk)); }
extern "C" void ora_qv_free(void* h) { delete (QvTab*)h; }

// bicubHQV_lin_min_max_ntr (bicubHQV_lin_mono.F90:20-172): fs = array starts (minx, miny, 1)
// of the nf fields; dom = minx maxx miny maxy ioff joff; outputs [f*n + p].
template <class TP>
static void qv_run(const QvTab& T, const float* const* fs, int nf, const TP* pos, long n, const int* dom, float* d,
                   float* l, float* mi, float* ma) {
  const long nit = dom[1] - dom[0] + 1, nij = nit * (dom[3] - dom[2] + 1);
#pragma omp parallel for schedule(static)
  for (long p = 0; p < n; p++) {
    double cx[4], cy[4], qz[6], lx[2], ly[2], lz[2];
    bool zq, zc;
    long o = qv_coef<TP>(pos[3 * p], pos[3 * p + 1], pos[3 * p + 2], T, dom[0], dom[2], nit, nij, dom[4], dom[5], cx,
                         cy, qz, lx, ly, lz, &zq, &zc);
    for (int f = 0; f < nf; f++)
      qv_zyx(d + (long)f * n + p, l + (long)f * n + p, mi + (long)f * n + p, ma + (long)f * n + p, fs[f] + o, cx, cy,
             qz, lx, ly, lz, nit, nij, zq, zc);
  }
}
extern "C" void ora_qv_m_n(const void* h, const void* const* fs, int nf, const void* pos, int pdt, long n,
                           const int* dom, float* d, float* l, float* mi, float* ma) {
  const QvTab& T = *(const QvTab*)h;
  if (pdt == 1) qv_run<float>(T, (const float* const*)fs, nf, (const float*)pos, n, dom, d, l, mi, ma);
  else qv_run<double>(T, (const float* const*)fs, nf, (const double*)pos, n, dom, d, l, mi, ma);
}
// the weights of one target (for the pins): cx cy qz lx ly lz (20 doubles), flags, offset
extern "C" long ora_qv_coef(const void* h, const float* pos, const int* dom, double* w, int* flags) {
  const QvTab& T = *(const QvTab*)h;
  const long nit = dom[1] - dom[0] + 1, nij = nit * (dom[3] - dom[2] + 1);
  bool zq, zc;
  long o = qv_coef<float>(pos[0], pos[1], pos[2], T, dom[0], dom[2], nit, nij, dom[4], dom[5], w, w + 4, w + 8,
                          w + 14, w + 16, w + 18, &zq, &zc);
  flags[0] = zq; flags[1] = zc;
  return o;
}
extern "C" void ora_qv_zyx(float* out4, const float* src, const double* w, long ni, long ninj, int zq, int zc) {
  qv_zyx(out4, out4 + 1, out4 + 2, out4 + 3, src, w, w + 4, w + 8, w + 14, w + 16, w + 18, ni, ninj, zq != 0,
         zc != 0);
}

// -------------------------------- ILMC (mono = 2) -----------------------------------------
#include "oracle_ilmc.h"

// dom = minx maxx miny maxy nk i0 in j0 jn k0 sweep_max min_max autobar
extern "C" void ora_ilmc(float* f, const float* mn, const float* mx, const float* air, const int* dom,
                         long long* counts) {
  IlmcDom d{dom[0], dom[1], dom[2], dom[3], dom[4], dom[5], dom[6], dom[7], dom[8], dom[9], dom[10], dom[11], dom[12]};
  ilmc_lam(f, mn, mx, air, d, counts);
}
// the cells of one sweep (array offsets, reference order); returns their number
extern "C" long ora_ilmc_shell(const int* dom, int i, int j, int k, int s, long* out) {
  IlmcDom d{dom[0], dom[1], dom[2], dom[3], dom[4], dom[5], dom[6], dom[7], dom[8], dom[9], dom[10], dom[11], dom[12]};
  std::vector<long> sh;
  ilmc_shell(d, i, j, k, s, sh);
  for (size_t q = 0; q < sh.size(); q++) out[q] = sh[q];
  return (long)sh.size();
}
