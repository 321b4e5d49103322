// Fused multi-field interpolation at departure points (adz_tricub_hlt,
// src/gemdyn/src/base/adz_interp_hlt.F90:28-187) and its batched, persistent form over
// several subdomains.
//
// Work is cut into tiles of 256 consecutive target points (i fastest) times a group of
// `fields_per_pass` fields.  One block of 256 lanes runs one tile: each lane computes the
// index/weights of its target once and interpolates the group's fields, writing straight
// into the destination arrays (the reference interpolates into a work buffer and scatters
// it, :165-180).  A persistent grid walks the tile list; with `xcd_chunks` the tile list is
// cut into 8 contiguous ranges, one per XCD (blocks are dispatched round-robin over the 8
// XCDs), so each XCD's L2 sees spatially adjacent tiles.
#include <algorithm>
#include <cstring>
#include <vector>

#include "adz_common.h"
#include "adz_eval.h"

namespace adz {

constexpr int kTile = 256;
constexpr int kInlineFields = 16;

struct DevItem {
  ZtabDev lv;
  const void* xyz;
  int i0, j0, k0, nx, ny, nz;
  int dni, dnj;
  int nf;
  long npts, ptiles, ngroups, tile0;
  const void* const* src;
  void* const* dst;
  void* const* lin;
  void* const* mn;
  void* const* mx;
};

struct PlanArgs {
  const DevItem* items;
  int nitems;
  long ntiles;
  int g;
  int points_outer;
  int xcd_chunks;
};

struct InlineFields {
  const void* src[kInlineFields];
  void* dst[kInlineFields];
  void* lin[kInlineFields];
  void* mn[kInlineFields];
  void* mx[kInlineFields];
};

// One tile: target points [ptile*256, ptile*256+256) of `it`, fields [f0, f1).
template <class TF, class TP, int MATH, int POST, class Fields>
__device__ __forceinline__ void run_tile(const DevItem& it, const Fields& F, long ptile, int f0, int f1) {
  long p = ptile * kTile + threadIdx.x;
  if (p >= it.npts) return;
  int i = it.i0 + (int)(p % it.nx);
  long r = p / it.nx;
  int j = it.j0 + (int)(r % it.ny);
  int k = it.k0 + (int)(r / it.ny);
  long o = (long)(i - 1) + (long)it.dni * ((j - 1) + (long)it.dnj * (k - 1));  // dst(i,j,k) from dst(1,1,1)
  double px, py, pz;
  load_pos<TP>(it.xyz, p, &px, &py, &pz);
  Index x;
  const bool ok = index_step(px, py, pz, it.lv, &x);
  const long ni = it.lv.ni;
  const long o1 = x.zlinear ? 0 : (long)it.lv.nij;
  const long o2 = o1 + it.lv.nij, o3 = o2 + o1;
  if (MATH == ADZ_MATH_REF) {
    double c[20];
    if (ok) weights_ref(x, it.lv, c);
    for (int f = f0; f < f1; f++) {
      TF* d = (TF*)F.dst[f];
      if (!ok) {
        TF nan = from_f64<TF>(__builtin_nan(""));
        d[o] = nan;
        if (POST & ADZ_POST_STORE) { ((TF*)F.lin[f])[o] = nan; ((TF*)F.mn[f])[o] = nan; ((TF*)F.mx[f])[o] = nan; }
        continue;
      }
      const TF* f1p = (const TF*)F.src[f] + x.off;
      double v = tricub_ref<TF>(f1p, c, ni, o1, o2, o3);
      if (POST) {
        double l, mi, ma;
        mono_extra<TF, double>(f1p, c, ni, o1, o2, &l, &mi, &ma);
        // adz_interp_hlt.F90:158-160 narrows each to real before clipping lin
        TF ls = from_f64<TF>(l), mis = from_f64<TF>(mi), mas = from_f64<TF>(ma);
        TF vs = from_f64<TF>(v);
        if (POST & ADZ_POST_CLIP) vs = from_f64<TF>(fmax(to_f64(mis), fmin(to_f64(mas), to_f64(vs))));
        d[o] = vs;
        if (POST & ADZ_POST_STORE) {
          ((TF*)F.lin[f])[o] = from_f64<TF>(fmax(to_f64(mis), fmin(to_f64(mas), to_f64(ls))));
          ((TF*)F.mn[f])[o] = mis;
          ((TF*)F.mx[f])[o] = mas;
        }
      } else {
        d[o] = from_f64<TF>(v);
      }
    }
  } else {
    float c[20];
    if (ok) weights_f32(x, it.lv, c);
    for (int f = f0; f < f1; f++) {
      TF* d = (TF*)F.dst[f];
      if (!ok) {
        TF nan = from_f32<TF>(__builtin_nanf(""));
        d[o] = nan;
        if (POST & ADZ_POST_STORE) { ((TF*)F.lin[f])[o] = nan; ((TF*)F.mn[f])[o] = nan; ((TF*)F.mx[f])[o] = nan; }
        continue;
      }
      const TF* f1p = (const TF*)F.src[f] + x.off;
      float v = tricub_f32<TF>(f1p, c, ni, o1, o2, o3);
      if (POST) {
        float l, mi, ma;
        mono_extra<TF, float>(f1p, c, ni, o1, o2, &l, &mi, &ma);
        TF ls = from_f32<TF>(l), mis = from_f32<TF>(mi), mas = from_f32<TF>(ma);
        TF vs = from_f32<TF>(v);
        if (POST & ADZ_POST_CLIP) vs = from_f32<TF>(fmaxf(to_f32(mis), fminf(to_f32(mas), to_f32(vs))));
        d[o] = vs;
        if (POST & ADZ_POST_STORE) {
          ((TF*)F.lin[f])[o] = from_f32<TF>(fmaxf(to_f32(mis), fminf(to_f32(mas), to_f32(ls))));
          ((TF*)F.mn[f])[o] = mis;
          ((TF*)F.mx[f])[o] = mas;
        }
      } else {
        d[o] = from_f32<TF>(v);
      }
    }
  }
}

struct ItemFields {  // field pointer tables of a plan item (device arrays)
  const void* const* src;
  void* const* dst;
  void* const* lin;
  void* const* mn;
  void* const* mx;
};

// Persistent kernel over the tile list of a plan.
template <class TF, class TP, int MATH, int POST>
__global__ __launch_bounds__(kTile) void k_plan(PlanArgs A) {
  long lo = 0, hi = A.ntiles, step = gridDim.x, t0 = blockIdx.x;
  if (A.xcd_chunks) {
    int nb = gridDim.x / 8;                 // blocks per XCD (grid is a multiple of 8)
    int x = blockIdx.x % 8, l = blockIdx.x / 8;
    lo = A.ntiles * x / 8;
    hi = A.ntiles * (x + 1) / 8;
    step = nb;
    t0 = lo + l;
  }
  for (long t = t0; t < hi; t += step) {
    int q = 0;
    while (q + 1 < A.nitems && A.items[q + 1].tile0 <= t) q++;
    const DevItem& it = A.items[q];
    long local = t - it.tile0;
    long ptile, grp;
    if (A.points_outer) { ptile = local / it.ngroups; grp = local % it.ngroups; }
    else { grp = local / it.ptiles; ptile = local % it.ptiles; }
    int f0 = (int)grp * A.g, f1 = min(it.nf, f0 + A.g);
    ItemFields F{it.src, it.dst, it.lin, it.mn, it.mx};
    run_tile<TF, TP, MATH, POST, ItemFields>(it, F, ptile, f0, f1);
  }
}

// One item with inline field pointers (the adz_tricub_hlt entry point).
template <class TF, class TP, int MATH, int POST>
__global__ __launch_bounds__(kTile) void k_hlt(DevItem it, InlineFields F) {
  run_tile<TF, TP, MATH, POST, InlineFields>(it, F, blockIdx.x, 0, it.nf);
}

}  // namespace adz

using namespace adz;

struct adz_plan {
  int fdtype, pdtype, math, post;
  adz_tune tune;
  PlanArgs args;
  int grid;
  void* dmem;
};

namespace {

bool fill_item(DevItem* d, const adz_ztab* geom, const void* xyz, int i0, int in, int j0, int jn, int k0, int nk,
               int dminx, int dmaxx, int dminy, int dmaxy, int ext, int nf) {
  if (!geom || !xyz || in < i0 || jn < j0 || nk < k0 || nf <= 0) return false;
  if (i0 < dminx || in > dmaxx || j0 < dminy || jn > dmaxy || k0 < 1) return false;  // box inside dst
  std::memset(d, 0, sizeof(*d));
  d->lv = geom->dev;
  if (!ext) {  // the source has the destination's (model) layout
    d->lv.ni = dmaxx - dminx + 1;
    d->lv.nj = dmaxy - dminy + 1;
    d->lv.nij = d->lv.ni * d->lv.nj;
    d->lv.ilo = dminx; d->lv.ihi = dmaxx; d->lv.jlo = dminy; d->lv.jhi = dmaxy;
  }
  d->xyz = xyz;
  d->i0 = i0; d->j0 = j0; d->k0 = k0;
  d->nx = in - i0 + 1; d->ny = jn - j0 + 1; d->nz = nk - k0 + 1;
  d->dni = dmaxx - dminx + 1; d->dnj = dmaxy - dminy + 1;
  d->nf = nf;
  d->npts = (long)d->nx * d->ny * d->nz;
  d->ptiles = (d->npts + kTile - 1) / kTile;
  return true;
}

// select the kernel instance for (fdtype, pdtype, math, post)
template <class TF, class TP, int M>
void* pick_plan(int post) {
  switch (post) {
    case 0: return (void*)&k_plan<TF, TP, M, 0>;
    case 1: return (void*)&k_plan<TF, TP, M, 1>;
    case 2: return (void*)&k_plan<TF, TP, M, 2>;
    default: return (void*)&k_plan<TF, TP, M, 3>;
  }
}
template <class TF, class TP, int M>
void* pick_hlt(int post) {
  switch (post) {
    case 0: return (void*)&k_hlt<TF, TP, M, 0>;
    case 1: return (void*)&k_hlt<TF, TP, M, 1>;
    case 2: return (void*)&k_hlt<TF, TP, M, 2>;
    default: return (void*)&k_hlt<TF, TP, M, 3>;
  }
}

template <bool PLAN>
void* pick(int fdt, int pdt, int math, int post) {
#define P_(TF, TP, M) return PLAN ? pick_plan<TF, TP, M>(post) : pick_hlt<TF, TP, M>(post)
  if (math == ADZ_MATH_REF) {
    if (fdt == ADZ_F64 && pdt == ADZ_F64) P_(double, double, ADZ_MATH_REF);
    if (fdt == ADZ_F64 && pdt == ADZ_F32) P_(double, float, ADZ_MATH_REF);
    if (fdt == ADZ_F32 && pdt == ADZ_F64) P_(float, double, ADZ_MATH_REF);
    if (fdt == ADZ_F32 && pdt == ADZ_F32) P_(float, float, ADZ_MATH_REF);
  } else if (math == ADZ_MATH_F32) {
    if (fdt == ADZ_F32 && pdt == ADZ_F64) P_(float, double, ADZ_MATH_F32);
    if (fdt == ADZ_F32 && pdt == ADZ_F32) P_(float, float, ADZ_MATH_F32);
    if (fdt == ADZ_BF16 && pdt == ADZ_F64) P_(bf16, double, ADZ_MATH_F32);
    if (fdt == ADZ_BF16 && pdt == ADZ_F32) P_(bf16, float, ADZ_MATH_F32);
  }
#undef P_
  return nullptr;
}

int device_cus() {
  static int cus = 0;
  if (!cus) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) !=
                                                 hipSuccess)
      cus = 256;
  }
  return cus;
}

}  // namespace

extern "C" {

int adz_tricub_hlt(const adz_field* fld, int nptr, const void* xyz, int pdtype, const adz_ztab* geom, int i0,
                   int in, int j0, int jn, int k0, int nk, int dminx, int dmaxx, int dminy, int dmaxy, int ext,
                   int post, int fdtype, int math, adz_stream stream) {
  if (!fld || nptr < 0) return fail(ADZ_EINVAL, "adz_tricub_hlt: bad field list");
  if (nptr == 0 || in < i0 || jn < j0) return ADZ_OK;  // nij <= 0: nothing to do (:61)
  if (post < 0 || post > 3) return fail(ADZ_EINVAL, "adz_tricub_hlt: post must be 0..3");
  void* kern = pick<false>(fdtype, pdtype, math, post);
  if (!kern) return fail(ADZ_EINVAL, "adz_tricub_hlt: unsupported (field dtype, position dtype, math)");
  for (int f = 0; f < nptr; f++) {
    if (!fld[f].src || !fld[f].dst) return fail(ADZ_EINVAL, "adz_tricub_hlt: null src/dst");
    if ((post & ADZ_POST_STORE) && (!fld[f].lin || !fld[f].mn || !fld[f].mx))
      return fail(ADZ_EINVAL, "adz_tricub_hlt: post arrays required");
  }
  for (int f0 = 0; f0 < nptr; f0 += kInlineFields) {
    int nf = std::min(kInlineFields, nptr - f0);
    DevItem it;
    if (!fill_item(&it, geom, xyz, i0, in, j0, jn, k0, nk, dminx, dmaxx, dminy, dmaxy, ext, nf))
      return fail(ADZ_EINVAL, "adz_tricub_hlt: bad box or descriptor");
    InlineFields F;
    std::memset(&F, 0, sizeof(F));
    for (int f = 0; f < nf; f++) {
      const adz_field& a = fld[f0 + f];
      F.src[f] = a.src; F.dst[f] = a.dst; F.lin[f] = a.lin; F.mn[f] = a.mn; F.mx[f] = a.mx;
    }
    void* args[] = {&it, &F};
    ADZ_HIP_CHECK(hipLaunchKernel(kern, dim3((unsigned)it.ptiles), dim3(kTile), args, 0, (hipStream_t)stream));
  }
  return ADZ_OK;
}

int adz_plan_create(const adz_item* items, int nitems, int fdtype, int math, int post, const adz_tune* tune,
                    adz_plan** out) {
  if (!items || nitems <= 0 || !out) return fail(ADZ_EINVAL, "adz_plan_create: no items");
  if (post < 0 || post > 3) return fail(ADZ_EINVAL, "adz_plan_create: post must be 0..3");
  int pdt = items[0].pdtype;
  for (int q = 0; q < nitems; q++)
    if (items[q].pdtype != pdt) return fail(ADZ_EINVAL, "adz_plan_create: items must share the position dtype");
  if (!pick<true>(fdtype, pdt, math, post)) return fail(ADZ_EINVAL, "adz_plan_create: unsupported dtype/math");
  adz_tune T{8, 0, 8, 1};
  if (tune) {
    if (tune->fields_per_pass > 0) T.fields_per_pass = std::min(64, tune->fields_per_pass);
    T.points_outer = tune->points_outer;
    if (tune->blocks_per_cu > 0) T.blocks_per_cu = tune->blocks_per_cu;
    T.xcd_chunks = tune->xcd_chunks;
  }
  std::vector<DevItem> dv(nitems);
  long nptr_total = 0;
  for (int q = 0; q < nitems; q++) nptr_total += items[q].nptr;
  std::vector<void*> ptrs(5 * nptr_total);
  long tiles = 0, pofs = 0;
  for (int q = 0; q < nitems; q++) {
    const adz_item& a = items[q];
    if (!a.fld) return fail(ADZ_EINVAL, "adz_plan_create: null field list");
    if (!fill_item(&dv[q], a.geom, a.xyz, a.i0, a.in, a.j0, a.jn, a.k0, a.nk, a.dminx, a.dmaxx, a.dminy, a.dmaxy,
                   a.ext, a.nptr))
      return fail(ADZ_EINVAL, "adz_plan_create: bad item");
    for (int f = 0; f < a.nptr; f++) {
      if (!a.fld[f].src || !a.fld[f].dst) return fail(ADZ_EINVAL, "adz_plan_create: null src/dst");
      if ((post & ADZ_POST_STORE) && (!a.fld[f].lin || !a.fld[f].mn || !a.fld[f].mx))
        return fail(ADZ_EINVAL, "adz_plan_create: post arrays required");
      ptrs[0 * nptr_total + pofs + f] = (void*)a.fld[f].src;
      ptrs[1 * nptr_total + pofs + f] = a.fld[f].dst;
      ptrs[2 * nptr_total + pofs + f] = a.fld[f].lin;
      ptrs[3 * nptr_total + pofs + f] = a.fld[f].mn;
      ptrs[4 * nptr_total + pofs + f] = a.fld[f].mx;
    }
    dv[q].ngroups = (a.nptr + T.fields_per_pass - 1) / T.fields_per_pass;
    dv[q].tile0 = tiles;
    tiles += dv[q].ptiles * dv[q].ngroups;
    dv[q].src = (const void* const*)(intptr_t)(0 * nptr_total + pofs);  // patched below
    dv[q].dst = (void* const*)(intptr_t)(1 * nptr_total + pofs);
    dv[q].lin = (void* const*)(intptr_t)(2 * nptr_total + pofs);
    dv[q].mn = (void* const*)(intptr_t)(3 * nptr_total + pofs);
    dv[q].mx = (void* const*)(intptr_t)(4 * nptr_total + pofs);
    pofs += a.nptr;
  }
  size_t ib = sizeof(DevItem) * nitems, pb = sizeof(void*) * ptrs.size();
  void* mem = nullptr;
  if (hipMalloc(&mem, ib + pb) != hipSuccess) return fail(ADZ_ENOMEM, "adz_plan_create: hipMalloc");
  void** dptr = (void**)((char*)mem + ib);
  for (auto& d : dv) {
    d.src = (const void* const*)(dptr + (intptr_t)d.src);
    d.dst = (void* const*)(dptr + (intptr_t)d.dst);
    d.lin = (void* const*)(dptr + (intptr_t)d.lin);
    d.mn = (void* const*)(dptr + (intptr_t)d.mn);
    d.mx = (void* const*)(dptr + (intptr_t)d.mx);
  }
  if (hipMemcpy(mem, dv.data(), ib, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(dptr, ptrs.data(), pb, hipMemcpyHostToDevice) != hipSuccess) {
    (void)hipFree(mem);
    return fail(ADZ_EHIP, "adz_plan_create: upload");
  }
  adz_plan* p = new adz_plan;
  p->fdtype = fdtype; p->pdtype = pdt; p->math = math; p->post = post; p->tune = T; p->dmem = mem;
  p->args = PlanArgs{(const DevItem*)mem, nitems, tiles, T.fields_per_pass, T.points_outer, T.xcd_chunks};
  long grid = (long)T.blocks_per_cu * device_cus();
  if (grid > tiles) grid = tiles;
  if (T.xcd_chunks) grid = std::max(8L, (grid / 8) * 8);
  p->grid = (int)grid;
  *out = p;
  return ADZ_OK;
}

int adz_plan_execute(const adz_plan* p, adz_stream stream) {
  if (!p) return fail(ADZ_EINVAL, "adz_plan_execute: null plan");
  void* kern = pick<true>(p->fdtype, p->pdtype, p->math, p->post);
  PlanArgs a = p->args;
  void* args[] = {&a};
  ADZ_HIP_CHECK(hipLaunchKernel(kern, dim3(p->grid), dim3(kTile), args, 0, (hipStream_t)stream));
  return ADZ_OK;
}

int adz_plan_tune(const adz_plan* p, adz_tune* out) {
  if (!p || !out) return fail(ADZ_EINVAL, "adz_plan_tune: null argument");
  *out = p->tune;
  return ADZ_OK;
}

int adz_plan_free(adz_plan* p) {
  if (!p) return ADZ_OK;
  hipError_t e = hipFree(p->dmem);
  delete p;
  if (e != hipSuccess) return fail(ADZ_EHIP, "adz_plan_free: hipFree");
  return ADZ_OK;
}

}  // extern "C"
