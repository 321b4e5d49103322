// Fused multi-field interpolation at departure points (adz_tricub_hlt,
// src/gemdyn/src/base/adz_interp_hlt.F90:28-187) and its batched, persistent form over
// several subdomains.
//
// Work is cut into tiles of 64 x 4 target points (i fastest: four rows of 64) times a group
// of `fields_per_pass` fields.  One block of 256 lanes runs one tile: each lane computes the
// index/weights of its target once and interpolates the group's fields, writing straight
// into the destination arrays (the reference interpolates into a work buffer and scatters
// it, :165-180).  A persistent grid walks the tile list; with `xcd_chunks` the tile list is
// cut into 8 contiguous ranges, one per XCD (blocks are dispatched round-robin over the 8
// XCDs), so each XCD's L2 sees spatially adjacent tiles.
#include <algorithm>
#include <cstring>
#include <type_traits>
#include <vector>

#include "adz_common.h"
#include "adz_eval.h"

namespace adz {

constexpr int kTI = 64;       // targets per tile along i (one wavefront)
constexpr int kTJ = 4;        // rows of targets per tile (one wavefront each)
constexpr int kTile = kTI * kTJ;
constexpr int kPitch = 128;       // LDS row pitch (elements): a box row spans at most 128 columns
constexpr int kMaxRows = 39;      // box rows (y x planes) per LDS buffer; two buffers per block
                                  // (2 x 39 x 128 x 4 B + tables: four FP32 blocks fit a CU's 160 KB)
constexpr int kInlineFields = 16;

struct DevItem {
  ZtabDev lv;
  const void* xyz;
  int i0, j0, k0, nx, ny, nz;
  int dni, dnj;
  int nf;
  long minoff;      // lowest source element a stencil may read, from element (1,1,1)
  unsigned span;    // bytes of FP32 source from minoff that stencils may read (0: too large
                    // for 32-bit buffer offsets; global loads are used)
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
  int stage;
  int level_inner;
  int spin_limit;   // staging mode 3: polls of one bounded wait before it gives up
  int* status;      // device word: number of staging waits that gave up (the plan's)
};

struct InlineFields {  // kernel-argument table (adz_tricub_hlt)
  const void* src[kInlineFields];
  void* dst[kInlineFields];
  void* lin[kInlineFields];
  void* mn[kInlineFields];
  void* mx[kInlineFields];
  __device__ const void* s(int f) const { return src[f]; }
  __device__ void* d(int f) const { return dst[f]; }
  __device__ void* l(int f) const { return lin[f]; }
  __device__ void* n(int f) const { return mn[f]; }
  __device__ void* x(int f) const { return mx[f]; }
};

struct ItemFields {  // field pointer tables of a plan item (device arrays in global memory)
  const void* const* src;
  void* const* dst;
  void* const* lin;
  void* const* mn;
  void* const* mx;
  typedef const void* cvp;
  typedef void* vp;
  // The tables do not change during a launch: reading them through the constant address
  // space lets the compiler use scalar loads for the (wave-uniform) field index, which
  // keeps them off the vector-memory counter.
#define ADZ_CONST __attribute__((address_space(4)))
  __device__ const void* s(int f) const { return ((const cvp ADZ_CONST*)src)[f]; }
  __device__ void* d(int f) const { return ((const vp ADZ_CONST*)dst)[f]; }
  __device__ void* l(int f) const { return ((const vp ADZ_CONST*)lin)[f]; }
  __device__ void* n(int f) const { return ((const vp ADZ_CONST*)mn)[f]; }
  __device__ void* x(int f) const { return ((const vp ADZ_CONST*)mx)[f]; }
  __device__ const void* sc(int f) const { return s(f); }
#undef ADZ_CONST
};

// Narrow and store one target's outputs at dst offset o: the value (clipped to [min, max]
// with ADZ_POST_CLIP) and, with ADZ_POST_STORE, lin clipped to [min, max], min and max.
// TC is the arithmetic type (FP64 for the reference's arithmetic, FP32 for the mixed form).
template <class TF, int MATH, int POST, class Fields, class TC>
__device__ __forceinline__ void store_out(const Fields& F, int f, long o, TC v, TC l, TC mi, TC ma) {
  typename GPtr<TF>::type d = (typename GPtr<TF>::type)F.d(f);
  if (MATH == ADZ_MATH_REF) {
    if (POST) {
      // adz_interp_hlt.F90:158-160 narrows each to real before clipping lin
      TF ls = from_f64<TF>(l), mis = from_f64<TF>(mi), mas = from_f64<TF>(ma);
      TF vs = from_f64<TF>(v);
      if (POST & ADZ_POST_CLIP) vs = from_f64<TF>(fmax(to_f64(mis), fmin(to_f64(mas), to_f64(vs))));
      d[o] = vs;
      if (POST & ADZ_POST_STORE) {
        ((typename GPtr<TF>::type)F.l(f))[o] = from_f64<TF>(fmax(to_f64(mis), fmin(to_f64(mas), to_f64(ls))));
        ((typename GPtr<TF>::type)F.n(f))[o] = mis;
        ((typename GPtr<TF>::type)F.x(f))[o] = mas;
      }
    } else {
      d[o] = from_f64<TF>(v);
    }
  } else {
    if (POST) {
      TF ls = from_f32<TF>(l), mis = from_f32<TF>(mi), mas = from_f32<TF>(ma);
      TF vs = from_f32<TF>(snap_to_range(v, mi, ma));
      if (POST & ADZ_POST_CLIP) vs = from_f32<TF>(fmaxf(to_f32(mis), fminf(to_f32(mas), to_f32(vs))));
      d[o] = vs;
      if (POST & ADZ_POST_STORE) {
        ((typename GPtr<TF>::type)F.l(f))[o] = from_f32<TF>(fmaxf(to_f32(mis), fminf(to_f32(mas), to_f32(ls))));
        ((typename GPtr<TF>::type)F.n(f))[o] = mis;
        ((typename GPtr<TF>::type)F.x(f))[o] = mas;
      }
    } else {
      d[o] = from_f32<TF>(v);
    }
  }
}

// Evaluate one target from its stencil accessor and store the outputs at dst offset o.
template <class TF, int MATH, int POST, class Fields, class Acc>
__device__ __forceinline__ void eval_store_a(const Fields& F, int f, Acc at, const void* cw, long o) {
  if (MATH == ADZ_MATH_REF) {
    const double* c = (const double*)cw;
    double v = tricub_ref_a<TF>(at, c), l = 0, mi = 0, ma = 0;
    if (POST) mono_extra_a<TF, double>(at, c, &l, &mi, &ma);
    store_out<TF, MATH, POST>(F, f, o, v, l, mi, ma);
  } else {
    const float* c = (const float*)cw;
    float v = tricub_f32_a<TF>(at, c), l = 0, mi = 0, ma = 0;
    if (POST) mono_extra_a<TF, float>(at, c, &l, &mi, &ma);
    store_out<TF, MATH, POST>(F, f, o, v, l, mi, ma);
  }
}

// The same from a source addressed with row stride `ni` and plane stride `nij` (global
// memory or the LDS box).
template <class TF, int MATH, int POST, class Fields, class P>
__device__ __forceinline__ void eval_store(const Fields& F, int f, P f1p, long ni, long nij, int zlinear,
                                           const void* cw, long o) {
  const long o1 = zlinear ? 0 : nij;
  const long o2 = o1 + nij, o3 = o2 + o1;
  eval_store_a<TF, MATH, POST>(F, f, PtrStencil<P>{f1p, ni, o1, o2, o3}, cw, o);
}

// FP32 sources through buffer loads: the 16 rows of a lane's stencil are read with
// buffer_load_dwordx4 at 32-bit byte offsets from the lowest address a stencil may touch.
// The offsets are computed once per target and reused for every field; per field only the
// (scalar) buffer descriptor changes.
typedef float f32x4 __attribute__((ext_vector_type(4)));
struct RegStencil {
  const f32x4 (*v)[4];
  __device__ __forceinline__ float operator()(int r, int z, int i) const { return v[r][z][i]; }
};

__device__ __forceinline__ __amdgpu_buffer_rsrc_t field_rsrc(const void* base, unsigned bytes) {
  // the field pointer is uniform across the wave: keep the descriptor in SGPRs
  unsigned long long a = (unsigned long long)base;
  unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)a), hi = __builtin_amdgcn_readfirstlane((unsigned)(a >> 32));
  void* p = (void*)(((unsigned long long)hi << 32) | lo);
  return __builtin_amdgcn_make_buffer_rsrc(p, (short)0, (int)__builtin_amdgcn_readfirstlane(bytes), 0x00020000);
}

template <class TF, int MATH, int POST, class Fields>
__device__ __forceinline__ void store_nan(const Fields& F, int f, long o) {
  TF nan = (MATH == ADZ_MATH_REF) ? from_f64<TF>(__builtin_nan("")) : from_f32<TF>(__builtin_nanf(""));
  ((typename GPtr<TF>::type)F.d(f))[o] = nan;
  if (POST & ADZ_POST_STORE) { ((typename GPtr<TF>::type)F.l(f))[o] = nan; ((typename GPtr<TF>::type)F.n(f))[o] = nan; ((typename GPtr<TF>::type)F.x(f))[o] = nan; }
}

// Minimum over the 64 lanes of a wave (every lane active), returned wave-uniform: lane
// pairs, quads, half-rows and rows of 16 by DPP moves within each row (xor 1, xor 2, the
// half-row and row mirrors), then the four row minima read as scalars.  No LDS traffic.
__device__ __forceinline__ int wave_min(int v) {
  v = min(v, __builtin_amdgcn_update_dpp(v, v, 0xB1, 0xF, 0xF, false));   // quad_perm [1,0,3,2]
  v = min(v, __builtin_amdgcn_update_dpp(v, v, 0x4E, 0xF, 0xF, false));   // quad_perm [2,3,0,1]
  v = min(v, __builtin_amdgcn_update_dpp(v, v, 0x141, 0xF, 0xF, false));  // row_half_mirror
  v = min(v, __builtin_amdgcn_update_dpp(v, v, 0x140, 0xF, 0xF, false));  // row_mirror
  return min(min(__builtin_amdgcn_readlane(v, 0), __builtin_amdgcn_readlane(v, 16)),
             min(__builtin_amdgcn_readlane(v, 32), __builtin_amdgcn_readlane(v, 48)));
}

// block-wide minimum of 6 values over the tile's TH waves of 64; `red` = 6 * TH ints of LDS
template <int TH = kTJ>
__device__ __forceinline__ void block_min6(int* v, int* red) {
#pragma unroll
  for (int q = 0; q < 6; q++) v[q] = wave_min(v[q]);
  if ((threadIdx.x & 63) == 0)
#pragma unroll
    for (int q = 0; q < 6; q++) red[q * TH + (threadIdx.x >> 6)] = v[q];
  __syncthreads();
#pragma unroll
  for (int q = 0; q < 6; q++) {
    int m = red[q * TH];
#pragma unroll
    for (int w = 1; w < TH; w++) m = min(m, red[q * TH + w]);
    v[q] = m;
  }
  __syncthreads();
}

// The outputs of one target (value, and with POST lin / min / max) from a stencil accessor.
template <class TF, int MATH, int POST, class Acc, class TC>
__device__ __forceinline__ void eval_vals(Acc at, const void* cw, TC* v, TC* l, TC* mi, TC* ma) {
  if constexpr (MATH == ADZ_MATH_REF) {
    *v = tricub_ref_a<TF>(at, (const double*)cw);
    if (POST) mono_extra_a<TF, double>(at, (const double*)cw, l, mi, ma);
  } else {
    *v = tricub_f32_a<TF>(at, (const float*)cw);
    if (POST) mono_extra_a<TF, float>(at, (const float*)cw, l, mi, ma);
  }
}

// Per-lane evaluation of fields [f0, f1) for one target: every lane reads its own 16
// stencil rows (buffer loads for FP32 sources, global loads otherwise).
template <class TF, int MATH, int POST, class Fields>
__device__ __forceinline__ void per_lane(const DevItem& it, const Fields& F, int f0, int f1, const Index& x, bool ok,
                                         bool valid, const double* cw, long o) {
  const long gni = it.lv.ni, gnij = it.lv.nij;
  if constexpr (std::is_same<TF, float>::value) {
    if (it.span) {
      // per-lane byte offsets of the 4 planes of row 0; rows r add r*ni, a uniform offset
      // that goes in the instruction's scalar offset
      int bo[4];
      if (ok) {
        const long o1 = x.zlinear ? 0 : gnij, o2 = o1 + gnij, o3 = o2 + o1;
        const long e = x.off - it.minoff;
#pragma unroll
        for (int z = 0; z < 4; z++) bo[z] = (int)((e + (z == 0 ? 0 : (z == 1 ? o1 : (z == 2 ? o2 : o3)))) * 4);
      }
      const int rstride = (int)(gni * 4);
      for (int f = f0; f < f1; f++) {
        if (ok) {
          const __amdgpu_buffer_rsrc_t rs = field_rsrc((const float*)F.s(f) + it.minoff, it.span);
          f32x4 v[4][4];
#pragma unroll
          for (int r = 0; r < 4; r++)
#pragma unroll
            for (int z = 0; z < 4; z++)
              v[r][z] = __builtin_bit_cast(f32x4, __builtin_amdgcn_raw_buffer_load_b128(rs, bo[z], r * rstride, 0));
          eval_store_a<TF, MATH, POST>(F, f, RegStencil{v}, cw, o);
        } else if (valid) {
          store_nan<TF, MATH, POST>(F, f, o);
        }
      }
      return;
    }
  }
  for (int f = f0; f < f1; f++) {
    if (ok)
      eval_store<TF, MATH, POST, Fields>(F, f, (typename GPtr<TF>::ctype)F.s(f) + x.off, gni, gnij, x.zlinear, cw, o);
    else if (valid) store_nan<TF, MATH, POST>(F, f, o);
  }
}

// Tile order within an item (ptile -> tile column ib, tile row jb, level kk):
//   level_inner 0: i fastest, then j, then k;
//   level_inner 1: the levels of one tile column first, then i, then j;
//   level_inner S >= 2: super-columns of S tile rows (4*S target rows) x every level, each
//     walked level by level (the S tile rows of a level together), super-columns along i,
//     then j.  Consecutive tiles then share source planes (level to level) and source rows
//     (within the super-column), which keeps the blocks resident on one XCD inside a
//     working set its L2 can hold.  Tiles past the last row (njb not a multiple of S) hold
//     no valid target.
struct TileXY {
  int ib, jb, kk;
};
// (32-bit arithmetic: the host refuses items of 2^31 tiles or more)
// Tiles of tl > 1 levels: kk is the tile's level block (levels kk * tl .. kk * tl + tl - 1).
__device__ __forceinline__ TileXY tile_at(const DevItem& it, long ptile_l, int level_inner, int tw = kTI,
                                         int th = kTJ, int tl = 1) {
  const unsigned ptile = (unsigned)ptile_l;
  const unsigned nib = (unsigned)(it.nx + tw - 1) / (unsigned)tw, njb = (unsigned)(it.ny + th - 1) / (unsigned)th;
  const unsigned nz = (unsigned)(it.nz + tl - 1) / (unsigned)tl;
  TileXY c;
  if (level_inner >= 2) {
    const unsigned S = (unsigned)level_inner, per = nz * S;
    const unsigned sc = ptile / per, r = ptile - sc * per;
    const unsigned kq = r / S, scj = sc / nib;
    c.kk = (int)kq;
    c.ib = (int)(sc - scj * nib);
    c.jb = (int)(scj * S + (r - kq * S));
  } else if (level_inner) {
    const unsigned q = ptile / nz, qj = q / nib;
    c.kk = (int)(ptile - q * nz);
    c.ib = (int)(q - qj * nib);
    c.jb = (int)qj;
  } else {
    const unsigned q = ptile / nib, qk = q / njb;
    c.ib = (int)(ptile - q * nib);
    c.jb = (int)(q - qk * njb);
    c.kk = (int)qk;
  }
  return c;
}
__device__ __forceinline__ void tile_coords(const DevItem& it, long ptile, int level_inner, int* ib, int* jb,
                                            int* kk, int tw = kTI) {
  const TileXY c = tile_at(it, ptile, level_inner, tw);
  *ib = c.ib; *jb = c.jb; *kk = c.kk;
}

// tiles of an item in the order above
inline long item_ptiles(int nx, int ny, int nz_levels, int level_inner, int tw = kTI, int th = kTJ, int tl = 1) {
  const long nib = (nx + tw - 1) / tw, njb = (ny + th - 1) / th, nz = (nz_levels + tl - 1) / tl;
  if (level_inner >= 2) return nib * ((njb + level_inner - 1) / level_inner) * level_inner * nz;
  return nib * njb * nz;
}

// One tile: TI x TJ targets (i fastest) of one level of item `it`, fields [f0, f1).
// The union of the tile's stencils (its "box") is staged in LDS once per field with
// coalesced loads, then every lane evaluates its stencil from LDS.  A box larger than
// the LDS buffer (a tile straddling very different departure points) falls back to
// per-lane global loads.  Both paths run the same arithmetic in the same order.
template <class TF, class TP, int MATH, int POST, class Fields>
__device__ __forceinline__ void run_tile(const DevItem& it, const Fields& F, long ptile, int f0, int f1, TF* boxA,
                                         TF* boxB, int* red, int* rowoff, int stage, int level_inner) {
  const int lane = threadIdx.x & (kTI - 1), wv = threadIdx.x / kTI;
  int ib, jb, kk;
  tile_coords(it, ptile, level_inner, &ib, &jb, &kk);
  const int ii = ib * kTI + lane, jj = jb * kTJ + wv;
  const bool valid = ii < it.nx && jj < it.ny;
  const int i = it.i0 + ii, j = it.j0 + jj, k = it.k0 + kk;
  const long p = (long)ii + (long)it.nx * (jj + (long)it.ny * kk);
  const long o = (long)(i - 1) + (long)it.dni * ((j - 1) + (long)it.dnj * (k - 1));  // dst(i,j,k) from dst(1,1,1)
  Index x;
  bool ok = false;
  if (valid) {
    double px, py, pz;
    load_pos<TP>(it.xyz, p, &px, &py, &pz);
    ok = index_step(px, py, pz, it.lv, &x);
  }
  double cw[20];  // weights (FP64, or FP32 in the first 10 slots' storage)
  if (ok) {
    if (MATH == ADZ_MATH_REF) weights_ref(x, it.lv, cw);
    else weights_f32(x, it.lv, (float*)cw);
  }
  if (stage == 2) {  // per-lane loads: no box needed
    per_lane<TF, MATH, POST>(it, F, f0, f1, x, ok, valid, cw, o);
    return;
  }
  // box = union of the tile's stencils (columns ix-1..ix+2, rows iy-1..iy+2, planes pz..)
  const int big = 1 << 30;
  int ext[6] = {ok ? x.ix - 1 : big, ok ? -(x.ix + 2) : big, ok ? x.iy - 1 : big, ok ? -(x.iy + 2) : big,
                ok ? x.pz : big, ok ? -(x.pz + (x.zlinear ? 1 : 3)) : big};
  block_min6(ext, red);
  const int x0 = ext[0], x1 = -ext[1], y0 = ext[2], y1 = -ext[3], z0 = ext[4], z1 = -ext[5];
  if (x0 == big) {  // no target of the tile has an in-range stencil
    if (valid)
      for (int f = f0; f < f1; f++) store_nan<TF, MATH, POST>(F, f, o);
    return;
  }
  const int nxb = x1 - x0 + 1, nyb = y1 - y0 + 1, nzb = z1 - z0 + 1;
  const long vol = (long)nxb * nyb * nzb;
  const long gni = it.lv.ni, gnij = it.lv.nij;
  const int nrows = nyb * nzb;
  if (nrows <= kMaxRows && nxb <= kPitch) {
    const long nijb = (long)kPitch * nyb;
    const long boff = ok ? (long)(x.ix - 1 - x0) + (long)(x.iy - 1 - y0) * kPitch + (long)(x.pz - z0) * nijb : 0;
    const long gbase = (long)(x0 - 1) + (long)(y0 - 1) * gni + (long)z0 * gnij;  // box corner from f(1,1,1)
    if (threadIdx.x < nrows) rowoff[threadIdx.x] = (int)((threadIdx.x % nyb) * gni + (threadIdx.x / nyb) * gnij);
    __syncthreads();
    // Asynchronous staging: each wavefront copies whole box rows straight from global
    // memory into LDS (global_load_lds, no registers), lanes along i.  Field f+1 is in
    // flight while field f is evaluated from the other buffer.
    constexpr int U = sizeof(TF) == 2 ? 2 : 4;          // bytes per lane per load
    const int units = nxb * (int)sizeof(TF) / U;         // loads per row
    auto issue = [&](int f, TF* buf) {
      const char* src = (const char*)((const TF*)F.s(f) + gbase);
      char* dst = (char*)buf;
      for (int r = wv; r < nrows; r += kTJ) {
        const char* grow = src + (long)rowoff[r] * sizeof(TF);
        char* lrow = dst + (long)r * kPitch * sizeof(TF);
        for (int c = 0; c < units; c += kTI) {
          if (c + lane < units) {
            const void* ga = grow + (long)(c + lane) * U;
            void* la = lrow + (long)c * U;
            if (U == 4)
              __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void*)ga,
                                               (__attribute__((address_space(3))) void*)la, 4, 0, 0);
            else
              __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void*)ga,
                                               (__attribute__((address_space(3))) void*)la, 2, 0, 0);
          }
        }
      }
    };
    // Two distinct LDS buffers with static roles (the loop is unrolled by two) let the
    // compiler prove the DMA target and the buffer being read are disjoint, so the reads
    // of field f do not wait for the loads of field f+1 still in flight.
    issue(f0, boxA);
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();
    for (int f = f0; f < f1; f += 2) {
      if (f + 1 < f1) issue(f + 1, boxB);
      if (ok) eval_store<TF, MATH, POST, Fields>(F, f, boxA + boff, kPitch, nijb, x.zlinear, cw, o);
      else if (valid) store_nan<TF, MATH, POST>(F, f, o);
      asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
      __syncthreads();
      if (f + 1 >= f1) break;
      if (f + 2 < f1) issue(f + 2, boxA);
      if (ok) eval_store<TF, MATH, POST, Fields>(F, f + 1, boxB + boff, kPitch, nijb, x.zlinear, cw, o);
      else if (valid) store_nan<TF, MATH, POST>(F, f + 1, o);
      asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
      __syncthreads();
    }
  } else {
    per_lane<TF, MATH, POST>(it, F, f0, f1, x, ok, valid, cw, o);
  }
}



// ---------------------------------------------------------------------------------
// Staging mode 3 (FP32 sources): a producer / consumer block of 5 wavefronts.
//
// Wave 0 only copies each field's box (the union of the tile's stencils) into a ring of
// LDS slots with 16-byte buffer-to-LDS DMA, several fields ahead.  It issues no stores,
// so its vector-memory counter counts its own DMA in order, and s_waitcnt vmcnt(k*R)
// tells it when an older field has landed; it then publishes that field in an LDS flag.
// Waves 1..kTJ (one target row each) wait for the flag, read their stencils from LDS,
// evaluate, store, and return the slot with an LDS atomic.  Every wait is a bounded spin
// (PlanArgs::spin_limit polls, 2^20 by default) instead of an unbounded one.  A wait that
// gives up raises the block's abort flag and counts itself in the plan's status word; the
// producer then stops filling slots, and every field whose slot was not published is
// evaluated from global memory by per-lane loads, so the result stays exact (only slower).
// One block barrier per tile separates tiles; an aborted block stages no later tile.
constexpr int kPCWaves = kTJ + 1;
constexpr int kPCThreads = kPCWaves * 64;
constexpr int kRing = 12288;   // floats of slot ring per block (48 KB)
constexpr int kMaxSlots = 8;
constexpr int kPCDepth = 3;    // fields in flight per producer

__device__ __forceinline__ void wait_vmcnt(int n) {
  switch (n) {
#define ADZ_VMW(k) case k: asm volatile("s_waitcnt vmcnt(" #k ")" ::: "memory"); break;
    ADZ_VMW(0) ADZ_VMW(1) ADZ_VMW(2) ADZ_VMW(3) ADZ_VMW(4) ADZ_VMW(5) ADZ_VMW(6) ADZ_VMW(7)
    ADZ_VMW(8) ADZ_VMW(9) ADZ_VMW(10) ADZ_VMW(11) ADZ_VMW(12) ADZ_VMW(13) ADZ_VMW(14) ADZ_VMW(15)
    ADZ_VMW(16) ADZ_VMW(17) ADZ_VMW(18) ADZ_VMW(19) ADZ_VMW(20) ADZ_VMW(21) ADZ_VMW(22) ADZ_VMW(23)
    ADZ_VMW(24) ADZ_VMW(25) ADZ_VMW(26) ADZ_VMW(27) ADZ_VMW(28) ADZ_VMW(29) ADZ_VMW(30) ADZ_VMW(31)
    ADZ_VMW(32) ADZ_VMW(33) ADZ_VMW(34) ADZ_VMW(35) ADZ_VMW(36) ADZ_VMW(37) ADZ_VMW(38) ADZ_VMW(39)
    ADZ_VMW(40) ADZ_VMW(41) ADZ_VMW(42) ADZ_VMW(43) ADZ_VMW(44) ADZ_VMW(45) ADZ_VMW(46) ADZ_VMW(47)
    ADZ_VMW(48) ADZ_VMW(49) ADZ_VMW(50) ADZ_VMW(51) ADZ_VMW(52) ADZ_VMW(53) ADZ_VMW(54) ADZ_VMW(55)
    ADZ_VMW(56) ADZ_VMW(57) ADZ_VMW(58) ADZ_VMW(59) ADZ_VMW(60) ADZ_VMW(61) ADZ_VMW(62)
#undef ADZ_VMW
    default: asm volatile("s_waitcnt vmcnt(63)" ::: "memory"); break;
  }
}

// Bounded spin on an LDS word until *p >= v or the block's abort flag is up.  A wait that
// runs out (`limit` polls) raises the abort flag and adds one to the plan's status word, and
// every later wait of the block returns at once, so a broken hand-off costs one timeout per
// block instead of one per wait.  Returns true when *p >= v.
typedef volatile __attribute__((address_space(3))) int lds_flag;  // LDS words read with ds_read
typedef __attribute__((address_space(3))) float lds_float;

__device__ __forceinline__ bool spin_ge(lds_flag* p, int v, lds_flag* abort_flag, int limit, int* status) {
  for (int it = 0;; it++) {
    if (*p >= v) return true;
    if (*abort_flag) return false;
    if (it >= limit) break;
    __builtin_amdgcn_s_sleep(1);
  }
  *abort_flag = 1;
  if ((threadIdx.x & 63) == 0) __hip_atomic_fetch_add(status, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return false;
}

// Publish a flag word.  Written in assembly: the compiler would otherwise wait for every
// outstanding LDS DMA before an LDS store it cannot prove disjoint from the DMA targets,
// which would serialise the producer's pipeline.  The caller has already waited for the
// DMA this flag announces.
__device__ __forceinline__ void lds_publish(lds_flag* p, int v) {
  const unsigned a = (unsigned)(unsigned long long)p;
  asm volatile("ds_write_b32 %0, %1" ::"v"(a), "v"(v) : "memory");
}

struct PCBox {  // the tile's staged box and ring geometry (uniform)
  int x0, y0, z0, nyb, cpr, nch, r, slot, ns;
};

template <int MATH, int POST, class TP>
__device__ __forceinline__ void pc_tile(const DevItem& it, const ItemFields& F, long ptile, int f0, int f1,
                                        lds_float* ring, lds_flag* ready, lds_flag* freed, int* red,
                                        lds_flag* abort_flag, int level_inner, int spin_limit, int* status) {
  const int wave = threadIdx.x / 64, lane = threadIdx.x & 63;
  const bool consumer = wave > 0;
  int ib, jb, kk;
  tile_coords(it, ptile, level_inner, &ib, &jb, &kk);
  const int ii = ib * kTI + lane, jj = jb * kTJ + (wave - 1);
  const bool valid = consumer && ii < it.nx && jj < it.ny;
  const int i = it.i0 + ii, j = it.j0 + jj, k = it.k0 + kk;
  const long p = (long)ii + (long)it.nx * (jj + (long)it.ny * kk);
  const long o = (long)(i - 1) + (long)it.dni * ((j - 1) + (long)it.dnj * (k - 1));
  Index x;
  bool ok = false;
  if (valid) {
    double px, py, pz;
    load_pos<TP>(it.xyz, p, &px, &py, &pz);
    ok = index_step(px, py, pz, it.lv, &x);
  }
  double cw[20];
  if (ok) {
    if (MATH == ADZ_MATH_REF) weights_ref(x, it.lv, cw);
    else weights_f32(x, it.lv, (float*)cw);
  }
  // box extents over the block (the producer contributes nothing)
  const int big = 1 << 30;
  int e[6] = {ok ? x.ix - 1 : big, ok ? -(x.ix + 2) : big, ok ? x.iy - 1 : big, ok ? -(x.iy + 2) : big,
              ok ? x.pz : big, ok ? -(x.pz + (x.zlinear ? 1 : 3)) : big};
#pragma unroll
  for (int q = 0; q < 6; q++)
    for (int sft = 32; sft > 0; sft >>= 1) e[q] = min(e[q], __shfl_xor(e[q], sft, 64));
  if (lane == 0)
#pragma unroll
    for (int q = 0; q < 6; q++) red[q * kPCWaves + wave] = e[q];
  __syncthreads();
#pragma unroll
  for (int q = 0; q < 6; q++) {
    int v = red[q * kPCWaves];
    for (int w = 1; w < kPCWaves; w++) v = min(v, red[q * kPCWaves + w]);
    e[q] = __builtin_amdgcn_readfirstlane(v);  // block-uniform (and known so to the compiler)
  }
  if (threadIdx.x < kMaxSlots) { ready[threadIdx.x] = 0; freed[threadIdx.x] = 0; }
  __syncthreads();  // everyone has read `red`; the flags are reset
  const long gni = it.lv.ni, gnij = it.lv.nij;
  // item values the DMA needs, read once and held in scalar registers
  const long minoff = ((long)__builtin_amdgcn_readfirstlane((int)(it.minoff >> 32)) << 32) |
                      (unsigned)__builtin_amdgcn_readfirstlane((int)it.minoff);
  const unsigned span = __builtin_amdgcn_readfirstlane(it.span);
  PCBox b;
  b.x0 = e[0]; b.y0 = e[2]; b.z0 = e[4];
  const int nxb = -e[1] - e[0] + 1, nyb = -e[3] - e[2] + 1, nzb = -e[5] - e[4] + 1;
  b.nyb = nyb;
  b.cpr = (nxb + 3 + 3) / 4;              // 16-byte chunks per row (room for the alignment shift)
  b.nch = nyb * nzb * b.cpr;
  b.r = (b.nch + 63) / 64;                // DMA instructions per field
  b.slot = b.r * 256;                     // floats per slot (each instruction fills 1 KB)
  b.ns = e[0] == big ? 0 : kRing / b.slot;
  if (b.ns > kMaxSlots) b.ns = kMaxSlots;
  const bool staged = e[0] != big && b.ns >= 2 && span && (gni & 3) == 0 && (gnij & 3) == 0 && b.r <= 21 &&
                      !*abort_flag;
  if (!staged) {  // per-lane loads for this tile; the producer has nothing to do
    if (consumer) per_lane<float, MATH, POST>(it, F, f0, f1, x, ok, valid, cw, o);
    __syncthreads();
    return;
  }
  const int nf = f1 - f0;
  const int depth = min(kPCDepth, b.ns);
  if (!consumer) {
    // ------------------------------------------------------------------ producer
    const float rcpr = 1.0f / (float)b.cpr, rnyb = 1.0f / (float)nyb;
    bool stopped = false;
    for (int q = 0; q < nf; q++) {
      const int sl = q % b.ns;
      // a slot is refilled only once its four consumers have returned it
      if (q >= b.ns && !spin_ge(&freed[sl], kTJ * (q / b.ns), abort_flag, spin_limit, status)) {
        stopped = true;
        break;
      }
      const float* base = (const float*)F.sc(f0 + q);
      const int sh = (int)((((unsigned long long)base >> 2) + (unsigned)(b.x0 - 1)) & 3);
      const __amdgpu_buffer_rsrc_t rs = field_rsrc(base + minoff, span);
      lds_float* dstslot = ring + sl * b.slot;
      for (int kq = 0; kq < b.r; kq++) {
        const int c = kq * 64 + lane;
        if (c < b.nch) {
          const int rr = (int)(((float)c + 0.5f) * rcpr);
          const int cc = c - rr * b.cpr;
          const int zz = (int)(((float)rr + 0.5f) * rnyb);
          const int yy = rr - zz * nyb;
          const long el = (long)(b.x0 - 1 - sh) + (long)(b.y0 - 1 + yy) * gni + (long)(b.z0 + zz) * gnij + 4 * cc -
                          minoff;
          __builtin_amdgcn_raw_ptr_buffer_load_lds(rs, (__attribute__((address_space(3))) void*)(dstslot + kq * 256),
                                                   16, (int)(el * 4), 0, 0, 0);
        }
      }
      if (q >= depth - 1) {  // the field issued depth-1 fields ago has landed
        wait_vmcnt((depth - 1) * b.r);
        if (lane == 0) lds_publish(&ready[(q - depth + 1) % b.ns], q - depth + 2);
      }
    }
    wait_vmcnt(0);
    if (lane == 0 && !stopped)
      for (int q = max(0, nf - depth + 1); q < nf; q++) lds_publish(&ready[q % b.ns], q + 1);
  } else {
    // ------------------------------------------------------------------ consumers
    // stencil row (r, plane z) starts at LDS float ((iy-1-y0+r) + nyb*(pz-z0+zp)) * W +
    // (ix-1-x0) + shift, W = 4*cpr
    const int W = 4 * b.cpr;
    int lrow[4] = {0, 0, 0, 0};
    if (ok) {
      const int zp1 = x.zlinear ? 0 : 1, zp2 = zp1 + 1, zp3 = zp2 + zp1;
      const int l0 = ((x.iy - 1 - b.y0) + nyb * (x.pz - b.z0)) * W + (x.ix - 1 - b.x0);
      lrow[0] = l0; lrow[1] = l0 + zp1 * nyb * W; lrow[2] = l0 + zp2 * nyb * W; lrow[3] = l0 + zp3 * nyb * W;
    }
    for (int q = 0; q < nf; q++) {
      const int sl = q % b.ns;
      // the slot holds field q once ready[sl] >= q + 1: the producer writes a slot only
      // after its consumers returned it, so it cannot hold a later field yet
      const bool landed = spin_ge(&ready[sl], q + 1, abort_flag, spin_limit, status);
      if (!landed) {  // the hand-off gave up: this field from global memory, per lane
        per_lane<float, MATH, POST>(it, F, f0 + q, f0 + q + 1, x, ok, valid, cw, o);
        continue;
      }
      const float* base = (const float*)F.sc(f0 + q);
      const int sh = (int)((((unsigned long long)base >> 2) + (unsigned)(b.x0 - 1)) & 3);
      if (ok) {
        const lds_float* slot = ring + sl * b.slot + sh;
        auto at = [&](int r, int z, int c) -> float { return slot[lrow[z] + r * W + c]; };
        eval_store_a<float, MATH, POST>(F, f0 + q, at, cw, o);
      } else if (valid) {
        store_nan<float, MATH, POST>(F, f0 + q, o);
      }
      asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");  // this wave's LDS reads are done
      if (lane == 0) __hip_atomic_fetch_add((__attribute__((address_space(3))) int*)&freed[sl], 1,
                                            __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
  }
  __syncthreads();
}

template <class TP, int MATH, int POST>
__global__ __launch_bounds__(kPCThreads) void k_plan_pc(PlanArgs A) {
  __shared__ float ring[kRing];
  __shared__ int ready[kMaxSlots];
  __shared__ int freed[kMaxSlots];
  __shared__ int red[6 * kPCWaves];
  __shared__ int abort_flag;
  if (threadIdx.x == 0) abort_flag = 0;
  __syncthreads();
  long lo = 0, hi = A.ntiles, step = gridDim.x, t0 = blockIdx.x;
  if (A.xcd_chunks) {
    int nb = gridDim.x / 8;
    int xc = blockIdx.x % 8, l = blockIdx.x / 8;
    lo = A.ntiles * xc / 8;
    hi = A.ntiles * (xc + 1) / 8;
    step = nb;
    t0 = lo + l;
  }
  for (long t = t0; t < hi; t += step) {
    int q = 0;
    const DevItem* items = A.items;
    while (q + 1 < A.nitems && items[q + 1].tile0 <= t) q++;
    const DevItem& it = items[q];
    long local = t - it.tile0;
    long ptile, grp;
    if (A.points_outer) { ptile = local / it.ngroups; grp = local % it.ngroups; }
    else { grp = local / it.ptiles; ptile = local % it.ptiles; }
    int f0 = (int)grp * A.g, f1 = min(it.nf, f0 + A.g);
    ItemFields F{it.src, it.dst, it.lin, it.mn, it.mx};
    pc_tile<MATH, POST, TP>(it, F, ptile, f0, f1, (lds_float*)ring, (lds_flag*)ready, (lds_flag*)freed, red,
                            (lds_flag*)&abort_flag, A.level_inner, A.spin_limit, A.status);
  }
}

// ---------------------------------------------------------------------------------
// Staging mode 4 (FP32 sources): stencil columns shared by the lanes of a wavefront.
//
// The targets of a wave are consecutive along i.  Where target l+c sits c source cells to
// the right of target l in the same row and vertical interval, column c of target l's
// stencil is column 0 of target l+c's.  Each lane therefore loads only column 0 of its own
// stencil (16 dword loads: 4 rows x 4 planes, coalesced across the wave) and reduces that
// column four times, over z and then y, with the weights of targets l, l-1, l-2 and l-3:
// the per-column part of the reference's operation order, so each partial is bit for bit
// what the owning target computes for that column itself.  The partials travel to their
// targets with three lane shifts; a column's min/max (target independent) with two.  The
// x combination then runs at the target in the usual order.  A tile is kSW = 61 targets
// wide, so lanes 61..63 load the columns 1..3 of the wave's last target.  A target whose
// right-hand neighbours do not continue its stencil (a cell skipped or repeated, another
// row or interval, a stencil out of range) is evaluated by per-lane row loads, as
// staging mode 2 does, in the same arithmetic.
constexpr int kSW = 61;

template <class T> __device__ __forceinline__ T fmx(T a, T b, T c);
template <> __device__ __forceinline__ float fmx<float>(float a, float b, float c) { return fmaf(a, b, c); }
template <> __device__ __forceinline__ double fmx<double>(double a, double b, double c) { return fma(a, b, c); }

// Weight slots one column reduction needs: y (4..7), z (8..11), linear y (16, 17), linear
// z (18, 19).
__device__ __forceinline__ constexpr int col_slot(int q) { return q < 8 ? 4 + q : 8 + q; }

// z then y over one column v[r][z] (rows r, planes z) with weights w[col_slot(.)]
template <class TC>
__device__ __forceinline__ TC col_tricub(const TC (&v)[4][4], const TC (&w)[12]) {
  TC ya = 0;
#pragma unroll
  for (int r = 0; r < 4; r++) {
    TC za = v[r][0] * w[4];
    za = fmx<TC>(v[r][1], w[5], za);
    za = fmx<TC>(v[r][2], w[6], za);
    za = fmx<TC>(v[r][3], w[7], za);
    ya = (r == 0) ? za * w[0] : fmx<TC>(za, w[r], ya);
  }
  return ya;
}

// the linear part of one column: rows 1..2, planes 1..2
template <class TC>
__device__ __forceinline__ TC col_lin(const TC (&v)[4][4], const TC (&w)[12]) {
  TC vz1 = v[1][1] * w[10];
  vz1 = fmx<TC>(v[1][2], w[11], vz1);
  TC vy = vz1 * w[8];
  TC vz2 = v[2][1] * w[10];
  vz2 = fmx<TC>(v[2][2], w[11], vz2);
  return fmx<TC>(vz2, w[9], vy);
}

// min / max of one column over rows 0..3 and planes 1..2: the operand rule of MINPD /
// MAXPD in FP64 (as mono_extra_a), IEEE minNum / maxNum in FP32
template <class TC>
__device__ __forceinline__ void col_minmax(const TC (&v)[4][4], TC* mi, TC* ma) {
  if constexpr (std::is_same<TC, float>::value) {
    float lo = fminf(v[0][1], v[0][2]), hi = fmaxf(v[0][1], v[0][2]);
#pragma unroll
    for (int r = 1; r < 4; r++) {
      lo = fminf(lo, fminf(v[r][1], v[r][2]));
      hi = fmaxf(hi, fmaxf(v[r][1], v[r][2]));
    }
    *mi = lo;
    *ma = hi;
  } else {
    TC lo = v[0][1], hi = v[0][1];
    lo = (lo < v[0][2]) ? lo : v[0][2];
    hi = (hi > v[0][2]) ? hi : v[0][2];
#pragma unroll
    for (int r = 1; r < 4; r++) {
      lo = (lo < v[r][1]) ? lo : v[r][1];
      hi = (hi > v[r][1]) ? hi : v[r][1];
      lo = (lo < v[r][2]) ? lo : v[r][2];
      hi = (hi > v[r][2]) ? hi : v[r][2];
    }
    *mi = lo;
    *ma = hi;
  }
}

template <int MATH, int POST, class TP, class Fields>
__device__ __forceinline__ void shared_tile(const DevItem& it, const Fields& F, long ptile, int f0, int f1,
                                            int level_inner) {
  typedef typename std::conditional<MATH == ADZ_MATH_REF, double, float>::type TC;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  int ib, jb, kk;
  tile_coords(it, ptile, level_inner, &ib, &jb, &kk, kSW);
  const int jj = jb * kTJ + wv;
  if (jj >= it.ny) return;  // a row past the item: this wave has no target (no barrier follows)
  const int ii = ib * kSW + lane;
  const bool target = lane < kSW && ii < it.nx;
  const int i = it.i0 + ii, j = it.j0 + jj, k = it.k0 + kk;
  const long p = (long)ii + (long)it.nx * (jj + (long)it.ny * kk);
  const long o = (long)(i - 1) + (long)it.dni * ((j - 1) + (long)it.dnj * (k - 1));
  Index x;
  bool ok = false;
  if (target) {
    double px, py, pz;
    load_pos<TP>(it.xyz, p, &px, &py, &pz);
    ok = index_step(px, py, pz, it.lv, &x);
  }
  double cw[20];  // own weights (FP64, or FP32 in the storage of the first 10 slots)
  if (ok) {
    if (MATH == ADZ_MATH_REF) weights_ref(x, it.lv, cw);
    else weights_f32(x, it.lv, (float*)cw);
  }
  const TC* c = (const TC*)cw;
  if (!it.span) {  // sources too large for 32-bit buffer offsets: per-lane global loads
    per_lane<float, MATH, POST>(it, F, f0, f1, x, ok, target, cw, o);
    return;
  }
  // stencil corner of each lane (element offset from minoff), its column, interval kind
  int e = ok ? (int)(x.off - it.minoff) : 0, ix = ok ? x.ix : 0, zl = ok ? x.zlinear : 0, ld = ok ? 1 : 0;
  // lanes past the wave's last target continue its stencil to the right (its columns 1..3)
  const int last = min(kSW - 1, it.nx - 1 - ib * kSW);
  {
    const int eL = __shfl(e, last, 64), ixL = __shfl(ix, last, 64), zlL = __shfl(zl, last, 64),
              okL = __shfl(ld, last, 64);
    if (lane > last) {
      const int d = lane - last;
      e = eL + d; ix = ixL + d; zl = zlL; ld = (okL && d <= 3) ? 1 : 0;
    }
  }
  // does column c of this target come from lane + c?
  bool fast = ok;
#pragma unroll
  for (int cc = 1; cc < 4; cc++) {
    const int en = __shfl_down(e, cc, 64), ixn = __shfl_down(ix, cc, 64), zln = __shfl_down(zl, cc, 64),
              ldn = __shfl_down(ld, cc, 64);
    fast = fast && ldn && en == e + cc && ixn == ix + cc && zln == zl;
  }
  const bool slow_any = __ballot(target && ok && !fast) != 0;
  // weights of targets lane-1, lane-2, lane-3 for their columns 1, 2, 3 held here
  TC w[4][12];
#pragma unroll
  for (int q = 0; q < 12; q++) w[0][q] = c[col_slot(q)];
#pragma unroll
  for (int cc = 1; cc < 4; cc++)
#pragma unroll
    for (int q = 0; q < 12; q++) w[cc][q] = __shfl_up(w[0][q], cc, 64);
  // row offsets of the 4 planes of row 0; rows r add r*ni (uniform: the scalar offset)
  int bo[4];
  {
    const int nij = (int)it.lv.nij;
    const int o1 = zl ? 0 : nij, o2 = o1 + nij, o3 = o2 + o1;
    bo[0] = e * 4; bo[1] = (e + o1) * 4; bo[2] = (e + o2) * 4; bo[3] = (e + o3) * 4;
  }
  const int rstride = (int)(it.lv.ni * 4);
  for (int f = f0; f < f1; f++) {
    TC Y[4] = {0, 0, 0, 0}, Lc[4] = {0, 0, 0, 0}, cmn = 0, cmx = 0;
    if (ld) {
      const __amdgpu_buffer_rsrc_t rs = field_rsrc((const float*)F.s(f) + it.minoff, it.span);
      TC v[4][4];
#pragma unroll
      for (int r = 0; r < 4; r++)
#pragma unroll
        for (int z = 0; z < 4; z++)
          v[r][z] = (TC)__builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rs, bo[z], r * rstride, 0));
#pragma unroll
      for (int cc = 0; cc < 4; cc++) Y[cc] = col_tricub<TC>(v, w[cc]);
      if (POST) {
#pragma unroll
        for (int cc = 0; cc < 4; cc++) Lc[cc] = col_lin<TC>(v, w[cc]);
        col_minmax<TC>(v, &cmn, &cmx);
      }
    }
    // column cc of target `lane` was reduced by lane + cc
#pragma unroll
    for (int cc = 1; cc < 4; cc++) Y[cc] = __shfl_down(Y[cc], cc, 64);
    TC mi1 = 0, mi2 = 0, ma1 = 0, ma2 = 0;
    if (POST) {
#pragma unroll
      for (int cc = 1; cc < 4; cc++) Lc[cc] = __shfl_down(Lc[cc], cc, 64);
      mi1 = __shfl_down(cmn, 1, 64); mi2 = __shfl_down(cmn, 2, 64);
      ma1 = __shfl_down(cmx, 1, 64); ma2 = __shfl_down(cmx, 2, 64);
    }
    if (fast) {
      TC v;
      if (MATH == ADZ_MATH_REF) {
        const TC y0 = Y[0] * c[0], y1 = Y[1] * c[1], y2 = Y[2] * c[2], y3 = Y[3] * c[3];
        v = (y0 + y2) + (y1 + y3);
      } else {
        v = fmx<TC>(Y[3], c[3], fmx<TC>(Y[2], c[2], fmx<TC>(Y[1], c[1], Y[0] * c[0])));
      }
      TC l = 0, mi = 0, ma = 0;
      if (POST) {
        const TC l0 = Lc[0] * c[12], l1 = Lc[1] * c[13], l2 = Lc[2] * c[14], l3 = Lc[3] * c[15];
        l = (l0 + l2) + (l1 + l3);
        if (MATH == ADZ_MATH_REF) {
          mi = (mi1 < mi2) ? mi1 : mi2;
          ma = (ma1 > ma2) ? ma1 : ma2;
        } else {
          mi = fminf(mi1, mi2);
          ma = fmaxf(ma1, ma2);
        }
      }
      store_out<float, MATH, POST>(F, f, o, v, l, mi, ma);
    } else if (target && !ok) {
      store_nan<float, MATH, POST>(F, f, o);
    }
  }
  // targets whose stencil is not continued by their neighbours: per-lane row loads
  if (slow_any) per_lane<float, MATH, POST>(it, F, f0, f1, x, target && ok && !fast, false, cw, o);
}

// ---------------------------------------------------------------------------------
// Staging mode 5 (FP32 sources): the tile's box (the union of its 256 stencils) is copied
// into LDS once per field by all four waves with 16-byte buffer-to-LDS DMA, and every lane
// then reads its 64 stencil values from LDS.  Per-lane loads spend 16 address slots of the
// texture unit per point-field; the box needs about 2 (a 68 x 7 x 4 box is 28 rows of 18
// quads for 256 targets).  LDS rows have a fixed pitch of 20 quads (80 floats): the rows of
// a buffer are dense, each DMA instruction fills 64 consecutive quads of them, and the 16
// values of one stencil plane lie within 1 KB of one per-lane base, so they are read in
// pairs at immediate offsets.  The DMA starts each row at a 16-byte boundary, so a field
// whose base is not 16-byte aligned lands shifted by `sh` floats (0..3), added to the
// bases.  Two LDS buffers with static roles: field f+1 is in flight while field f is
// evaluated; one block barrier per field.  Tiles whose box does not fit (more than
// kMaxRows5 rows or 77 columns), rows not a multiple of 4 floats apart, or sources beyond
// 32-bit buffer offsets use per-lane loads.  The arithmetic and its order are those of the
// per-lane path, so both give the same bits.  Staging mode 6 is the same with tiles of 64 x 8
// targets and blocks of eight waves: a box of 8 target rows spans ~12 source rows against
// ~7 for 4, so about a fifth fewer DMA bytes per target, and the tile's fixed work (box
// extents, DMA offsets) is shared by twice the targets.  Staging modes 7 and 8 take tiles of
// 64 x 4 targets over 2 and 4 consecutive levels (blocks of 8 and 16 waves, one wave per
// target row and level).  A source plane is read by the stencils of about four levels, so
// the planes of a box dominate its volume: at c5, a box over 64 x 4 x 4 targets holds ~7
// rows of ~7 planes against ~11 rows of 4 planes for 64 x 8 x 1, about 14 against 25 DMA
// bytes per target.
constexpr int kQRow = 20;                               // 16-byte quads per LDS row
constexpr int kPitch5 = 4 * kQRow;                      // floats per LDS row
constexpr int kMaxRows5 = kMaxRows * kPitch / kPitch5;  // rows per buffer (the stage-1 buffers)
constexpr int kTJ6 = 8;                                 // target rows of a staging-mode-6 tile
constexpr int kMaxRows8 = 96;                           // rows per buffer of the 16-wave blocks
// rows per LDS buffer for blocks of W waves: 16-wave blocks run one per CU and take larger
// buffers (2 x 30 KB); the others keep the stage-1 buffer size, two blocks per CU
template <int W> constexpr int box_rows() { return W >= 16 ? kMaxRows8 : kMaxRows5; }
template <int W> constexpr int dma_max() { return (box_rows<W>() * kQRow + 64 * W - 1) / (64 * W); }
static_assert(4 * dma_max<4>() * 64 >= kMaxRows5 * kQRow && 8 * dma_max<8>() * 64 >= kMaxRows5 * kQRow &&
                  16 * dma_max<16>() * 64 >= kMaxRows8 * kQRow && dma_max<4>() <= 32,
              "DMA slots cover a full buffer");
static_assert(kMaxRows5 < 128 && kMaxRows8 < 128, "row split by a float reciprocal below");

template <int MATH, int POST, class TP, int TH, int TL, class Fields>
__device__ __forceinline__ void dma_tile(const DevItem& it, const Fields& F, long ptile, int f0, int f1,
                                         float* boxA, float* boxB, int* red, int level_inner) {
  constexpr int W = TH * TL;                  // waves of the block: one per target row and level
  constexpr int kRows = box_rows<W>();
  constexpr int kDmaMax = dma_max<W>();       // DMA instructions per wave and field
  const int lane = threadIdx.x & (kTI - 1), wv = threadIdx.x / kTI;
  const TileXY tc = tile_at(it, ptile, level_inner, kTI, TH, TL);
  const int ii = tc.ib * kTI + lane, jj = tc.jb * TH + wv % TH, kk = tc.kk * TL + wv / TH;
  const bool valid = ii < it.nx && jj < it.ny && kk < it.nz;
  const int i = it.i0 + ii, j = it.j0 + jj, k = it.k0 + kk;
  const long p = (long)ii + (long)it.nx * (jj + (long)it.ny * kk);
  const long o = (long)(i - 1) + (long)it.dni * ((j - 1) + (long)it.dnj * (k - 1));
  Index x;
  bool ok = false;
  if (valid) {
    double px, py, pz;
    load_pos<TP>(it.xyz, p, &px, &py, &pz);
    ok = index_step(px, py, pz, it.lv, &x);
  }
  double cw[20];
  if (ok) {
    if (MATH == ADZ_MATH_REF) weights_ref(x, it.lv, cw);
    else weights_f32(x, it.lv, (float*)cw);
  }
  const int big = 1 << 30;
  int ext[6] = {ok ? x.ix - 1 : big, ok ? -(x.ix + 2) : big, ok ? x.iy - 1 : big, ok ? -(x.iy + 2) : big,
                ok ? x.pz : big, ok ? -(x.pz + (x.zlinear ? 1 : 3)) : big};
  block_min6<W>(ext, red);
#pragma unroll
  for (int q = 0; q < 6; q++) ext[q] = __builtin_amdgcn_readfirstlane(ext[q]);  // block-uniform
  const int x0 = ext[0], y0 = ext[2], z0 = ext[4];
  if (x0 == big) {  // no target of the tile has an in-range stencil
    if (valid)
      for (int f = f0; f < f1; f++) store_nan<float, MATH, POST>(F, f, o);
    return;
  }
  const int nxb = -ext[1] - x0 + 1, nyb = -ext[3] - y0 + 1, nzb = -ext[5] - z0 + 1;
  const int nrows = nyb * nzb, cpr = (nxb + 3 + 3) / 4;  // quads per row, room for the shift
  const long gni = it.lv.ni, gnij = it.lv.nij;
  // every quad the DMA reads, for any shift 0..3, inside the buffer's range [minoff,
  // minoff + span): a quad reaching outside would read zeros for its in-range floats too
  const long first = (long)(x0 - 1 - 3) + (long)(y0 - 1) * gni + (long)z0 * gnij - it.minoff;
  const long last = (long)(x0 - 1) + 4L * cpr - 1 + (long)(y0 + nyb - 2) * gni + (long)(z0 + nzb - 1) * gnij -
                    it.minoff;
  if (nrows > kRows || cpr > kQRow || !it.span || (gni & 3) || (gnij & 3) || first < 0 ||
      (last + 1) * 4 > (long)it.span) {
    per_lane<float, MATH, POST>(it, F, f0, f1, x, ok, valid, cw, o);
    return;
  }
  // this lane's DMA slots: instruction q of wave wv fills the 64 quads g*64 .. g*64+63 of
  // the buffer (g = wv + W q), quad Q = row Q / kQRow, column quad Q % kQRow; element
  // offsets from minoff of the lane's quad (before the field's shift)
  // (32-bit arithmetic: every offset is below span < 2^31 bytes; R / nyb by a float
  // reciprocal, exact since R < 128 and nyb < 128 keep (R + 0.5) / nyb >= 1/256 from an
  // integer, far above the float product's error; slots with R >= nrows are inactive)
  const int rel0 = (int)__builtin_amdgcn_readfirstlane((int)((long)(x0 - 1) + (long)(y0 - 1) * gni +
                                                             (long)z0 * gnij - it.minoff));
  const int ni32 = (int)gni, nij32 = (int)gnij;
  const float rnyb = 1.0f / (float)nyb;
  int rel[kDmaMax];
  unsigned act = 0;
#pragma unroll
  for (int q = 0; q < kDmaMax; q++) {
    const int Q = (wv + W * q) * 64 + lane;
    const int R = Q / kQRow, cq = Q - R * kQRow;
    const int zz = (int)(((float)R + 0.5f) * rnyb), yy = R - zz * nyb;
    rel[q] = rel0 + 4 * cq + yy * ni32 + zz * nij32;
    if (R < nrows && cq < cpr) act |= 1u << q;
  }
  // plane bases of this lane's stencil in a buffer (floats, before the field's shift)
  int lb[4] = {0, 0, 0, 0};
  if (ok) {
    const int r0 = x.iy - 1 - y0, pz0 = x.pz - z0, c0 = x.ix - 1 - x0;
#pragma unroll
    for (int z = 0; z < 4; z++) {
      const int dz = x.zlinear ? (z >> 1) : z;
      lb[z] = (r0 + nyb * (pz0 + dz)) * kPitch5 + c0;
    }
  }
  auto shift = [&](int f) {  // floats between the field's 16-byte boundary and column x0
    const unsigned long long b = (unsigned long long)F.s(f);
    return (int)(((b >> 2) + (unsigned)(x0 - 1)) & 3);
  };
  auto issue = [&](int f, float* buf) {
    const int sh = shift(f);
    const __amdgpu_buffer_rsrc_t rs = field_rsrc((const float*)F.s(f) + it.minoff, it.span);
#pragma unroll
    for (int q = 0; q < kDmaMax; q++) {
      if (act & (1u << q))
        __builtin_amdgcn_raw_ptr_buffer_load_lds(
            rs, (__attribute__((address_space(3))) void*)(buf + (wv + W * q) * 256), 16,
            (rel[q] - sh) * 4, 0, 0, 0);
    }
  };
  // Every lane evaluates (one without a stencil reads its buffer's first values and stores
  // NaN where valid): with one store site the field index and output pointers stay scalar.
  typedef typename std::conditional<MATH == ADZ_MATH_REF, double, float>::type TC;
  auto eval = [&](int f, const float* buf) {
    const int sh = shift(f);
    const float* b0 = buf + lb[0] + sh;
    const float* b1 = buf + lb[1] + sh;
    const float* b2 = buf + lb[2] + sh;
    const float* b3 = buf + lb[3] + sh;
    auto at = [&](int r, int z, int c) -> float {
      const float* bz = z == 0 ? b0 : (z == 1 ? b1 : (z == 2 ? b2 : b3));
      return bz[r * kPitch5 + c];
    };
    TC v = 0, l = 0, mi = 0, ma = 0;
    eval_vals<float, MATH, POST>(at, cw, &v, &l, &mi, &ma);
    if (!ok) {
      const TC nan = (TC)__builtin_nanf("");
      v = nan; l = nan; mi = nan; ma = nan;
    }
    if (valid) store_out<float, MATH, POST>(F, f, o, v, l, mi, ma);
  };
  issue(f0, boxA);
  asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
  __syncthreads();
  for (int f = f0; f < f1; f += 2) {
    if (f + 1 < f1) issue(f + 1, boxB);
    eval(f, boxA);
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();
    if (f + 1 >= f1) break;
    if (f + 2 < f1) issue(f + 2, boxA);
    eval(f + 1, boxB);
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();
  }
}

// Persistent loop over the tile list of a plan.
template <class TF, class TP, int MATH, int POST, int STAGE>
__device__ __forceinline__ void plan_loop(const PlanArgs& A, TF* boxA, TF* boxB, int* red, int* rowoff) {
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
    // The item table does not change during a launch: read through the constant address
    // space, the (uniform) item's fields come in with scalar loads.
    typedef const DevItem __attribute__((address_space(4))) CItem;
    const CItem* items = (const CItem*)A.items;
    while (q + 1 < A.nitems && items[q + 1].tile0 <= t) q++;
    DevItem it;
    {
      static_assert(sizeof(DevItem) % 8 == 0, "DevItem is copied in 8-byte words");
      typedef const unsigned long long __attribute__((address_space(4))) CWord;
      const CWord* w = (const CWord*)(A.items + q);
      unsigned long long* d = (unsigned long long*)&it;
#pragma unroll
      for (int e = 0; e < (int)(sizeof(DevItem) / 8); e++) d[e] = w[e];
    }
    const long local = t - it.tile0;
    long ptile, grp;
    if (local < (1L << 31)) {  // 32-bit division (the usual case)
      const unsigned lo32 = (unsigned)local;
      if (A.points_outer) { const unsigned q = lo32 / (unsigned)it.ngroups; ptile = q; grp = lo32 - q * (unsigned)it.ngroups; }
      else { const unsigned q = lo32 / (unsigned)it.ptiles; grp = q; ptile = lo32 - q * (unsigned)it.ptiles; }
    } else {
      if (A.points_outer) { ptile = local / it.ngroups; grp = local % it.ngroups; }
      else { grp = local / it.ptiles; ptile = local % it.ptiles; }
    }
    int f0 = (int)grp * A.g, f1 = min(it.nf, f0 + A.g);
    ItemFields F{it.src, it.dst, it.lin, it.mn, it.mx};
    if constexpr (STAGE == 4 && std::is_same<TF, float>::value)
      shared_tile<MATH, POST, TP>(it, F, ptile, f0, f1, A.level_inner);
    else if constexpr (STAGE == 5 && std::is_same<TF, float>::value)
      dma_tile<MATH, POST, TP, kTJ, 1>(it, F, ptile, f0, f1, (float*)boxA, (float*)boxB, red, A.level_inner);
    else if constexpr (STAGE == 6 && std::is_same<TF, float>::value)
      dma_tile<MATH, POST, TP, kTJ6, 1>(it, F, ptile, f0, f1, (float*)boxA, (float*)boxB, red, A.level_inner);
    else if constexpr (STAGE == 7 && std::is_same<TF, float>::value)
      dma_tile<MATH, POST, TP, kTJ, 2>(it, F, ptile, f0, f1, (float*)boxA, (float*)boxB, red, A.level_inner);
    else if constexpr (STAGE == 8 && std::is_same<TF, float>::value)
      dma_tile<MATH, POST, TP, kTJ, 4>(it, F, ptile, f0, f1, (float*)boxA, (float*)boxB, red, A.level_inner);
    else
      run_tile<TF, TP, MATH, POST, ItemFields>(it, F, ptile, f0, f1, boxA, boxB, red, rowoff, STAGE, A.level_inner);
  }
}

// Persistent kernel: STAGE 1 stages each tile's source box in LDS, STAGE 2 reads per lane,
// STAGE 4 shares stencil columns across the lanes of a wave (separate instances, so the
// per-lane ones carry neither the LDS buffers nor the staging path's registers).
template <class TF, class TP, int MATH, int POST, int STAGE>
__global__ __launch_bounds__(kTile) void k_plan(PlanArgs A) {
  if constexpr (STAGE == 1) {
    __shared__ TF boxA[kMaxRows * kPitch];
    __shared__ TF boxB[kMaxRows * kPitch];
    __shared__ int red[6 * kTJ];
    __shared__ int rowoff[kMaxRows];
    plan_loop<TF, TP, MATH, POST, 1>(A, boxA, boxB, red, rowoff);
  } else {
    plan_loop<TF, TP, MATH, POST, STAGE>(A, nullptr, nullptr, nullptr, nullptr);
  }
}

// Staging modes 5 to 8 (FP32 sources, tiles of TH target rows x TL levels, one wave per row
// and level): their own kernel.  In the FP32 arithmetic it is held to 128 VGPRs, so that the
// blocks its two LDS buffers allow per CU keep four waves on every SIMD; the reference
// arithmetic needs more registers than that without spilling (and has no 16-wave form).
template <int TH, int TL> constexpr int dma_stage() { return TL == 4 ? 8 : TL == 2 ? 7 : TH == kTJ ? 5 : 6; }
template <class TP, int MATH, int POST, int TH, int TL>
__global__ __launch_bounds__(64 * TH * TL, MATH == ADZ_MATH_F32 ? 4 : 1) void k_plan5(PlanArgs A) {
  __shared__ float boxA[box_rows<TH * TL>() * kPitch5];
  __shared__ float boxB[box_rows<TH * TL>() * kPitch5];
  __shared__ int red[6 * TH * TL];
  plan_loop<float, TP, MATH, POST, dma_stage<TH, TL>()>(A, boxA, boxB, red, nullptr);
}

// One item with inline field pointers (the adz_tricub_hlt entry point).
template <class TF, class TP, int MATH, int POST>
__global__ __launch_bounds__(kTile) void k_hlt(DevItem it, InlineFields F) {
  run_tile<TF, TP, MATH, POST, InlineFields>(it, F, blockIdx.x, 0, it.nf, (TF*)nullptr, (TF*)nullptr, nullptr,
                                             nullptr, 2, 0);
}

// The same for FP32 sources with LDS-staged boxes (staging mode 6: tiles of 64 x 8 targets,
// blocks of eight waves): every field of the call in one pass per tile, tiles in
// super-columns of kHltOrder tile rows, and block b on tile
// (b % 8) * per + b / 8, so that each XCD (blocks go round-robin over the 8 XCDs) walks one
// contiguous eighth of the tile list and its L2 sees neighbouring tiles.  One tile per
// block: without lin/min/max the kernel needs 72 VGPRs, so seven waves share each SIMD.
// (The 16-wave form of 64 x 4 x 4 tiles moves fewer bytes, but one such block is resident per
// CU and its per-field DMA waits are not covered: at c5 with 16 fields and no lin/min/max it
// gave 150-154 against 200 G point-fields/s.)
constexpr int kHltOrder = 3;

template <class TP, int MATH, int POST>
__global__ __launch_bounds__(64 * kTJ6, MATH == ADZ_MATH_F32 ? 4 : 1) void k_hlt5(DevItem it, InlineFields F,
                                                                                   long ntiles) {
  __shared__ float boxA[box_rows<kTJ6>() * kPitch5];
  __shared__ float boxB[box_rows<kTJ6>() * kPitch5];
  __shared__ int red[6 * kTJ6];
  const long per = (ntiles + 7) / 8;
  const long t = (long)(blockIdx.x % 8) * per + blockIdx.x / 8;
  if (t >= ntiles) return;  // the whole block leaves: no barrier is reached
  dma_tile<MATH, POST, TP, kTJ6, 1>(it, F, t, 0, it.nf, boxA, boxB, red, kHltOrder);
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
  {
    const ZtabDev& v = d->lv;
    d->minoff = (long)(v.ilo - 1) + (long)(v.jlo - 1) * v.ni;
    long maxoff = (long)(v.ihi - 1) + (long)(v.jhi - 1) * v.ni + (long)(v.nk - 1) * v.nij;
    long bytes = (maxoff - d->minoff + 1) * 4;
    d->span = (bytes > 0 && bytes < 0x7fffffffL) ? (unsigned)bytes : 0u;
  }
  d->npts = (long)d->nx * d->ny * d->nz;
  d->ptiles = (long)((d->nx + kTI - 1) / kTI) * ((d->ny + kTJ - 1) / kTJ) * d->nz;
  return true;
}

// select the kernel instance for (fdtype, pdtype, math, post)
template <class TF, class TP, int M>
void* pick_plan(int post, int stage) {
  if constexpr (std::is_same<TF, float>::value) {
    if (stage == 4) {
      switch (post) {
        case 0: return (void*)&k_plan<TF, TP, M, 0, 4>;
        case 1: return (void*)&k_plan<TF, TP, M, 1, 4>;
        case 2: return (void*)&k_plan<TF, TP, M, 2, 4>;
        default: return (void*)&k_plan<TF, TP, M, 3, 4>;
      }
    }
    if (stage == 5) {
      switch (post) {
        case 0: return (void*)&k_plan5<TP, M, 0, kTJ, 1>;
        case 1: return (void*)&k_plan5<TP, M, 1, kTJ, 1>;
        case 2: return (void*)&k_plan5<TP, M, 2, kTJ, 1>;
        default: return (void*)&k_plan5<TP, M, 3, kTJ, 1>;
      }
    }
    if (stage == 6) {
      switch (post) {
        case 0: return (void*)&k_plan5<TP, M, 0, kTJ6, 1>;
        case 1: return (void*)&k_plan5<TP, M, 1, kTJ6, 1>;
        case 2: return (void*)&k_plan5<TP, M, 2, kTJ6, 1>;
        default: return (void*)&k_plan5<TP, M, 3, kTJ6, 1>;
      }
    }
    if (stage == 7) {
      switch (post) {
        case 0: return (void*)&k_plan5<TP, M, 0, kTJ, 2>;
        case 1: return (void*)&k_plan5<TP, M, 1, kTJ, 2>;
        case 2: return (void*)&k_plan5<TP, M, 2, kTJ, 2>;
        default: return (void*)&k_plan5<TP, M, 3, kTJ, 2>;
      }
    }
    if constexpr (M == ADZ_MATH_F32) {
      if (stage == 8) {
        switch (post) {
          case 0: return (void*)&k_plan5<TP, M, 0, kTJ, 4>;
          case 1: return (void*)&k_plan5<TP, M, 1, kTJ, 4>;
          case 2: return (void*)&k_plan5<TP, M, 2, kTJ, 4>;
          default: return (void*)&k_plan5<TP, M, 3, kTJ, 4>;
        }
      }
    }
    if (stage == 3) {
      switch (post) {
        case 0: return (void*)&k_plan_pc<TP, M, 0>;
        case 1: return (void*)&k_plan_pc<TP, M, 1>;
        case 2: return (void*)&k_plan_pc<TP, M, 2>;
        default: return (void*)&k_plan_pc<TP, M, 3>;
      }
    }
  }
  if (stage == 1) {
    switch (post) {
      case 0: return (void*)&k_plan<TF, TP, M, 0, 1>;
      case 1: return (void*)&k_plan<TF, TP, M, 1, 1>;
      case 2: return (void*)&k_plan<TF, TP, M, 2, 1>;
      default: return (void*)&k_plan<TF, TP, M, 3, 1>;
    }
  }
  switch (post) {
    case 0: return (void*)&k_plan<TF, TP, M, 0, 2>;
    case 1: return (void*)&k_plan<TF, TP, M, 1, 2>;
    case 2: return (void*)&k_plan<TF, TP, M, 2, 2>;
    default: return (void*)&k_plan<TF, TP, M, 3, 2>;
  }
}
template <class TF, class TP, int M>
void* pick_hlt(int post) {
  if constexpr (std::is_same<TF, float>::value) {
    switch (post) {
      case 0: return (void*)&k_hlt5<TP, M, 0>;
      case 1: return (void*)&k_hlt5<TP, M, 1>;
      case 2: return (void*)&k_hlt5<TP, M, 2>;
      default: return (void*)&k_hlt5<TP, M, 3>;
    }
  }
  switch (post) {
    case 0: return (void*)&k_hlt<TF, TP, M, 0>;
    case 1: return (void*)&k_hlt<TF, TP, M, 1>;
    case 2: return (void*)&k_hlt<TF, TP, M, 2>;
    default: return (void*)&k_hlt<TF, TP, M, 3>;
  }
}

template <bool PLAN>
void* pick(int fdt, int pdt, int math, int post, int stage = 2) {
#define P_(TF, TP, M) return PLAN ? pick_plan<TF, TP, M>(post, stage) : pick_hlt<TF, TP, M>(post)
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

// Built-in tile parameters, ten fields per pass: for FP32 fields in the FP32 arithmetic,
// LDS-staged boxes over tiles of 64 x 4 targets x 4 levels (stage 8) with one 16-wave block
// per CU, the levels of a tile column walked first; otherwise LDS-staged boxes over 64 x 4
// tiles (stage 5; other field types fall back to per-lane loads) in super-columns of three
// tile rows.
adz_tune default_tune(int fdtype, int math) {
  return fdtype == ADZ_F32 && math == ADZ_MATH_F32 ? adz_tune{10, 0, 1, 1, 8, 1} : adz_tune{10, 0, 4, 1, 5, 3};
}
// targets per tile along i: kSW for the shared-column staging of FP32 sources, else kTI

int tile_width(int stage, int fdtype) { return (stage == 4 && fdtype == ADZ_F32) ? kSW : kTI; }
int tile_height(int stage, int fdtype) { return (stage == 6 && fdtype == ADZ_F32) ? kTJ6 : kTJ; }
int tile_levels(int stage, int fdtype) {
  return fdtype != ADZ_F32 ? 1 : stage == 7 ? 2 : stage == 8 ? 4 : 1;
}
// threads per block of a plan's kernel
unsigned plan_threads(int stage, int fdtype) {
  if (fdtype != ADZ_F32) return kTile;
  return stage == 3 ? kPCThreads : stage == 6 || stage == 7 ? 512 : stage == 8 ? 1024 : kTile;
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
    if (item_ptiles(it.nx, it.ny, it.nz, kHltOrder) >= (1L << 31) - 8)
      return fail(ADZ_EINVAL, "adz_tricub_hlt: a box of 2^31 tiles or more");
    InlineFields F;
    std::memset(&F, 0, sizeof(F));
    for (int f = 0; f < nf; f++) {
      const adz_field& a = fld[f0 + f];
      F.src[f] = a.src; F.dst[f] = a.dst; F.lin[f] = a.lin; F.mn[f] = a.mn; F.mx[f] = a.mx;
    }
    if (fdtype == ADZ_F32) {  // k_hlt5: LDS-staged boxes, an eighth of the tile list per XCD
      long ntiles = item_ptiles(it.nx, it.ny, it.nz, kHltOrder, kTI, kTJ6);
      long grid = 8 * ((ntiles + 7) / 8);  // one block per tile
      void* args[] = {&it, &F, &ntiles};
      ADZ_HIP_CHECK(hipLaunchKernel(kern, dim3((unsigned)grid), dim3(64 * kTJ6), args, 0, (hipStream_t)stream));
    } else {
      void* args[] = {&it, &F};
      ADZ_HIP_CHECK(hipLaunchKernel(kern, dim3((unsigned)it.ptiles), dim3(kTile), args, 0, (hipStream_t)stream));
    }
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
  if (tune && (tune->stage < 0 || tune->stage > 8)) return fail(ADZ_EINVAL, "adz_plan_create: stage must be 0..8");
  if (tune && tune->stage == 8 && fdtype == ADZ_F32 && math != ADZ_MATH_F32)
    return fail(ADZ_EINVAL, "adz_plan_create: stage 8 needs the FP32 arithmetic");
  if (tune && (tune->xcd_chunks < 0 || tune->xcd_chunks > 1))
    return fail(ADZ_EINVAL, "adz_plan_create: xcd_chunks must be 0 or 1");
  adz_tune T = default_tune(fdtype, math);
  if (tune) {
    if (tune->fields_per_pass > 0) T.fields_per_pass = std::min(64, tune->fields_per_pass);
    T.points_outer = tune->points_outer;
    if (tune->blocks_per_cu > 0) T.blocks_per_cu = tune->blocks_per_cu;
    T.xcd_chunks = tune->xcd_chunks;
    if (tune->stage > 0) {  // stage 0 keeps the built-in staging and its tile order
      T.stage = tune->stage;
      T.level_inner = tune->level_inner;
    }
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
    dv[q].ptiles = item_ptiles(dv[q].nx, dv[q].ny, dv[q].nz, T.level_inner, tile_width(T.stage, fdtype),
                               tile_height(T.stage, fdtype), tile_levels(T.stage, fdtype));
    if (dv[q].ptiles >= (1L << 31)) return fail(ADZ_EINVAL, "adz_plan_create: an item of 2^31 tiles or more");
    dv[q].tile0 = tiles;
    tiles += dv[q].ptiles * dv[q].ngroups;
    dv[q].src = (const void* const*)(intptr_t)(0 * nptr_total + pofs);  // patched below
    dv[q].dst = (void* const*)(intptr_t)(1 * nptr_total + pofs);
    dv[q].lin = (void* const*)(intptr_t)(2 * nptr_total + pofs);
    dv[q].mn = (void* const*)(intptr_t)(3 * nptr_total + pofs);
    dv[q].mx = (void* const*)(intptr_t)(4 * nptr_total + pofs);
    pofs += a.nptr;
  }
  size_t ib = sizeof(DevItem) * nitems, pb = sizeof(void*) * ptrs.size(), sb = 64;  // sb: the status word
  void* mem = nullptr;
  if (hipMalloc(&mem, ib + pb + sb) != hipSuccess) return fail(ADZ_ENOMEM, "adz_plan_create: hipMalloc");
  void** dptr = (void**)((char*)mem + ib);
  for (auto& d : dv) {
    d.src = (const void* const*)(dptr + (intptr_t)d.src);
    d.dst = (void* const*)(dptr + (intptr_t)d.dst);
    d.lin = (void* const*)(dptr + (intptr_t)d.lin);
    d.mn = (void* const*)(dptr + (intptr_t)d.mn);
    d.mx = (void* const*)(dptr + (intptr_t)d.mx);
  }
  int* status = (int*)((char*)mem + ib + pb);
  if (hipMemcpy(mem, dv.data(), ib, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(dptr, ptrs.data(), pb, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemset(status, 0, sb) != hipSuccess) {
    (void)hipFree(mem);
    return fail(ADZ_EHIP, "adz_plan_create: upload");
  }
  adz_plan* p = new adz_plan;
  p->fdtype = fdtype; p->pdtype = pdt; p->math = math; p->post = post; p->tune = T; p->dmem = mem;
  p->args = PlanArgs{(const DevItem*)mem, nitems, tiles, T.fields_per_pass, T.points_outer, T.xcd_chunks, T.stage,
                     T.level_inner, 1 << 20, status};
  long grid = (long)T.blocks_per_cu * device_cus();
  if (grid > tiles) grid = tiles;
  if (T.xcd_chunks) grid = std::max(8L, (grid / 8) * 8);
  p->grid = (int)grid;
  *out = p;
  return ADZ_OK;
}

int adz_plan_execute(const adz_plan* p, adz_stream stream) {
  if (!p) return fail(ADZ_EINVAL, "adz_plan_execute: null plan");
  void* kern = pick<true>(p->fdtype, p->pdtype, p->math, p->post, p->tune.stage);
  PlanArgs a = p->args;
  void* args[] = {&a};
  const unsigned nthr = plan_threads(p->tune.stage, p->fdtype);
  ADZ_HIP_CHECK(hipLaunchKernel(kern, dim3(p->grid), dim3(nthr), args, 0, (hipStream_t)stream));
  return ADZ_OK;
}

int adz_plan_status(adz_plan* p, int* aborted_waits) {
  if (!p || !aborted_waits) return fail(ADZ_EINVAL, "adz_plan_status: null argument");
  int v = 0;
  ADZ_HIP_CHECK(hipMemcpy(&v, p->args.status, sizeof(int), hipMemcpyDeviceToHost));
  if (v) ADZ_HIP_CHECK(hipMemset(p->args.status, 0, sizeof(int)));
  *aborted_waits = v;
  return ADZ_OK;
}

int adz_plan_set_spin_limit(adz_plan* p, int polls) {
  if (!p || polls < 0) return fail(ADZ_EINVAL, "adz_plan_set_spin_limit: bad argument");
  p->args.spin_limit = polls;
  return ADZ_OK;
}

int adz_plan_tune(const adz_plan* p, adz_tune* out) {
  if (!p || !out) return fail(ADZ_EINVAL, "adz_plan_tune: null argument");
  *out = p->tune;
  return ADZ_OK;
}

// Autotuner: a search over the tile parameters, each candidate timed with HIP events on
// the caller's plan (its destinations are written, with the values any execution writes).
// Staging and tile order are searched jointly, then fields per pass, tile order again,
// persistent grid size, loop order and XCD chunking one coordinate at a time; each step
// keeps the best value found before moving on.
int adz_plan_autotune(const adz_item* items, int nitems, int fdtype, int math, int post, int reps,
                      adz_stream stream, adz_tune* best, float* best_ms) {
  if (!best) return fail(ADZ_EINVAL, "adz_plan_autotune: null result");
  if (reps <= 0) reps = 3;
  int maxf = 1;
  for (int q = 0; q < nitems && items; q++) maxf = std::max(maxf, items[q].nptr);
  hipStream_t s = (hipStream_t)stream;
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess) return fail(ADZ_EHIP, "adz_plan_autotune: event");
  if (hipEventCreate(&e1) != hipSuccess) { (void)hipEventDestroy(e0); return fail(ADZ_EHIP, "adz_plan_autotune: event"); }
  auto time_one = [&](const adz_tune& t, float* ms) -> int {
    adz_plan* p = nullptr;
    int rc = adz_plan_create(items, nitems, fdtype, math, post, &t, &p);
    if (rc != ADZ_OK) return rc;
    rc = adz_plan_execute(p, stream);  // warm-up
    if (rc == ADZ_OK && hipEventRecord(e0, s) != hipSuccess) rc = fail(ADZ_EHIP, "adz_plan_autotune: record");
    for (int r = 0; r < reps && rc == ADZ_OK; r++) rc = adz_plan_execute(p, stream);
    if (rc == ADZ_OK && (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                         hipEventElapsedTime(ms, e0, e1) != hipSuccess))
      rc = fail(ADZ_EHIP, "adz_plan_autotune: timing");
    if (rc == ADZ_OK) *ms /= reps;
    // a candidate whose staging waits gave up timed a fallback, not itself: drop it
    int aborted = 0;
    if (rc == ADZ_OK) rc = adz_plan_status(p, &aborted);
    if (rc == ADZ_OK && aborted) *ms = 3.0e38f;
    int rf = adz_plan_free(p);
    return rc != ADZ_OK ? rc : rf;
  };
  adz_tune cur = default_tune(fdtype, math);
  cur.fields_per_pass = std::min(cur.fields_per_pass, maxf);
  float cur_ms = 0.f;
  int rc = time_one(cur, &cur_ms);
  const int stages[] = {1, 2, 3, 4, 5, 6, 7, 8}, fpps[] = {2, 3, 4, 5, 8, 10, 16, 20, 32, 40, 64},
            bpcs[] = {1, 2, 4, 8, 16}, flags[] = {0, 1}, orders[] = {0, 1, 2, 3, 4, 6, 8};
  const int nst = fdtype != ADZ_F32 ? 2 : math == ADZ_MATH_F32 ? 8 : 7, nor = 7;
  // staging and tile order together first: the LDS-staged forms depend on the L2 reuse the
  // tile order gives, so one is not searched at the other's best value.  Each staging is
  // timed at 4 blocks per CU, blocks of 16 waves at 1 (one of them is resident per CU).
  for (int s = 0; s < nst && rc == ADZ_OK; s++)
    for (int v = 0; v < nor && rc == ADZ_OK; v++) {
      adz_tune t = cur;
      t.stage = stages[s];
      t.level_inner = orders[v];
      t.blocks_per_cu = plan_threads(t.stage, fdtype) >= 1024 ? 1 : 4;
      if (t.stage == cur.stage && t.level_inner == cur.level_inner && t.blocks_per_cu == cur.blocks_per_cu) continue;
      float ms = 0.f;
      rc = time_one(t, &ms);
      if (rc == ADZ_OK && ms < cur_ms) { cur = t; cur_ms = ms; }
    }
  // then coordinates: 2 fields per pass, 1 tile order again (they interact through the L2
  // working set), 3 blocks per CU, 4 loop order, 5 XCD chunking
  const int coords[] = {2, 1, 3, 4, 5};
  for (int ci = 0; ci < 5 && rc == ADZ_OK; ci++) {
    const int coord = coords[ci];
    const int* vals = coord == 1 ? orders : coord == 2 ? fpps : coord == 3 ? bpcs : flags;
    int nv = coord == 1 ? nor : coord == 2 ? 11 : coord == 3 ? 5 : 2;
    for (int v = 0; v < nv && rc == ADZ_OK; v++) {
      adz_tune t = cur;
      int* slot = coord == 1 ? &t.level_inner : coord == 2 ? &t.fields_per_pass
                : coord == 3 ? &t.blocks_per_cu : coord == 4 ? &t.points_outer : &t.xcd_chunks;
      if (*slot == vals[v]) continue;
      if (coord == 2 && v > 0 && vals[v] > maxf && vals[v - 1] >= maxf) continue;  // one candidate covers all fields
      *slot = coord == 2 ? std::min(vals[v], maxf) : vals[v];
      if (coord == 2 && t.fields_per_pass == cur.fields_per_pass) continue;
      float ms = 0.f;
      rc = time_one(t, &ms);
      if (rc == ADZ_OK && ms < cur_ms) { cur = t; cur_ms = ms; }
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc != ADZ_OK) return rc;
  *best = cur;
  if (best_ms) *best_ms = cur_ms;
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
