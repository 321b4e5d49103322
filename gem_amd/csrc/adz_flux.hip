// Aranami et al. (2015) boundary fluxes of the LAM Bermejo-Conde fixer on gfx950:
// adz_BC_LAM_Aranami_hlt.F90:18-236, with the masks of adz_BC_LAM_setup.F90:18-92 and the
// flux masses of mass_tr_PM.F90:101-127.
//
// For every fixer tracer the reference interpolates two masked copies at the Adz_pb
// positions (:52-155): mask_o*f (f on the core, 0 elsewhere) and mask_i*f (f on the tile
// outside the core).  It then zeroes FLUX_out on the core and FLUX_in outside it
// (:157-227).  FLUX_out is therefore nonzero only at targets outside the core whose stencil
// reaches into the core, and FLUX_in only at core targets whose stencil reaches outside it:
// a few rings of points along the core's edges.  Every other target interpolates zeros.
//
// The GPU form works on those rings only:
//   1. k_flux_mark: one lane per target: index step, class (inactive, FLUX_out, FLUX_in),
//      count per block;
//   2. k_flux_scan, k_flux_scan_chunks: exclusive scan of the block counts;
//   3. k_flux_emit: the active targets, in target order, into a list;
//   4. k_flux_eval: per (group of tracers, slice of the list): the masked tricubic
//      interpolation in the reference's arithmetic, optional store into fo / fi,
//      double-double flux masses;
//   5. k_flux_fin: per tracer, a fixed-order sum of the partials -> mass_fo, mass_fi.
// Steps 1-3 depend only on the positions and run once per call for all tracers.
//
// Values at inactive targets are zeros in both forms.  The reference's zero there is a sum
// of products with +0 or -0, so its sign may differ; the flux masses are unaffected (a sum
// that starts from +0 is unchanged by adding -0).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "adz_common.h"
#include "adz_dd.h"
#include "adz_eval.h"

namespace adz {

constexpr int kFxBlock = 256;
constexpr int kFxGroup = 4;  // tracers per block of k_flux_eval

struct FluxTr {
  const void* src;  // tracer, element (1,1,1) of the model layout
  void* fo;         // FLUX_out / FLUX_in arrays (same layout), or null
  void* fi;
};

struct FluxArgs {
  ZtabDev lv;                    // thermo-level descriptor over the tracers' layout
  const void* xyz;               // Adz_pb (3, i0b:inb, j0b:jnb, 1:nk)
  int i0b, j0b, nbx, nby, nk;    // target box
  int ni, nj;                    // the tile: 1..ni x 1..nj x 1..nk
  int ic0, ic1, jc0, jc1, kc0;   // the core: ic0..ic1 x jc0..jc1 x kc0..nk
  int nxa, nya;                  // array extents (row and plane pitch)
  long npts;
  int ntr, nslices, store;
  int autobar;                   // Schm_autobar_L: masses over level 1 of FLUX * mask (area mask)
  const FluxTr* tr;
  unsigned char* cls;            // per target: bit0 active, bit1 in the core, bit2 bad position
  int* cnt;                      // per mark block: count, then exclusive offset within its chunk
  int* chunk;                    // per 1024 mark blocks: total, then exclusive offset
  int* total;
  unsigned* list;                // active targets: index | core << 30 | bad << 31
  double* part;                  // [ntr][nslices][4]: FLUX_out hi lo, FLUX_in hi lo
  const void* air;               // air mass at time M, element (1,1,1)
  const double* mask;            // geomh_mask_8, element (1,1)
  double* mass;                  // [ntr][2]: mass_fo, mass_fi
};

__device__ __forceinline__ void flux_target(const FluxArgs& A, long p, int* i, int* j, int* k) {
  *i = A.i0b + (int)(p % A.nbx);
  const long r = p / A.nbx;
  *j = A.j0b + (int)(r % A.nby);
  *k = 1 + (int)(r / A.nby);
}

__device__ __forceinline__ bool in_core(const FluxArgs& A, int i, int j, int k) {
  return i >= A.ic0 && i <= A.ic1 && j >= A.jc0 && j <= A.jc1 && k >= A.kc0;
}

// Class of one target (adz_BC_LAM_setup.F90:50-76 masks, Aranami_hlt :157-227 zeroing).
template <class TP>
__global__ __launch_bounds__(kFxBlock) void k_flux_mark(FluxArgs A) {
  const long p = (long)blockIdx.x * kFxBlock + threadIdx.x;
  int c = 0;
  if (p < A.npts) {
    int i, j, k;
    flux_target(A, p, &i, &j, &k);
    const bool core = in_core(A, i, j, k);
    double px, py, pz;
    load_pos<TP>(A.xyz, p, &px, &py, &pz);
    Index x;
    if (!index_step(px, py, pz, A.lv, &x)) {
      c = 1 | 4;  // evaluated as NaN
    } else {
      const int ka = x.pz + 1, kb = x.pz + (x.zlinear ? 2 : 4);  // stencil levels (1-based)
      bool act;
      if (core) {  // FLUX_in: the stencil meets the tile outside the core
        const int bx0 = max(x.ix - 1, 1), bx1 = min(x.ix + 2, A.ni);
        const int by0 = max(x.iy - 1, 1), by1 = min(x.iy + 2, A.nj);
        act = bx0 <= bx1 && by0 <= by1 &&
              !(bx0 >= A.ic0 && bx1 <= A.ic1 && by0 >= A.jc0 && by1 <= A.jc1 && ka >= A.kc0);
      } else {     // FLUX_out: the stencil meets the core
        act = x.ix + 2 >= A.ic0 && x.ix - 1 <= A.ic1 && x.iy + 2 >= A.jc0 && x.iy - 1 <= A.jc1 && kb >= A.kc0;
      }
      c = act ? 1 : 0;
    }
    if (core) c |= 2;
    A.cls[p] = (unsigned char)c;
  }
  const int n = __syncthreads_count(c & 1);
  if (threadIdx.x == 0) A.cnt[blockIdx.x] = n;
}

// Exclusive scan of the n block counts in two levels, with coalesced loads: each block of
// 1024 lanes scans 1024 consecutive counts in place and writes its total to chunk[b];
// k_flux_scan_chunks then scans the chunk totals (one block; chunk[nchunks] = the total).
constexpr int kScan = 1024;
__device__ __forceinline__ int block_scan_excl(int v, int* s) {  // s: kScan ints of LDS
  const int t = threadIdx.x;
  s[t] = v;
  __syncthreads();
  for (int o = 1; o < kScan; o <<= 1) {
    const int u = t >= o ? s[t - o] : 0;
    __syncthreads();
    s[t] += u;
    __syncthreads();
  }
  return s[t] - v;
}

__global__ __launch_bounds__(kScan) void k_flux_scan(int* cnt, int n, int* chunk) {
  __shared__ int s[kScan];
  const long b = (long)blockIdx.x * kScan + threadIdx.x;
  const int v = b < n ? cnt[b] : 0;
  const int e = block_scan_excl(v, s);
  if (b < n) cnt[b] = e;
  if (threadIdx.x == kScan - 1) chunk[blockIdx.x] = e + v;
}

__global__ __launch_bounds__(kScan) void k_flux_scan_chunks(int* chunk, int nchunks, int* total) {
  __shared__ int s[kScan];
  int run = 0;
  for (int c0 = 0; c0 < nchunks; c0 += kScan) {  // nchunks <= 2^30 / 256 / 1024: a few passes at most
    const int c = c0 + (int)threadIdx.x;
    const int v = c < nchunks ? chunk[c] : 0;
    const int e = block_scan_excl(v, s);
    if (c < nchunks) chunk[c] = run + e;
    run += s[kScan - 1];
    __syncthreads();
  }
  if (threadIdx.x == 0) *total = run;
}

__global__ __launch_bounds__(kFxBlock) void k_flux_emit(FluxArgs A) {
  __shared__ int wsum[kFxBlock / 64];
  const long p = (long)blockIdx.x * kFxBlock + threadIdx.x;
  const int c = p < A.npts ? A.cls[p] : 0;
  const bool act = (c & 1) != 0;
  const unsigned long long m = __ballot(act);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int below = __popcll(m & ((1ull << lane) - 1ull));
  if (lane == 0) wsum[w] = __popcll(m);
  __syncthreads();
  int base = A.cnt[blockIdx.x] + A.chunk[blockIdx.x / kScan];
  for (int q = 0; q < w; q++) base += wsum[q];
  if (act) A.list[base + below] = (unsigned)p | ((unsigned)(c & 6) << 29);
}

// Masked 4x4x4 (or 4x4x2) stencil: element (r, z, i) is f times 1 where the mask is 1 and
// times 0 elsewhere, as the reference's `real` product mask*f (Aranami_hlt :72-73).
template <class TF>
struct MaskedStencil {
  const ADZ_GLOBAL TF* f1;
  long ni, o1, o2, o3;
  unsigned colc, cold, rowc, rowd, levc;  // 4-bit masks: in the core / in the tile
  bool outside;                           // FLUX_in (mask_i) rather than FLUX_out (mask_o)
  __device__ __forceinline__ TF operator()(int r, int z, int i) const {
    const long zo = z == 0 ? 0 : (z == 1 ? o1 : (z == 2 ? o2 : o3));
    const bool c = ((colc >> i) & (rowc >> r) & (levc >> z) & 1u) != 0;
    const bool d = ((cold >> i) & (rowd >> r) & 1u) != 0;
    const bool on = outside ? (d && !c) : c;
    return f1[r * ni + zo + i] * (on ? (TF)1 : (TF)0);
  }
};

// One block: a slice of the active list for up to kFxGroup tracers; the index step, the
// weights and the masks of a target are computed once for the group.
template <class TF, class TP>
__global__ __launch_bounds__(kFxBlock) void k_flux_eval(FluxArgs A) {
  const int t0 = blockIdx.x * kFxGroup, nt = min(kFxGroup, A.ntr - t0);
  const int total = *A.total;
  const ADZ_GLOBAL TF* air = (const ADZ_GLOBAL TF*)A.air;
  const ADZ_GLOBAL double* maskg = (const ADZ_GLOBAL double*)A.mask;
  // per-lane double-double accumulators in LDS (each lane owns a column): [4g + 0/1]
  // FLUX_out hi/lo, [4g + 2/3] FLUX_in hi/lo of the group's tracer g
  __shared__ double acc[4 * kFxGroup][kFxBlock];
#pragma unroll
  for (int q = 0; q < 4 * kFxGroup; q++) acc[q][threadIdx.x] = 0.0;
  for (long e = (long)blockIdx.y * kFxBlock + threadIdx.x; e < total; e += (long)gridDim.y * kFxBlock) {
    const unsigned code = A.list[e];
    const long p = code & 0x3fffffffu;
    const bool core = ((code >> 30) & 1u) != 0, bad = (code >> 31) != 0;
    int i, j, k;
    flux_target(A, p, &i, &j, &k);
    const long q3 = (long)(i - 1) + (long)A.nxa * ((j - 1) + (long)A.nya * (k - 1));
    const long q2 = (long)(i - 1) + (long)A.nxa * (j - 1);
    const TF airv = air[q3];
    const double mk = maskg[q2];
    Index x;
    double c[20];
    MaskedStencil<TF> at;
    if (!bad) {
      double px, py, pz;
      load_pos<TP>(A.xyz, p, &px, &py, &pz);
      index_step(px, py, pz, A.lv, &x);
      weights_ref(x, A.lv, c);
      at.ni = A.lv.ni;
      at.o1 = x.zlinear ? 0 : A.lv.nij;
      at.o2 = at.o1 + A.lv.nij;
      at.o3 = at.o2 + at.o1;
      at.colc = at.cold = at.rowc = at.rowd = at.levc = 0u;
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const int col = x.ix - 1 + q, row = x.iy - 1 + q;
        const int lev = x.pz + 1 + (x.zlinear ? (q >> 1) : q);
        at.colc |= (unsigned)(col >= A.ic0 && col <= A.ic1) << q;
        at.cold |= (unsigned)(col >= 1 && col <= A.ni) << q;
        at.rowc |= (unsigned)(row >= A.jc0 && row <= A.jc1) << q;
        at.rowd |= (unsigned)(row >= 1 && row <= A.nj) << q;
        at.levc |= (unsigned)(lev >= A.kc0) << q;
      }
      at.outside = core;
    }
#pragma unroll 1
    for (int g = 0; g < nt; g++) {
      const FluxTr tr = A.tr[t0 + g];
      double v = __builtin_nan("");
      if (!bad) {
        MaskedStencil<TF> st = at;
        st.f1 = (const ADZ_GLOBAL TF*)tr.src + x.off;
        // the 64 mask factors are cheap to form again per tracer; an opaque copy of the
        // mask bits keeps the compiler from holding all of them in registers across tracers
        asm volatile("" : "+v"(st.colc), "+v"(st.cold), "+v"(st.rowc), "+v"(st.rowd), "+v"(st.levc));
        v = tricub_ref_a<TF>(st, c);
      }
      const TF out = (TF)v;
      // mass_tr_PM.F90:105-125: FLUX * air mass * mask, or on level 1 FLUX * area mask
      const double term = A.autobar ? (k == 1 ? (double)out * mk : 0.0) : mass_term<TF>(out, airv, mk);
      const int a = 4 * g + (core ? 2 : 0);
      DD d{acc[a][threadIdx.x], acc[a + 1][threadIdx.x]};
      dd_add(&d, term);
      acc[a][threadIdx.x] = d.hi;
      acc[a + 1][threadIdx.x] = d.lo;
      if (A.store) ((ADZ_GLOBAL TF*)(core ? tr.fi : tr.fo))[q3] = out;
    }
  }
  // fixed-order block tree over the lanes' accumulators
  __syncthreads();
  for (int w = kFxBlock / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
      for (int q = 0; q < 2 * nt; q++) {
        DD a{acc[2 * q][threadIdx.x], acc[2 * q + 1][threadIdx.x]};
        DD b{acc[2 * q][threadIdx.x + w], acc[2 * q + 1][threadIdx.x + w]};
        DD r = dd_merge(a, b);
        acc[2 * q][threadIdx.x] = r.hi; acc[2 * q + 1][threadIdx.x] = r.lo;
      }
    }
    __syncthreads();
  }
  if (threadIdx.x < 4 * nt) {
    const int g = threadIdx.x / 4, q = threadIdx.x % 4;
    A.part[((long)(t0 + g) * A.nslices + blockIdx.y) * 4 + q] = acc[threadIdx.x][0];
  }
}

// Per tracer: lane l merges slices l, l+64, ... in order, then a fixed tree over lanes.
__global__ __launch_bounds__(64) void k_flux_fin(FluxArgs A) {
  const int t = blockIdx.x, l = threadIdx.x;
  DD so{0.0, 0.0}, si{0.0, 0.0};
  const double* pp = A.part + (long)t * A.nslices * 4;
  for (int s = l; s < A.nslices; s += 64) {
    so = dd_merge(so, DD{pp[4 * s], pp[4 * s + 1]});
    si = dd_merge(si, DD{pp[4 * s + 2], pp[4 * s + 3]});
  }
  __shared__ double sh[4][64];
  sh[0][l] = so.hi; sh[1][l] = so.lo; sh[2][l] = si.hi; sh[3][l] = si.lo;
  __syncthreads();
  for (int w = 32; w > 0; w >>= 1) {
    if (l < w) {
      DD a = dd_merge(DD{sh[0][l], sh[1][l]}, DD{sh[0][l + w], sh[1][l + w]});
      DD b = dd_merge(DD{sh[2][l], sh[3][l]}, DD{sh[2][l + w], sh[3][l + w]});
      sh[0][l] = a.hi; sh[1][l] = a.lo; sh[2][l] = b.hi; sh[3][l] = b.lo;
    }
    __syncthreads();
  }
  if (l == 0) {
    A.mass[2 * t] = sh[0][0] + sh[1][0] + 0.0;      // a sum that starts from +0, as
    A.mass[2 * t + 1] = sh[2][0] + sh[3][0] + 0.0;  // mass_tr_PM's accumulators do
  }
}

// The reference's zeroing leaves zeros at every target it does not overwrite; with stored
// arrays, every target of both arrays starts at zero before the active ones are written.
template <class TF>
__global__ __launch_bounds__(kFxBlock) void k_flux_zero(FluxArgs A) {
  const long p = (long)blockIdx.x * kFxBlock + threadIdx.x;
  if (p >= A.npts) return;
  const FluxTr tr = A.tr[blockIdx.y];
  int i, j, k;
  flux_target(A, p, &i, &j, &k);
  const long q3 = (long)(i - 1) + (long)A.nxa * ((j - 1) + (long)A.nya * (k - 1));
  ((ADZ_GLOBAL TF*)tr.fo)[q3] = (TF)0;
  ((ADZ_GLOBAL TF*)tr.fi)[q3] = (TF)0;
}

}  // namespace adz

using namespace adz;

struct adz_flux {
  int tdtype, pdtype;
  FluxArgs a;
  int nmark;
  void* mem;
};

extern "C" {

int adz_flux_create(const adz_flux_tracer* tr, int ntr, const adz_flux_domain* dom, const adz_ztab* geom, int tdtype,
                    int pdtype, adz_flux** out) {
  if (!tr || ntr <= 0 || !dom || !geom || !out) return fail(ADZ_EINVAL, "adz_flux_create: bad argument");
  if (tdtype != ADZ_F32 && tdtype != ADZ_F64) return fail(ADZ_EINVAL, "adz_flux_create: tdtype must be F32 or F64");
  if (pdtype != ADZ_F32 && pdtype != ADZ_F64) return fail(ADZ_EINVAL, "adz_flux_create: pdtype must be F32 or F64");
  const adz_flux_domain& d = *dom;
  if (d.inb < d.i0b || d.jnb < d.j0b || d.nk < 4 || d.i0b < d.minx || d.inb > d.maxx || d.j0b < d.miny ||
      d.jnb > d.maxy || d.ni < 1 || d.nj < 1 || d.ni > d.maxx || d.nj > d.maxy || d.minx > 1 || d.miny > 1)
    return fail(ADZ_EINVAL, "adz_flux_create: target box or tile outside the arrays");
  if (d.nk != geom->dev.nk) return fail(ADZ_EINVAL, "adz_flux_create: nk differs from the descriptor's");
  const int store = tr[0].fo != nullptr;
  for (int n = 0; n < ntr; n++) {
    if (!tr[n].src) return fail(ADZ_EINVAL, "adz_flux_create: null tracer");
    if ((tr[n].fo != nullptr) != store || (tr[n].fi != nullptr) != store)
      return fail(ADZ_EINVAL, "adz_flux_create: fo/fi must be given for every tracer or for none");
  }
  FluxArgs a;
  std::memset(&a, 0, sizeof(a));
  a.lv = geom->dev;  // the tracers have the model layout (as adz_tricub_hlt with F_ext_L = .false.)
  a.lv.ni = d.maxx - d.minx + 1;
  a.lv.nj = d.maxy - d.miny + 1;
  a.lv.nij = a.lv.ni * a.lv.nj;
  a.lv.ilo = d.minx; a.lv.ihi = d.maxx; a.lv.jlo = d.miny; a.lv.jhi = d.maxy;
  a.i0b = d.i0b; a.j0b = d.j0b; a.nbx = d.inb - d.i0b + 1; a.nby = d.jnb - d.j0b + 1; a.nk = d.nk;
  a.ni = d.ni; a.nj = d.nj;
  a.ic0 = d.ic0; a.ic1 = d.ic1; a.jc0 = d.jc0; a.jc1 = d.jc1; a.kc0 = d.kc0;
  a.nxa = a.lv.ni; a.nya = a.lv.nj;
  a.npts = (long)a.nbx * a.nby * d.nk;
  if (a.npts >= (1L << 30)) return fail(ADZ_EINVAL, "adz_flux_create: more than 2^30 targets");
  a.ntr = ntr;
  a.store = store;
  a.autobar = d.autobar != 0;
  const int ngroups = (ntr + kFxGroup - 1) / kFxGroup;
  a.nslices = std::max(16, std::min(1024, 2048 / ngroups));
  const int nmark = (int)((a.npts + kFxBlock - 1) / kFxBlock);
  std::vector<FluxTr> h(ntr);
  for (int n = 0; n < ntr; n++) h[n] = FluxTr{tr[n].src, tr[n].fo, tr[n].fi};
  auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
  const int nchunk = (nmark + kScan - 1) / kScan;
  const size_t tb = up(sizeof(FluxTr) * ntr), pb = up(sizeof(double) * 4 * ntr * a.nslices),
               cb = up(sizeof(int) * (nmark + 1 + nchunk)), lb = up(sizeof(unsigned) * a.npts),
               kb = up((size_t)a.npts);
  char* mem = nullptr;
  if (hipMalloc((void**)&mem, tb + pb + cb + lb + kb) != hipSuccess) return fail(ADZ_ENOMEM, "adz_flux_create: hipMalloc");
  if (hipMemcpy(mem, h.data(), sizeof(FluxTr) * ntr, hipMemcpyHostToDevice) != hipSuccess) {
    (void)hipFree(mem);
    return fail(ADZ_EHIP, "adz_flux_create: upload");
  }
  a.tr = (const FluxTr*)mem;
  a.part = (double*)(mem + tb);
  a.cnt = (int*)(mem + tb + pb);
  a.total = a.cnt + nmark;
  a.chunk = a.total + 1;
  a.list = (unsigned*)(mem + tb + pb + cb);
  a.cls = (unsigned char*)(mem + tb + pb + cb + lb);
  adz_flux* f = new adz_flux;
  f->tdtype = tdtype; f->pdtype = pdtype; f->a = a; f->nmark = nmark; f->mem = mem;
  *out = f;
  return ADZ_OK;
}

int adz_bc_lam_flux(adz_flux* fx, const void* xyz, const void* airm0, const double* mask, double* mass,
                    adz_stream stream) {
  if (!fx || !xyz || !airm0 || !mask || !mass) return fail(ADZ_EINVAL, "adz_bc_lam_flux: null argument");
  FluxArgs a = fx->a;
  a.xyz = xyz; a.air = airm0; a.mask = mask; a.mass = mass;
  hipStream_t st = (hipStream_t)stream;
  const dim3 gm((unsigned)fx->nmark), b(kFxBlock);
  if (fx->pdtype == ADZ_F32) hipLaunchKernelGGL((k_flux_mark<float>), gm, b, 0, st, a);
  else hipLaunchKernelGGL((k_flux_mark<double>), gm, b, 0, st, a);
  const int nchunk = (fx->nmark + kScan - 1) / kScan;
  hipLaunchKernelGGL(k_flux_scan, dim3((unsigned)nchunk), dim3(kScan), 0, st, a.cnt, fx->nmark, a.chunk);
  hipLaunchKernelGGL(k_flux_scan_chunks, dim3(1), dim3(kScan), 0, st, a.chunk, nchunk, a.total);
  hipLaunchKernelGGL(k_flux_emit, gm, b, 0, st, a);
  if (a.store) {
    const dim3 gz((unsigned)fx->nmark, (unsigned)a.ntr);
    if (fx->tdtype == ADZ_F32) hipLaunchKernelGGL((k_flux_zero<float>), gz, b, 0, st, a);
    else hipLaunchKernelGGL((k_flux_zero<double>), gz, b, 0, st, a);
  }
  const dim3 ge((unsigned)((a.ntr + kFxGroup - 1) / kFxGroup), (unsigned)a.nslices);
  if (fx->tdtype == ADZ_F32) {
    if (fx->pdtype == ADZ_F32) hipLaunchKernelGGL((k_flux_eval<float, float>), ge, b, 0, st, a);
    else hipLaunchKernelGGL((k_flux_eval<float, double>), ge, b, 0, st, a);
  } else {
    if (fx->pdtype == ADZ_F32) hipLaunchKernelGGL((k_flux_eval<double, float>), ge, b, 0, st, a);
    else hipLaunchKernelGGL((k_flux_eval<double, double>), ge, b, 0, st, a);
  }
  hipLaunchKernelGGL(k_flux_fin, dim3((unsigned)a.ntr), dim3(64), 0, st, a);
  ADZ_HIP_CHECK(hipGetLastError());
  return ADZ_OK;
}

int adz_flux_active(const adz_flux* fx, int* count) {
  if (!fx || !count) return fail(ADZ_EINVAL, "adz_flux_active: null argument");
  ADZ_HIP_CHECK(hipMemcpy(count, fx->a.total, sizeof(int), hipMemcpyDeviceToHost));
  return ADZ_OK;
}

int adz_flux_free(adz_flux* fx) {
  if (!fx) return ADZ_OK;
  hipError_t e = hipFree(fx->mem);
  delete fx;
  if (e != hipSuccess) return fail(ADZ_EHIP, "adz_flux_free: hipFree");
  return ADZ_OK;
}

}  // extern "C"

// ====================================================================== ZLF
// The ZLF form of the fixer's LAM boundary treatment (Adz_BC_LAM_flux = 2):
// adz_BC_LAM_zlf_hlt.F90:18-319, four element-wise steps on model-layout arrays of a tile
// whose four sides are LAM edges.  Every element of the tile 1..ni x 1..nj x 1..nk is
// visited once; the halo is not touched (the reference's loops cover the tile only).
//   setup 1 (F1 = source): zero outside the extension E = [1+pil_w-ext, ni-pil_e+ext] x
//           [1+pil_s-ext, nj-pil_n+ext], and inside E below level nint(k0/2);
//   setup 2 (F1 = destination, F2 = piloting copy): F2 = F1 outside the core
//           C = [1+pil_w, ni-pil_e] x [1+pil_s, nj-pil_n] x [k0, nk];
//   setup 3 (F1 = destination): zero outside [1+4, ni-4] x [1+4, nj-4] (BCS_BASE = 4);
//   setup 4 (F1 = destination, F2 = piloting copy): F1 = F2 outside C.
namespace adz {

constexpr int kZlfFields = 16;

struct ZlfArgs {
  int setup, ni, nj, nk, nxa, nya;
  int i0e, ine, j0e, jne, k0d2;   // setup 1
  int i0c, inc, j0c, jnc, k0;     // setups 2 and 4
  int i0b, inb, j0b, jnb;         // setup 3
  int nf;
  void* f1[kZlfFields];
  void* f2[kZlfFields];
};

template <class T>
__global__ __launch_bounds__(256) void k_zlf(ZlfArgs A) {
  const long p = (long)blockIdx.x * 256 + threadIdx.x;
  const long n = (long)A.ni * A.nj * A.nk;
  if (p >= n) return;
  const int i = 1 + (int)(p % A.ni);
  const long r = p / A.ni;
  const int j = 1 + (int)(r % A.nj), k = 1 + (int)(r / A.nj);
  const long q = (long)(i - 1) + (long)A.nxa * ((j - 1) + (long)A.nya * (k - 1));
  bool act;
  if (A.setup == 1) {
    const bool inE = i >= A.i0e && i <= A.ine && j >= A.j0e && j <= A.jne;
    act = !inE || k < A.k0d2;
  } else if (A.setup == 3) {
    act = !(i >= A.i0b && i <= A.inb && j >= A.j0b && j <= A.jnb);
  } else {
    act = !(i >= A.i0c && i <= A.inc && j >= A.j0c && j <= A.jnc && k >= A.k0);
  }
  if (!act) return;
  for (int f = 0; f < A.nf; f++) {
    ADZ_GLOBAL T* a = (ADZ_GLOBAL T*)A.f1[f];
    ADZ_GLOBAL T* b = (ADZ_GLOBAL T*)A.f2[f];
    if (A.setup == 1 || A.setup == 3) a[q] = (T)0;
    else if (A.setup == 2) b[q] = a[q];
    else a[q] = b[q];
  }
}

}  // namespace adz

extern "C" int adz_bc_lam_zlf(int setup, void* const* f1, void* const* f2, int nf, const adz_zlf_domain* dom,
                              int dtype, adz_stream stream) {
  if (setup < 1 || setup > 4) return fail(ADZ_EINVAL, "ADZ_BC_LAM_ZLF: SETUP not defined");
  if (!f1 || !dom || nf < 0) return fail(ADZ_EINVAL, "adz_bc_lam_zlf: bad argument");
  if ((setup == 2 || setup == 4) && !f2) return fail(ADZ_EINVAL, "adz_bc_lam_zlf: setups 2 and 4 need F2");
  if (dtype != ADZ_F32 && dtype != ADZ_F64) return fail(ADZ_EINVAL, "adz_bc_lam_zlf: dtype must be F32 or F64");
  const adz_zlf_domain& d = *dom;
  if (d.ni < 1 || d.nj < 1 || d.nk < 1 || d.minx > 1 || d.miny > 1 || d.maxx < d.ni || d.maxy < d.nj)
    return fail(ADZ_EINVAL, "adz_bc_lam_zlf: tile outside the arrays");
  ZlfArgs a;
  std::memset(&a, 0, sizeof(a));
  a.setup = setup; a.ni = d.ni; a.nj = d.nj; a.nk = d.nk;
  a.nxa = d.maxx - d.minx + 1; a.nya = d.maxy - d.miny + 1;
  a.i0e = 1 + d.pil_w - d.ext; a.ine = d.ni - d.pil_e + d.ext;
  a.j0e = 1 + d.pil_s - d.ext; a.jne = d.nj - d.pil_n + d.ext;
  a.k0d2 = (int)std::lround(d.k0 / 2.0);  // nint(k0/2.): halves round away from zero
  a.i0c = 1 + d.pil_w; a.inc = d.ni - d.pil_e; a.j0c = 1 + d.pil_s; a.jnc = d.nj - d.pil_n; a.k0 = d.k0;
  a.i0b = 1 + 4; a.inb = d.ni - 4; a.j0b = 1 + 4; a.jnb = d.nj - 4;
  const long n = (long)d.ni * d.nj * d.nk;
  for (int f0 = 0; f0 < nf; f0 += kZlfFields) {
    a.nf = std::min(kZlfFields, nf - f0);
    for (int f = 0; f < a.nf; f++) {
      a.f1[f] = f1[f0 + f];
      a.f2[f] = f2 ? f2[f0 + f] : nullptr;
      if (!a.f1[f] || ((setup == 2 || setup == 4) && !a.f2[f]))
        return fail(ADZ_EINVAL, "adz_bc_lam_zlf: null array");
    }
    const dim3 g((unsigned)((n + 255) / 256)), b(256);
    if (dtype == ADZ_F32) hipLaunchKernelGGL((k_zlf<float>), g, b, 0, (hipStream_t)stream, a);
    else hipLaunchKernelGGL((k_zlf<double>), g, b, 0, (hipStream_t)stream, a);
  }
  ADZ_HIP_CHECK(hipGetLastError());
  return ADZ_OK;
}
