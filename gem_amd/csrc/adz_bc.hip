// Bermejo-Conde mass fixer on gfx950 (src/gemdyn/src/base/Bermejo_Conde.F90:18-320, with
// the global FP64 sums of mass_tr_PM.F90:75-175 and mass_tr_X.F90:66-111).
//
// The fixer needs four global sums per tracer, each one gating the next elementwise step:
//   1. masses of p (time P) and m (time M)        -> deficit D = M_m - (M_p - eps)
//   2. mass of the weights w(m - lin, sign D)     -> lambda = D / W (skip if W is ~0)
//   3. corrected m (min/max reset) and its mass  -> proportion = M_target / M_corrected
//   4. proportional fix of m and its mass        -> new carried deficit eps
// Sums 1 and 2 share one pass: a weight is nonzero only where sign(m - lin) = sign(D), so
// the pass accumulates the weight mass of both signs and the finishing step keeps the one
// sign(D) selects (negation commutes with every rounding involved, so W is the same value
// a separate pass would give).  Every pass runs all tracers in one launch.  Sums are deterministic:
// each lane accumulates a fixed set of elements in double-double (TwoSum), blocks reduce
// in a fixed tree, and a finalising kernel adds the block partials in index order; the
// result is the exactly rounded sum except in astronomically rare ties, and independent of
// how the domain is split.  The weights are recomputed in passes 2 and 3 instead of being
// stored (w depends only on m, lin, pkps and the sign of D), which saves one array per
// tracer.  Scalars stay on the device, so a timestep can be captured in a hipGraph.
#include <algorithm>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "adz_common.h"
#include "adz_dd.h"

namespace adz {

#ifndef ADZ_BC_UNROLL
#define ADZ_BC_UNROLL 4  // elements gathered per lane before accumulating
#endif
constexpr int kBcBlock = 256;
constexpr int kSc = 16;  // per-tracer scalars
enum { SC_MP = 0, SC_MM, SC_MTP, SC_D, SC_W, SC_LAMBDA, SC_CYCLE, SC_MBCR, SC_PROP, SC_MBC };
enum { P_PM = 0, P_W = 1, P_CORR = 2, P_PROP = 3 };

struct BcTr {
  const void* p;
  void* m;
  const void* lin;
  const void* mn;
  const void* mx;
  int legacy, weight;
  double pexp;
};

struct BcArgs {
  const BcTr* tr;
  int ntr;
  int nxa, nya;              // array extents
  int i0, j0, k0, nx, ny;    // scope
  long npts;
  int min_max;
  int autobar;               // Schm_autobar_L: masses over level 1 of x * mask (area mask)
  int nblk;                  // blocks per tracer
  int nrows;                 // scope rows (j fastest, then k)
  const void* airm1;
  const void* airm0;
  const void* pkps;
  const double* mask;
  double* deficit;
  double* diag;
  const double* flux;        // [ntr][2] masses of FLUX_out, FLUX_in (LAM Aranami), or null
  double* part;              // [ntr][nblk][8]
  double* sc;                // [ntr][kSc]
};

// almost_zero.F90:16-36 (see the oracle's note: true for the patterns 0 and 1 only)
__device__ __forceinline__ bool d_almost_zero(double a) {
  unsigned long long u = __builtin_bit_cast(unsigned long long, a);
  long long abit = (long long)u;
  if (abit < 0) abit = (long long)(0x80000000ULL - u);
  return abit >= -1 && abit <= 1;
}

__device__ __forceinline__ double sgn1(double b) { return signbit(b) ? -1.0 : 1.0; }

// w of Bermejo_Conde.F90:112-172 for one element (exactly the stored `real` weight)
template <class T>
__device__ __forceinline__ T bc_weight(T m, T lin, double pkps, const BcTr& t, double sd) {
  T hml = m - lin;
  double H = (double)hml;
  T w = (T)0;
  if (sd == sgn1(H)) {
    double a = fabs(H);
    double v = (t.pexp == 1.0) ? a : (t.pexp == 2.0 ? a * a : pow(a, t.pexp));
    v = v * sd;
    if (t.weight == 2) v = v * (double)(m > (T)0 ? m : (T)0);
    if (t.weight == 3) v = v * pkps;
    w = (T)v;
  }
  return w;
}

// one element's contribution to a mass: tracer * air mass * geomh_mask_8 (mass_tr_X.F90:86-93),
// or, with Schm_autobar_L, tracer * geomh_area_mask_8 on level 1 only (:73-79)
template <class T>
__device__ __forceinline__ double bc_term(int autobar, int k, T x, T air, double mask) {
  if (autobar) return k == 1 ? (double)x * mask : 0.0;
  return mass_term<T>(x, air, mask);
}

// |w| and the sign of m - lin for the fused first pass (w = sign * |w| when the signs match)
template <class T>
__device__ __forceinline__ T bc_weight_abs(T m, T lin, double pkps, const BcTr& t, bool* pos) {
  T hml = m - lin;
  double H = (double)hml;
  *pos = !signbit(H);
  double a = fabs(H);
  double v = (t.pexp == 1.0) ? a : (t.pexp == 2.0 ? a * a : pow(a, t.pexp));
  if (t.weight == 2) v = v * (double)(m > (T)0 ? m : (T)0);
  if (t.weight == 3) v = v * pkps;
  return (T)v;
}

// Each block owns whole rows (a fixed partition) for kBcWaves tracers, one wave per tracer,
// lanes striding along i.  The waves of a block walk the same rows together, so the air
// masses, mask and pkps they share are fetched once for the block and served to the other
// waves from the CU's caches.
constexpr int kBcWaves = kBcBlock / 64;
template <class T, int PASS>
__global__ __launch_bounds__(kBcBlock) void k_bc_pass(BcArgs A) {
  // 1-D grid of quads*nblk blocks (a multiple of 8): blocks are dispatched round-robin over
  // the 8 XCDs, so XCD x takes the x-th eighth of the (row block, tracer group) list,
  // tracer group fastest: the blocks over the same rows share one XCD's L2
  const long quads = (A.ntr + kBcWaves - 1) / kBcWaves;
  const long nlog = quads * A.nblk, per = nlog / 8;
  const long lg = (long)(blockIdx.x % 8) * per + blockIdx.x / 8;
  if (lg >= nlog) return;
  const int lane = threadIdx.x & 63;
  const int n = (int)(lg % quads) * kBcWaves + (int)(threadIdx.x / 64);
  const int by = (int)(lg / quads);
  if (n >= A.ntr) return;  // a wave without a tracer (the last group): nothing to add
  const BcTr t = A.tr[n];
  const double* sc = A.sc + n * kSc;
  const bool cyc = (PASS >= P_CORR) && sc[SC_CYCLE] != 0.0;
  const double sd = sgn1(sc[SC_D]);
  const double lambda = sc[SC_LAMBDA], prop = sc[SC_PROP];
  DD s0{0.0, 0.0}, s1{0.0, 0.0}, s2{0.0, 0.0}, s3{0.0, 0.0};
  const int r0 = (int)((long)A.nrows * by / A.nblk), r1 = (int)((long)A.nrows * (by + 1) / A.nblk);
  const ADZ_GLOBAL T* air1 = (const ADZ_GLOBAL T*)A.airm1;
  const ADZ_GLOBAL T* air0 = (const ADZ_GLOBAL T*)A.airm0;
  const ADZ_GLOBAL double* maskg = (const ADZ_GLOBAL double*)A.mask;
  ADZ_GLOBAL T* mg = (ADZ_GLOBAL T*)t.m;
  // Each lane gathers the inputs of kU elements (kU * 64 apart along the row) before it
  // accumulates them in order, so a wave keeps kU loads per array in flight.
  constexpr int kU = ADZ_BC_UNROLL;
  const ADZ_GLOBAL T* lin = (const ADZ_GLOBAL T*)t.lin;
  const ADZ_GLOBAL T* pk = (const ADZ_GLOBAL T*)A.pkps;
  const ADZ_GLOBAL T* pg = (const ADZ_GLOBAL T*)t.p;
  const ADZ_GLOBAL T* mng = (const ADZ_GLOBAL T*)t.mn;
  const ADZ_GLOBAL T* mxg = (const ADZ_GLOBAL T*)t.mx;
  for (int r = r0; r < r1; r++) {
    const int j = A.j0 + r % A.ny, k = A.k0 + r / A.ny;
    const long q2r = (long)(A.i0 - 1) + (long)A.nxa * (j - 1);
    const long qr = q2r + (long)A.nxa * A.nya * (k - 1);
    for (int ib = 0; ib < A.nx; ib += kU * 64) {
      T mv[kU], av[kU], bv[kU], lv[kU], nv[kU], xv[kU];
      double mk[kU], pkv[kU];
#pragma unroll
      for (int u = 0; u < kU; u++) {
        const int ii = ib + u * 64 + lane;
        const bool in = ii < A.nx;
        const long q = qr + (in ? ii : 0), q2 = q2r + (in ? ii : 0);
        mk[u] = in ? maskg[q2] : 0.0;
        mv[u] = mg[q];
        av[u] = air0[q];
        if (PASS == P_PM) { bv[u] = air1[q]; nv[u] = pg[q]; }
        if (PASS == P_PM || PASS == P_W || (PASS == P_CORR && !cyc)) {
          lv[u] = lin[q];
          pkv[u] = (t.weight == 3) ? (double)pk[q] : 0.0;
        }
        // the correction lambda*w has the sign of D wherever it is nonzero (w and W both
        // carry sign(D)), so only the bound on that side can trigger the reset
        if (PASS == P_CORR && !cyc && A.min_max) {
          if (sd > 0.0) { nv[u] = mng[q]; xv[u] = (T)0; }
          else { xv[u] = mxg[q]; nv[u] = (T)0; }
        }
      }
#pragma unroll
      for (int u = 0; u < kU; u++) {
        const int ii = ib + u * 64 + lane;
        if (ii >= A.nx) break;
        ADZ_GLOBAL T* mp = mg + qr + ii;
        T m = mv[u];
        if (PASS == P_PM) {
          dd_add(&s0, bc_term<T>(A.autobar, k, nv[u], bv[u], mk[u]));
          dd_add(&s1, bc_term<T>(A.autobar, k, m, av[u], mk[u]));
          bool pos;
          const double wt = bc_term<T>(A.autobar, k, bc_weight_abs<T>(m, lv[u], pkv[u], t, &pos), av[u], mk[u]);
          dd_add(pos ? &s2 : &s3, wt);
        } else if (PASS == P_W) {
          T w = bc_weight<T>(m, lv[u], pkv[u], t, sd);
          dd_add(&s0, bc_term<T>(A.autobar, k, w, av[u], mk[u]));
        } else if (PASS == P_CORR) {
          if (!cyc) {
            T w = bc_weight<T>(m, lv[u], pkv[u], t, sd);
            double corr = lambda * (double)w;
            T bc = (T)((double)m - corr);
            if (A.min_max) {
              if (sd > 0.0) { if (corr > 0.0 && bc < nv[u]) bc = nv[u]; }
              else if (corr < 0.0 && bc > xv[u]) bc = xv[u];
            }
            m = bc;
            *mp = m;
          }
          dd_add(&s0, bc_term<T>(A.autobar, k, m, av[u], mk[u]));
        } else {  // P_PROP
          if (!cyc) {
            m = (T)(prop * (double)m);
            *mp = m;
          }
          dd_add(&s0, bc_term<T>(A.autobar, k, m, av[u], mk[u]));
        }
      }
    }
  }
  // fixed-order wave tree (P_PM carries four sums, the other passes one)
  constexpr int NS = PASS == P_PM ? 4 : 1;
  DD acc[4] = {s0, s1, s2, s3};
#pragma unroll
  for (int q = 0; q < NS; q++)
    for (int off = 32; off > 0; off >>= 1) {
      const DD o{__shfl_down(acc[q].hi, off, 64), __shfl_down(acc[q].lo, off, 64)};
      acc[q] = dd_merge(acc[q], o);
    }
  if (lane == 0) {
    double* o = A.part + ((long)n * A.nblk + by) * 8;
#pragma unroll
    for (int q = 0; q < NS; q++) { o[2 * q] = acc[q].hi; o[2 * q + 1] = acc[q].lo; }
  }
}

// One block per tracer: lane l merges the block partials l, l+256, ... in order, a fixed
// tree merges the lanes, and lane 0 runs the scalar steps that follow each global sum.
constexpr int kFinBlock = 256;
template <int PASS>
__global__ __launch_bounds__(kFinBlock) void k_bc_fin(BcArgs A) {
  const int n = blockIdx.x, l = threadIdx.x;
  constexpr int NS = PASS == P_PM ? 4 : 1;
  DD s[4] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
  const double* pp = A.part + (long)n * A.nblk * 8;
  for (int b = l; b < A.nblk; b += kFinBlock)
#pragma unroll
    for (int q = 0; q < NS; q++) s[q] = dd_merge(s[q], DD{pp[8 * b + 2 * q], pp[8 * b + 2 * q + 1]});
  __shared__ double sh[2 * NS][kFinBlock];
#pragma unroll
  for (int q = 0; q < NS; q++) { sh[2 * q][l] = s[q].hi; sh[2 * q + 1][l] = s[q].lo; }
  __syncthreads();
  for (int w = kFinBlock / 2; w > 0; w >>= 1) {
    if (l < w) {
#pragma unroll
      for (int q = 0; q < NS; q++) {
        DD c = dd_merge(DD{sh[2 * q][l], sh[2 * q + 1][l]}, DD{sh[2 * q][l + w], sh[2 * q + 1][l + w]});
        sh[2 * q][l] = c.hi; sh[2 * q + 1][l] = c.lo;
      }
    }
    __syncthreads();
  }
  if (l != 0) return;
#pragma unroll
  for (int q = 0; q < NS; q++) s[q] = DD{sh[2 * q][0], sh[2 * q + 1][0]};
  double sum0 = s[0].hi + s[0].lo + 0.0;
  double* sc = A.sc + n * kSc;
  if (PASS == P_PM) {
    double sum1 = s[1].hi + s[1].lo + 0.0;
    // Bermejo_Conde.F90:92-105; the flux masses are zero without the Aranami flux
    double mass_bflux = A.flux ? A.flux[2 * n + 1] - A.flux[2 * n] : 0.0 - 0.0;
    double mass_p = sum0 - A.deficit[n];
    sc[SC_MP] = mass_p;
    sc[SC_MM] = sum1;
    sc[SC_MTP] = mass_p + mass_bflux;
    sc[SC_D] = sum1 - sc[SC_MTP];
    sc[SC_CYCLE] = 0.0;
    // weight mass of the sign of D (:112-172): the positive sum, or minus the magnitude
    // sum (a sum of zeros is +0, as a separate pass over signed weights gives)
    double wpos = s[2].hi + s[2].lo + 0.0, wneg = s[3].hi + s[3].lo + 0.0;
    double W = sgn1(sc[SC_D]) > 0.0 ? wpos : (wneg == 0.0 ? 0.0 : -wneg);
    sc[SC_W] = W;
    if (d_almost_zero(W)) sc[SC_CYCLE] = 1.0;                   // :190-198
    else sc[SC_LAMBDA] = sc[SC_D] / W;                          // :200
  } else if (PASS == P_W) {
    sc[SC_W] = sum0;
    if (d_almost_zero(sum0)) sc[SC_CYCLE] = 1.0;               // :190-198
    else sc[SC_LAMBDA] = sc[SC_D] / sum0;                       // :200
  } else if (PASS == P_CORR && A.min_max) {
    sc[SC_MBCR] = sum0;
    if (sc[SC_CYCLE] == 0.0) sc[SC_PROP] = sc[SC_MTP] / sum0;   // :273
  } else {  // final mass: P_PROP, or P_CORR without min/max
    sc[SC_MBC] = sum0;
    if (sc[SC_CYCLE] == 0.0) A.deficit[n] = sum0 - sc[SC_MTP];  // :298-303
    if (A.diag) {
      double* d = A.diag + 6 * n;
      d[0] = sc[SC_MTP]; d[1] = sc[SC_MM]; d[2] = sc[SC_W]; d[3] = sum0; d[4] = sc[SC_CYCLE]; d[5] = A.deficit[n];
    }
  }
}

}  // namespace adz

using namespace adz;

struct adz_bc {
  int tdtype;
  int has_w3;  // some tracer uses weight 3 (needs pkps)
  BcArgs a;
  void* mem;
};

namespace {

constexpr long kBcMaxBlocks = 2048;  // row blocks per tracer (partials kept per block)

// Validate the tracers and scope of a fixer call and fill the host tracer table h[0..ntr)
// and the launch arguments (without device addresses); `who` names the entry point.
int bc_setup(const adz_bc_tracer* tr, int ntr, const adz_bc_domain* dom, const char* who, BcTr* h, BcArgs* out,
             int* has_w3) {
  const adz_bc_domain& d = *dom;
  if (d.in < d.i0 || d.jn < d.j0 || d.nk < d.k0 || d.i0 < d.minx || d.in > d.maxx || d.j0 < d.miny ||
      d.jn > d.maxy || d.k0 < 1)
    return fail(ADZ_EINVAL, std::string(who) + ": scope outside the arrays");
  *has_w3 = 0;
  for (int n = 0; n < ntr; n++) {
    const adz_bc_tracer& t = tr[n];
    if (!t.p || !t.m || !t.lin || !t.mn || !t.mx) return fail(ADZ_EINVAL, std::string(who) + ": null tracer array");
    if (t.weight < 1 || t.weight > 3) return fail(ADZ_EINVAL, "BERMEJO-CONDE: BC WEIGHT not valid");
    if (!t.legacy && (t.pexp_n < 1 || t.pexp_n > 9)) return fail(ADZ_EINVAL, "BERMEJO-CONDE: BC PEXP not valid");
    h[n] = BcTr{t.p, t.m, t.lin, t.mn, t.mx, t.legacy, t.weight,
                t.legacy ? 1.0 : (double)d.pexp_list[t.pexp_n - 1]};
    *has_w3 |= (t.weight == 3);
  }
  BcArgs a;
  std::memset(&a, 0, sizeof(a));
  a.ntr = ntr;
  a.nxa = d.maxx - d.minx + 1; a.nya = d.maxy - d.miny + 1;
  a.i0 = d.i0; a.j0 = d.j0; a.k0 = d.k0; a.nx = d.in - d.i0 + 1; a.ny = d.jn - d.j0 + 1;
  a.npts = (long)a.nx * a.ny * (d.nk - d.k0 + 1);
  a.min_max = d.min_max;
  a.autobar = d.autobar != 0;
  if (a.autobar && d.k0 != 1) return fail(ADZ_EINVAL, std::string(who) + ": Schm_autobar_L needs the scope from level 1");
  a.nrows = a.ny * (d.nk - d.k0 + 1);
  long per = (long)kBcBlock * 16;
  long nblk = (a.npts + per - 1) / per;
  if (nblk > kBcMaxBlocks) nblk = kBcMaxBlocks;
  if (nblk > a.nrows) nblk = a.nrows;
  if (nblk < 1) nblk = 1;
  nblk = std::max(8L, nblk / 8 * 8);  // ntr*nblk a multiple of 8 (one eighth per XCD)
  a.nblk = (int)nblk;
  *out = a;
  return ADZ_OK;
}

}  // namespace

extern "C" {

int adz_bc_create(const adz_bc_tracer* tr, int ntr, const adz_bc_domain* dom, int tdtype, adz_bc** out) {
  if (!tr || ntr <= 0 || !dom || !out) return fail(ADZ_EINVAL, "adz_bc_create: bad argument");
  if (tdtype != ADZ_F32 && tdtype != ADZ_F64) return fail(ADZ_EINVAL, "adz_bc_create: tdtype must be F32 or F64");
  std::vector<BcTr> h(ntr);
  BcArgs a;
  int has_w3 = 0;
  if (int rc = bc_setup(tr, ntr, dom, "adz_bc_create", h.data(), &a, &has_w3)) return rc;
  const long nblk = a.nblk;
  size_t tb = sizeof(BcTr) * ntr, pb = sizeof(double) * 8 * ntr * nblk, sb = sizeof(double) * kSc * ntr;
  void* mem = nullptr;
  if (hipMalloc(&mem, tb + pb + sb) != hipSuccess) return fail(ADZ_ENOMEM, "adz_bc_create: hipMalloc");
  if (hipMemcpy(mem, h.data(), tb, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemset((char*)mem + tb, 0, pb + sb) != hipSuccess) {
    (void)hipFree(mem);
    return fail(ADZ_EHIP, "adz_bc_create: upload");
  }
  a.tr = (const BcTr*)mem;
  a.part = (double*)((char*)mem + tb);
  a.sc = (double*)((char*)mem + tb + pb);
  adz_bc* b = new adz_bc;
  b->tdtype = tdtype; b->a = a; b->mem = mem;
  b->has_w3 = has_w3;
  *out = b;
  return ADZ_OK;
}

}  // extern "C"

namespace {

// The fixer's passes on the stream: a.tr, a.part and a.sc address device memory.
int bc_launch(BcArgs a, int tdtype, int has_w3, const void* airm1, const void* airm0, const void* pkps,
              const double* mask, const double* flux, double* deficit, double* diag, adz_stream stream) {
  if (!airm1 || !airm0 || !mask || !deficit) return fail(ADZ_EINVAL, "adz_bermejo_conde: null argument");
  a.airm1 = airm1; a.airm0 = airm0; a.pkps = pkps; a.mask = mask; a.deficit = deficit; a.diag = diag;
  a.flux = flux;
  if (has_w3 && !pkps) return fail(ADZ_EINVAL, "adz_bermejo_conde: weight 3 needs pkps");
  hipStream_t st = (hipStream_t)stream;
  dim3 g((unsigned)((long)((a.ntr + kBcWaves - 1) / kBcWaves) * a.nblk)), b(kBcBlock);
  dim3 gf((unsigned)a.ntr), bf(kFinBlock);
#define ADZ_BC_PASSES(T)                                                       \
  hipLaunchKernelGGL((k_bc_pass<T, P_PM>), g, b, 0, st, a);                   \
  hipLaunchKernelGGL((k_bc_fin<P_PM>), gf, bf, 0, st, a);                     \
  hipLaunchKernelGGL((k_bc_pass<T, P_CORR>), g, b, 0, st, a);                 \
  hipLaunchKernelGGL((k_bc_fin<P_CORR>), gf, bf, 0, st, a);                   \
  if (a.min_max) {                                                            \
    hipLaunchKernelGGL((k_bc_pass<T, P_PROP>), g, b, 0, st, a);               \
    hipLaunchKernelGGL((k_bc_fin<P_PROP>), gf, bf, 0, st, a);                 \
  }
  if (tdtype == ADZ_F32) {
    ADZ_BC_PASSES(float)
  } else {
    ADZ_BC_PASSES(double)
  }
#undef ADZ_BC_PASSES
  ADZ_HIP_CHECK(hipGetLastError());
  return ADZ_OK;
}

int bc_apply(adz_bc* bc, const void* airm1, const void* airm0, const void* pkps, const double* mask,
             const double* flux, double* deficit, double* diag, adz_stream stream) {
  if (!bc) return fail(ADZ_EINVAL, "adz_bermejo_conde: null argument");
  return bc_launch(bc->a, bc->tdtype, bc->has_w3, airm1, airm0, pkps, mask, flux, deficit, diag, stream);
}

}  // namespace

extern "C" {

int adz_bermejo_conde(adz_bc* bc, const void* airm1, const void* airm0, const void* pkps, const double* mask,
                      double* deficit, double* diag, adz_stream stream) {
  return bc_apply(bc, airm1, airm0, pkps, mask, nullptr, deficit, diag, stream);
}

int adz_bermejo_conde_lam(adz_bc* bc, const void* airm1, const void* airm0, const void* pkps, const double* mask,
                          const double* flux_mass, double* deficit, double* diag, adz_stream stream) {
  if (!flux_mass) return fail(ADZ_EINVAL, "adz_bermejo_conde_lam: null flux masses");
  return bc_apply(bc, airm1, airm0, pkps, mask, flux_mass, deficit, diag, stream);
}

}  // extern "C"

// ---------------------------------------------------------------------------------
// The fixer without allocation.  The tracer table, the block partials and the per-tracer
// scalars of one call live in `scratch`, a file-scope device array of fixed size that each
// GPU holds for the life of the process; the host reaches it through the address that
// hipGetSymbolAddress gives on that device (the array's name in host code is only a
// stand-in).  Calls that overlap on one GPU share the array and race.
namespace {

constexpr int kBcStaticTracers = 64;
constexpr size_t kBcStaticTb = (sizeof(BcTr) * kBcStaticTracers + 255) / 256 * 256;
constexpr size_t kBcStaticPb = sizeof(double) * 8 * kBcStaticTracers * kBcMaxBlocks;
constexpr size_t kBcStaticSb = sizeof(double) * kSc * kBcStaticTracers;
constexpr size_t kBcScratchBytes = kBcStaticTb + kBcStaticPb + kBcStaticSb;  // 8.4 MB
constexpr int kFillChunk = 16;
constexpr int kMaxDevices = 64;

struct BcTrChunk {
  BcTr t[kFillChunk];
  int n0, cnt;
};

}  // namespace

static __device__ __align__(256) unsigned char scratch[kBcScratchBytes];

namespace {

// tracers n0 .. n0+cnt-1 into the table; their scalars start at zero
__global__ __launch_bounds__(256) void k_bc_static_fill(BcTrChunk c, BcTr* table, double* sc) {
  const int l = threadIdx.x;
  if (l < c.cnt) table[c.n0 + l] = c.t[l];
  for (int q = l; q < c.cnt * kSc; q += 256) sc[(long)c.n0 * kSc + q] = 0.0;
}

std::mutex g_scratch_mu;
void* g_scratch[kMaxDevices];

// the device address of `scratch` on the current device (looked up once per device)
int scratch_address(void** out) {
  int dev = 0;
  ADZ_HIP_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(ADZ_EINVAL, "adz_bermejo_conde_static: device index too large");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(ADZ_EHIP, std::string("adz_bermejo_conde_static: hipGetSymbolAddress: ") + hipGetErrorString(e));
    g_scratch[dev] = p;
  }
  *out = g_scratch[dev];
  return ADZ_OK;
}

}  // namespace

extern "C" {

int adz_bermejo_conde_static(const adz_bc_tracer* tr, int ntr, const adz_bc_domain* dom, int tdtype,
                             const void* airm1, const void* airm0, const void* pkps, const double* mask,
                             const double* flux_mass, double* deficit, double* diag, adz_stream stream) {
  if (!tr || ntr <= 0 || !dom) return fail(ADZ_EINVAL, "adz_bermejo_conde_static: bad argument");
  if (ntr > kBcStaticTracers)
    return fail(ADZ_EINVAL, "adz_bermejo_conde_static: at most 64 tracers per call (the scratch array's size)");
  if (tdtype != ADZ_F32 && tdtype != ADZ_F64)
    return fail(ADZ_EINVAL, "adz_bermejo_conde_static: tdtype must be F32 or F64");
  BcTr h[kBcStaticTracers];
  BcArgs a;
  int has_w3 = 0;
  if (int rc = bc_setup(tr, ntr, dom, "adz_bermejo_conde_static", h, &a, &has_w3)) return rc;
  if (!airm1 || !airm0 || !mask || !deficit) return fail(ADZ_EINVAL, "adz_bermejo_conde_static: null argument");
  if (has_w3 && !pkps) return fail(ADZ_EINVAL, "adz_bermejo_conde_static: weight 3 needs pkps");
  void* base = nullptr;
  if (int rc = scratch_address(&base)) return rc;
  char* p = (char*)base;
  BcTr* table = (BcTr*)p;
  a.tr = table;
  a.part = (double*)(p + kBcStaticTb);
  a.sc = (double*)(p + kBcStaticTb + kBcStaticPb);
  hipStream_t st = (hipStream_t)stream;
  for (int n0 = 0; n0 < ntr; n0 += kFillChunk) {
    BcTrChunk c;
    std::memset(&c, 0, sizeof(c));
    c.n0 = n0;
    c.cnt = std::min(kFillChunk, ntr - n0);
    for (int q = 0; q < c.cnt; q++) c.t[q] = h[n0 + q];
    hipLaunchKernelGGL(k_bc_static_fill, dim3(1), dim3(256), 0, st, c, table, a.sc);
  }
  ADZ_HIP_CHECK(hipGetLastError());
  return bc_launch(a, tdtype, has_w3, airm1, airm0, pkps, mask, flux_mass, deficit, diag, stream);
}

int adz_bc_free(adz_bc* bc) {
  if (!bc) return ADZ_OK;
  hipError_t e = hipFree(bc->mem);
  delete bc;
  if (e != hipSuccess) return fail(ADZ_EHIP, "adz_bc_free: hipFree");
  return ADZ_OK;
}

}  // extern "C"
