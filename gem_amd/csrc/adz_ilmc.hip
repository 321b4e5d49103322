// ILMC shape preservation (mono = 2; Sorensen et al. 2013, GMD 6, 1029-1042) on gfx950:
// ILMC_LAM.F90:18-273 with its sweep loop ilmc_lam_loop.inc:16-139, for a tile that touches
// the four LAM edges (the domain is F_i0..F_in x F_j0..F_jn x Adz_k0t..nk,
// ILMC_LAM.F90:81-86).
//
// The reference visits every cell of the domain in (k, j, i) order.  A cell above its max
// (below its min) moves mass into (out of) the cells of sweep 1, 2, ..., S around it -- the
// cells at Chebyshev distance s, clipped to the domain -- up to their own max (min), and a
// later visit sees the values the earlier ones left.  That order is kept exactly, in
// parallel:
//   * a cell inside [min, max] never leaves it: mass is added up to a cell's max and
//     removed down to its min, and only where there is room.  So the cells to visit are the
//     cells that violate their bounds before the first visit; the loop skips the others
//     (ilmc_lam_loop.inc:26) and so does this code;
//   * a visit reads and writes only the cube of half-width S around its cell, so two
//     visits commute unless their cells lie within Chebyshev distance 2S;
//   * with a = 2S+1 and b = a*a, the hyperplane number t = di + a*dj + b*dk (offsets from
//     the domain's first cell; b = 0 with Schm_autobar_L, whose sweeps stay on one level)
//     is strictly larger for the later of two cells within distance 2S, and two distinct
//     cells with the same t are more than 2S apart.
// The violating cells are bucketed by t, and one launch per hyperplane visits its bucket
// in parallel: each visit sees exactly the values the sequential loop gives it.  One wave
// visits one cell: the lanes gather the sweep's cells, lane 0 adds their masses in the
// reference's order (j, then k, then i; cube positions that are not on the sweep add +0,
// which leaves a sum of non-negative FP32 terms unchanged), and the lanes redistribute.
// Arithmetic is the reference's `real` arithmetic as written (built without contraction).
#include <algorithm>
#include <cstring>
#include <vector>

#include "adz_common.h"

namespace adz {

constexpr int kIlBlock = 256;
constexpr int kIlScan = 1024;

struct IlmcTr {
  float* f;
  const float* mn;
  const float* mx;
};

struct IlmcArgs {
  const IlmcTr* tr;
  int ntr;
  long nx, nxy;                 // array pitches (pointers address element (1,1,1))
  int i0, j0, k0, ni, nj, nkk;  // domain origin and extents
  int a, b;                     // hyperplane coefficients
  int nplanes;
  const float* air;
  unsigned long long* cnt;  // [nplanes] cells per hyperplane, then fill cursors
  long long* off;           // [nplanes + 1] bucket starts
  uint64_t* list;  // cells by hyperplane: i | j << 16 | k << 32 | tracer << 44 (domain offsets)
  long long* counts;
};

__device__ __forceinline__ float il_max(float a, float b) { return (b > a) ? b : a; }
__device__ __forceinline__ float il_min(float a, float b) { return (b < a) ? b : a; }

__global__ __launch_bounds__(kIlBlock) void k_il_zero(long long* p, long n) {
  const long q = (long)blockIdx.x * kIlBlock + threadIdx.x;
  if (q < n) p[q] = 0;
}

// One block per (row of the domain, tracer): find the cells outside [min, max]; count them
// per hyperplane (fill = 0) or append them to their hyperplane's bucket (fill = 1).  The
// order inside a bucket does not matter: its cells are more than 2S apart.
__global__ __launch_bounds__(kIlBlock) void k_il_mark(IlmcArgs A, int fill) {
  const int row = blockIdx.x, n = blockIdx.y;
  const int dj = row % A.nj, dk = row / A.nj;
  const IlmcTr T = A.tr[n];
  const long base = (long)(A.i0 - 1) + A.nx * (A.j0 + dj - 1) + A.nxy * (A.k0 + dk - 1);
  const int t0 = A.a * dj + A.b * dk;
  for (int di = threadIdx.x; di < A.ni; di += kIlBlock) {
    const long c = base + di;
    const float f = T.f[c];
    const float o = il_max(0.f, f - T.mx[c]), u = il_max(0.f, T.mn[c] - f);
    if (o == 0.f && u == 0.f) continue;
    const int t = t0 + di;
    if (!fill) {
      atomicAdd(&A.cnt[t], 1ull);
    } else {
      const unsigned long long pos = atomicAdd(&A.cnt[t], 1ull);
      A.list[pos] = (uint64_t)di | ((uint64_t)dj << 16) | ((uint64_t)dk << 32) | ((uint64_t)n << 44);
    }
  }
}

// One block: off = exclusive prefix sum of cnt, off[nplanes] = total; cnt becomes the fill
// cursor (= off).
__global__ __launch_bounds__(kIlScan) void k_il_scan(IlmcArgs A) {
  __shared__ long long s[kIlScan];
  const int t = threadIdx.x;
  long long run = 0;
  for (int b0 = 0; b0 < A.nplanes; b0 += kIlScan) {
    const int q = b0 + t;
    const long long v = q < A.nplanes ? (long long)A.cnt[q] : 0;
    s[t] = v;
    __syncthreads();
    for (int o = 1; o < kIlScan; o <<= 1) {
      const long long w = t >= o ? s[t - o] : 0;
      __syncthreads();
      s[t] += w;
      __syncthreads();
    }
    if (q < A.nplanes) {
      A.off[q] = run + s[t] - v;
      A.cnt[q] = (unsigned long long)(run + s[t] - v);
    }
    run += s[kIlScan - 1];
    __syncthreads();
  }
  if (t == 0) A.off[A.nplanes] = run;
}

// Sweep s of the visit of cell c (wave-wide; lane 0 holds the cell's current value).
// Returns 1 if the sweep moved mass, plus 2 if the cell is back in bounds: every later sweep
// then computes m_use = min(m_sweep, 0) <= 0 and moves nothing (ilmc_lam_loop.inc:47,51),
// so the visit may stop there.
template <int s, bool AB>
__device__ __forceinline__ int il_sweep(const IlmcArgs& A, const IlmcTr& T, long c, int di, int dj, int dk,
                                        bool over, float& fc, float& o, float& u, float mnc, float mxc, float airc,
                                        float* buf, float* bc) {
  constexpr int W = 2 * s + 1, P = AB ? W * W : W * W * W, NQ = (P + 63) / 64;
  const int lane = threadIdx.x;
  float sw[NQ], fq[NQ], bq[NQ];
  long qq[NQ];
#pragma unroll
  for (int r = 0; r < NQ; r++) {
    const int p = lane + 64 * r;
    sw[r] = 0.f;
    fq[r] = 0.f;
    bq[r] = 0.f;
    qq[r] = 0;
    if (p < P) {
      const int ei = p % W - s;
      const int ek = AB ? 0 : (p / W) % W - s;
      const int ej = (AB ? p / W : p / (W * W)) - s;
      const int m = max(abs(ei), max(abs(ej), abs(ek)));
      float prod = 0.f;
      if (m == s && di + ei >= 0 && di + ei < A.ni && dj + ej >= 0 && dj + ej < A.nj && dk + ek >= 0 &&
          dk + ek < A.nkk) {
        const long q = c + ei + A.nx * ej + A.nxy * ek;
        const float f = T.f[q];
        const float bnd = over ? T.mx[q] : T.mn[q];
        const float w = over ? il_max(0.f, bnd - f) : il_max(0.f, f - bnd);
        prod = w * A.air[q];
        sw[r] = w;
        fq[r] = f;
        bq[r] = bnd;
        qq[r] = q;
      }
      buf[p] = prod;
    }
  }
  __syncthreads();
  if (lane == 0) {
    float msum = 0.f;  // ilmc_lam_loop.inc:34-45 (:86-97): masses in the sweep's order
    for (int p = 0; p < P; p++) msum = buf[p] + msum;
    const float m_use = il_min(msum, (over ? o : u) * airc);  // :47 (:99)
    float ratio = 0.f, mv = 0.f;
    if (m_use > 0.f) {  // :51-76 (:103-128)
      mv = 1.f;
      if (over) {
        fc = il_max(fc - m_use / airc, mnc);
        o = il_max(0.f, fc - mxc);
      } else {
        fc = il_min(fc + m_use / airc, mxc);
        u = il_max(0.f, mnc - fc);
      }
      T.f[c] = fc;
      ratio = m_use / msum;
    }
    bc[0] = ratio;
    bc[1] = mv;
    bc[2] = ((over ? o : u) == 0.f) ? 1.f : 0.f;
  }
  __syncthreads();
  const float ratio = bc[0];
  const int moved = bc[1] != 0.f;
  const int done = bc[2] != 0.f;
  if (moved) {
#pragma unroll
    for (int r = 0; r < NQ; r++)
      if (sw[r] > 0.f) T.f[qq[r]] = over ? il_min(fq[r] + sw[r] * ratio, bq[r]) : il_max(fq[r] - sw[r] * ratio, bq[r]);
  }
  __syncthreads();  // buf and bc are reused by the next sweep
  return moved | (done << 1);
}

// The visits of hyperplane t: one wave per cell, waves stride over the bucket.
template <int S, bool AB>
__global__ __launch_bounds__(64) void k_il_plane(IlmcArgs A, int t) {
  constexpr int W = 2 * S + 1, P = AB ? W * W : W * W * W;
  __shared__ float buf[P];
  __shared__ float bc[3];
  const long long beg = A.off[t], end = A.off[t + 1];
  for (long long e = beg + blockIdx.x; e < end; e += gridDim.x) {
    const uint64_t code = A.list[e];
    const int di = (int)(code & 0xffff), dj = (int)((code >> 16) & 0xffff), dk = (int)((code >> 32) & 0xfff);
    const int n = (int)(code >> 44);
    const IlmcTr T = A.tr[n];
    const long c = (long)(A.i0 + di - 1) + A.nx * (A.j0 + dj - 1) + A.nxy * (A.k0 + dk - 1);
    float fc = T.f[c];
    const float mnc = T.mn[c], mxc = T.mx[c], airc = A.air[c];
    float o = il_max(0.f, fc - mxc), u = il_max(0.f, mnc - fc);  // :23-26
    if (o == 0.f && u == 0.f) continue;                          // put back in bounds by earlier visits
    const bool over = o > 0.f;
    int r = il_sweep<1, AB>(A, T, c, di, dj, dk, over, fc, o, u, mnc, mxc, airc, buf, bc);
    int moved = r & 1;
    if constexpr (S >= 2) {
      if (!(r & 2)) {
        r = il_sweep<2, AB>(A, T, c, di, dj, dk, over, fc, o, u, mnc, mxc, airc, buf, bc);
        moved += r & 1;
      }
    }
    if constexpr (S >= 3) {
      if (!(r & 2)) {
        r = il_sweep<3, AB>(A, T, c, di, dj, dk, over, fc, o, u, mnc, mxc, airc, buf, bc);
        moved += r & 1;
      }
    }
    if constexpr (S >= 4) {
      if (!(r & 2)) {
        r = il_sweep<4, AB>(A, T, c, di, dj, dk, over, fc, o, u, mnc, mxc, airc, buf, bc);
        moved += r & 1;
      }
    }
    if (A.counts && threadIdx.x == 0 && moved) atomicAdd((unsigned long long*)&A.counts[3 * n], (unsigned long long)moved);
  }
}

// ILMC_LAM.F90:249-273: reset to [min, max]; counts of both resets.
__global__ __launch_bounds__(kIlBlock) void k_il_reset(IlmcArgs A) {
  const int row = blockIdx.x, n = blockIdx.y;
  const int dj = row % A.nj, dk = row / A.nj;
  const IlmcTr T = A.tr[n];
  const long base = (long)(A.i0 - 1) + A.nx * (A.j0 + dj - 1) + A.nxy * (A.k0 + dk - 1);
  int nmin = 0, nmax = 0;
  for (int di = threadIdx.x; di < A.ni; di += kIlBlock) {
    const long c = base + di;
    float f = T.f[c];
    const float mn = T.mn[c], mx = T.mx[c];
    const bool lo = f < mn;
    if (lo) f = mn;
    const bool hi = f > mx;
    if (hi) f = mx;
    if (lo || hi) T.f[c] = f;
    nmin += lo;
    nmax += hi;
  }
  if (A.counts) {
    for (int o = 32; o > 0; o >>= 1) {
      nmin += __shfl_down(nmin, o);
      nmax += __shfl_down(nmax, o);
    }
    if ((threadIdx.x & 63) == 0) {
      if (nmin) atomicAdd((unsigned long long*)&A.counts[3 * n + 1], (unsigned long long)nmin);
      if (nmax) atomicAdd((unsigned long long*)&A.counts[3 * n + 2], (unsigned long long)nmax);
    }
  }
}

}  // namespace adz

using namespace adz;

struct adz_ilmc {
  IlmcArgs a;
  int sweep_max, autobar, min_max;
  int plane_blocks;
  void* mem;
};

namespace {

template <int S, bool AB>
void launch_planes(const IlmcArgs& a, int blocks, hipStream_t st) {
  for (int t = 0; t < a.nplanes; t++) hipLaunchKernelGGL((k_il_plane<S, AB>), dim3(blocks), dim3(64), 0, st, a, t);
}

}  // namespace

extern "C" {

int adz_ilmc_create(const adz_ilmc_tracer* tr, int ntr, const adz_ilmc_domain* dom, adz_ilmc** out) {
  if (!tr || ntr <= 0 || !dom || !out) return fail(ADZ_EINVAL, "adz_ilmc_create: bad argument");
  const adz_ilmc_domain& d = *dom;
  if (d.in < d.i0 || d.jn < d.j0 || d.nk < d.k0 || d.i0 < d.minx || d.in > d.maxx || d.j0 < d.miny ||
      d.jn > d.maxy || d.k0 < 1)
    return fail(ADZ_EINVAL, "adz_ilmc_create: domain outside the arrays");
  if (d.sweep_max < 1 || d.sweep_max > 4)
    return fail(ADZ_EINVAL, "adz_ilmc_create: ILMC_sweep_max must be 1..4 (the sweep arrays hold 400 cells)");
  const int ni = d.in - d.i0 + 1, nj = d.jn - d.j0 + 1, nkk = d.nk - d.k0 + 1;
  // cell offsets are packed in 16 + 16 + 12 bits, tracers in the remaining 20; the row
  // kernels launch one 256-lane block per row of the domain (at most 2^32 lanes)
  if (ni > 65535 || nj > 65535 || nkk > 4095 || ntr > 65535 || (long)nj * nkk > (1L << 24))
    return fail(ADZ_EINVAL, "adz_ilmc_create: domain or tracer count too large");
  std::vector<IlmcTr> h(ntr);
  for (int n = 0; n < ntr; n++) {
    if (!tr[n].f || !tr[n].mn || !tr[n].mx) return fail(ADZ_EINVAL, "adz_ilmc_create: null tracer array");
    h[n] = IlmcTr{(float*)tr[n].f, (const float*)tr[n].mn, (const float*)tr[n].mx};
  }
  IlmcArgs a;
  std::memset(&a, 0, sizeof(a));
  a.ntr = ntr;
  a.nx = d.maxx - d.minx + 1;
  a.nxy = a.nx * (long)(d.maxy - d.miny + 1);
  a.i0 = d.i0; a.j0 = d.j0; a.k0 = d.k0; a.ni = ni; a.nj = nj; a.nkk = nkk;
  a.a = 2 * d.sweep_max + 1;
  a.b = d.autobar ? 0 : a.a * a.a;
  const long np = (long)(ni - 1) + (long)a.a * (nj - 1) + (long)a.b * (nkk - 1) + 1;
  const long cap = (long)ntr * ni * nj * nkk;
  if (np > (1L << 30)) return fail(ADZ_EINVAL, "adz_ilmc_create: domain too large");
  a.nplanes = (int)np;
  const size_t tb = (sizeof(IlmcTr) * ntr + 255) / 256 * 256;
  const size_t cb = (sizeof(long long) * np + 255) / 256 * 256, ob = (sizeof(long long) * (np + 1) + 255) / 256 * 256;
  const size_t lb = sizeof(uint64_t) * cap;
  void* mem = nullptr;
  if (hipMalloc(&mem, tb + cb + ob + lb) != hipSuccess) return fail(ADZ_ENOMEM, "adz_ilmc_create: hipMalloc");
  if (hipMemcpy(mem, h.data(), sizeof(IlmcTr) * ntr, hipMemcpyHostToDevice) != hipSuccess) {
    (void)hipFree(mem);
    return fail(ADZ_EHIP, "adz_ilmc_create: upload");
  }
  char* p = (char*)mem;
  a.tr = (const IlmcTr*)p;
  a.cnt = (unsigned long long*)(p + tb);
  a.off = (long long*)(p + tb + cb);
  a.list = (uint64_t*)(p + tb + cb + ob);
  adz_ilmc* r = new adz_ilmc;
  r->a = a;
  r->sweep_max = d.sweep_max; r->autobar = d.autobar != 0; r->min_max = d.min_max != 0;
  // waves per hyperplane launch: enough for the busiest buckets when ~1/4 of the cells
  // violate, bounded to 32 per CU
  const long want = cap / np / 4;
  r->plane_blocks = (int)std::min(8192L, std::max(64L, want));
  r->mem = mem;
  *out = r;
  return ADZ_OK;
}

int adz_ilmc_lam(adz_ilmc* il, const void* air_m, long long* counts, adz_stream stream) {
  if (!il || !air_m) return fail(ADZ_EINVAL, "adz_ilmc_lam: null argument");
  hipStream_t st = (hipStream_t)stream;
  IlmcArgs a = il->a;
  a.air = (const float*)air_m;
  a.counts = counts;
  if (counts)
    hipLaunchKernelGGL(k_il_zero, dim3((unsigned)((3L * a.ntr + kIlBlock - 1) / kIlBlock)), dim3(kIlBlock), 0, st,
                       counts, 3L * a.ntr);
  hipLaunchKernelGGL(k_il_zero, dim3((unsigned)((a.nplanes + kIlBlock - 1) / kIlBlock)), dim3(kIlBlock), 0, st,
                     (long long*)a.cnt, (long)a.nplanes);
  const dim3 gr((unsigned)(a.nj * a.nkk), (unsigned)a.ntr);
  hipLaunchKernelGGL(k_il_mark, gr, dim3(kIlBlock), 0, st, a, 0);
  hipLaunchKernelGGL(k_il_scan, dim3(1), dim3(kIlScan), 0, st, a);
  hipLaunchKernelGGL(k_il_mark, gr, dim3(kIlBlock), 0, st, a, 1);
  const int nb = il->plane_blocks;
  switch (il->sweep_max * 2 + il->autobar) {
    case 2: launch_planes<1, false>(a, nb, st); break;
    case 3: launch_planes<1, true>(a, nb, st); break;
    case 4: launch_planes<2, false>(a, nb, st); break;
    case 5: launch_planes<2, true>(a, nb, st); break;
    case 6: launch_planes<3, false>(a, nb, st); break;
    case 7: launch_planes<3, true>(a, nb, st); break;
    case 8: launch_planes<4, false>(a, nb, st); break;
    default: launch_planes<4, true>(a, nb, st); break;
  }
  if (il->min_max) hipLaunchKernelGGL(k_il_reset, gr, dim3(kIlBlock), 0, st, a);
  ADZ_HIP_CHECK(hipGetLastError());
  return ADZ_OK;
}

int adz_ilmc_free(adz_ilmc* il) {
  if (!il) return ADZ_OK;
  hipError_t e = hipFree(il->mem);
  delete il;
  if (e != hipSuccess) return fail(ADZ_EHIP, "adz_ilmc_free: hipFree");
  return ADZ_OK;
}

}  // extern "C"
