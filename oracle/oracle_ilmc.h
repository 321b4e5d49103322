// CPU ORACLE -- test infrastructure only (tests/, smoke() and bench.py's CPU baseline).
//
// ILMC shape preservation (mono = 2; Sorensen et al. 2013), restated from
//   - ILMC_LAM.F90:166-214: the cells of each sweep around (i,j,k), in the order j, k, i
//     (outer to inner), clipped to [il:ir] x [jl:jr] x [k0t:nk]; with Schm_autobar_L the
//     level k only;
//   - ilmc_lam_loop.inc:16-139: the sequential visit in (k, j, i) order of every cell of
//     the domain; an over- or undershoot moves mass from the cell into the cells of sweep
//     1, 2, ..., up to their local max (min);
//   - ILMC_LAM.F90:249-273: the min/max reset of Adz_ILMC_min_max_L.
// The arithmetic is the reference's `real` (FP32) arithmetic, evaluated as written (no
// contraction into FMA).  For a tile that touches the four LAM edges (il = F_i0, ir =
// F_in, jl = F_j0, jr = F_jn; ILMC_LAM.F90:81-86).
#pragma once
#include <vector>

namespace ora {

struct IlmcDom {
  int minx, maxx, miny, maxy, nk;  // array layout (l_minx:l_maxx, l_miny:l_maxy, nk)
  int i0, in, j0, jn, k0;          // F_i0..F_in x F_j0..F_jn x Adz_k0t..nk
  int sweep_max;                   // Adz_ILMC_sweep_max
  int min_max;                     // Adz_ILMC_min_max_L
  int autobar;                     // Schm_autobar_L
};

inline float il_max(float a, float b) { return (b > a) ? b : a; }
inline float il_min(float a, float b) { return (b < a) ? b : a; }

// The cells of sweep s around (i,j,k) as array offsets, in the reference's order
// (ILMC_LAM.F90:177-207).
inline void ilmc_shell(const IlmcDom& d, int i, int j, int k, int s, std::vector<long>& out) {
  const long nx = d.maxx - d.minx + 1, nxy = nx * (d.maxy - d.miny + 1);
  out.clear();
  int kx_m = k, kx_p = k;
  if (!d.autobar) {
    kx_m = std::max(k - s, d.k0);
    kx_p = std::min(k + s, d.nk);
  }
  const int j1 = std::max(j - s, d.j0), j2 = std::min(j + s, d.jn);
  for (int jx = j1; jx <= j2; jx++)
    for (int kx = kx_m; kx <= kx_p; kx++) {
      bool limit_i = (jx != j - s && jx != j + s);
      if (!d.autobar) limit_i = limit_i && (kx > k - s && kx < k + s);
      for (int ix = std::max(i - s, d.i0); ix <= std::min(i + s, d.in); ix++) {
        if (limit_i && ix != i - s && ix != i + s) continue;
        out.push_back((ix - d.minx) + nx * (jx - d.miny) + nxy * (kx - 1));
      }
    }
}

// In place on f (flat arrays of the layout).  counts (may be null): [0] sweeps that moved
// mass (reset(:,5)), [1] min resets (reset(:,3)), [2] max resets (reset(:,4)).
inline void ilmc_lam(float* f, const float* mn, const float* mx, const float* air, const IlmcDom& d,
                     long long* counts) {
  const long nx = d.maxx - d.minx + 1, nxy = nx * (d.maxy - d.miny + 1);
  long long moved = 0, rmin = 0, rmax = 0;
  std::vector<long> sh;
  std::vector<float> sw;
  for (int k = d.k0; k <= d.nk; k++)
    for (int j = d.j0; j <= d.jn; j++)
      for (int i = d.i0; i <= d.in; i++) {
        const long c = (i - d.minx) + nx * (j - d.miny) + nxy * (k - 1);
        float o_shoot = il_max(0.f, f[c] - mx[c]);
        float u_shoot = il_max(0.f, mn[c] - f[c]);
        if (o_shoot == 0.f && u_shoot == 0.f) continue;
        const bool over = o_shoot > 0.f;
        for (int s = 1; s <= d.sweep_max; s++) {
          ilmc_shell(d, i, j, k, s, sh);
          sw.assign(sh.size(), 0.f);
          float m_sweep = 0.f;
          for (size_t ii = 0; ii < sh.size(); ii++) {
            const long q = sh[ii];
            sw[ii] = over ? il_max(0.f, mx[q] - f[q]) : il_max(0.f, f[q] - mn[q]);
            m_sweep = sw[ii] * air[q] + m_sweep;
          }
          const float m_use = il_min(m_sweep, (over ? o_shoot : u_shoot) * air[c]);
          if (!(m_use > 0.f)) continue;
          moved++;
          if (over) {
            f[c] = il_max(f[c] - m_use / air[c], mn[c]);
            o_shoot = il_max(0.f, f[c] - mx[c]);
          } else {
            f[c] = il_min(f[c] + m_use / air[c], mx[c]);
            u_shoot = il_max(0.f, mn[c] - f[c]);
          }
          const float ratio = m_use / m_sweep;
          for (size_t ii = 0; ii < sh.size(); ii++) {
            if (!(sw[ii] > 0.f)) continue;
            const long q = sh[ii];
            f[q] = over ? il_min(f[q] + sw[ii] * ratio, mx[q]) : il_max(f[q] - sw[ii] * ratio, mn[q]);
          }
        }
      }
  if (d.min_max)
    for (int k = d.k0; k <= d.nk; k++)
      for (int j = d.j0; j <= d.jn; j++)
        for (int i = d.i0; i <= d.in; i++) {
          const long c = (i - d.minx) + nx * (j - d.miny) + nxy * (k - 1);
          if (f[c] < mn[c]) { rmin++; f[c] = mn[c]; }
          if (f[c] > mx[c]) { rmax++; f[c] = mx[c]; }
        }
  if (counts) { counts[0] = moved; counts[1] = rmin; counts[2] = rmax; }
}

}  // namespace ora
