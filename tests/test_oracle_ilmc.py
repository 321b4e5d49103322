"""CPU: the oracle's ILMC (ILMC_LAM.F90:18-273, ilmc_lam_loop.inc:16-139) and the premise of
the GPU schedule.

Pins:
  * the reference's own sweep loop, compiled from ilmc_lam_loop.inc inside
    oracle/ilmc_ref_host.F90 (oracle/_ref), agrees with the oracle bit for bit;
  * the sweep cells equal, in order, a brute-force enumeration of the cells at Chebyshev
    distance s (j, then k, then i), clipped to the domain;
  * one isolated overshoot with room in sweep 1 has the closed-form result (the cell drops
    to its max, each neighbour gains its room times o * air_c / sum(room * air));
  * cells inside [min, max] never leave it; mass (sum f * air) moves only by FP32 rounding;
  * the hyperplane numbering t = i + a j + b k orders every pair of conflicting cells the
    way the sequential loop does, and separates the cells of one hyperplane by > 2S.
"""
import itertools
import os

import numpy as np
import pytest

from oracle import oracle as O

MH = 2


def _case(ni, nj, nk, seed, spread=0.3):
    layout = (1 - MH, ni + MH, 1 - MH, nj + MH, nk)
    shape = (nk, nj + 2 * MH, ni + 2 * MH)
    n = int(np.prod(shape))
    rng = np.random.default_rng(seed)
    src = rng.lognormal(0.0, 1.0, n).astype(np.float32)
    mn = (src * rng.uniform(1.0 - spread, 1.0, n)).astype(np.float32)
    mx = (src * rng.uniform(1.0, 1.0 + spread, n)).astype(np.float32)
    f = (src * rng.uniform(1.0 - 2 * spread, 1.0 + 2 * spread, n)).astype(np.float32)
    air = rng.uniform(0.5, 2.0, n).astype(np.float32)
    return layout, shape, f, mn, mx, air


@pytest.mark.skipif(not os.path.exists(O.REF_ILMC), reason="oracle/_ref not built (reference tree absent)")
@pytest.mark.parametrize("sweep_max", [1, 2, 3, 4])
@pytest.mark.parametrize("autobar", [False, True])
def test_oracle_equals_reference_loop(sweep_max, autobar):
    layout, shape, f, mn, mx, air = _case(22, 18, 9, 10 * sweep_max + autobar)
    box = (3, 20, 2, 17, 2)
    a, b = f.copy(), f.copy()
    counts = O.ilmc(a, mn, mx, air, layout, box, sweep_max=sweep_max, min_max=False, autobar=autobar)
    moved = O.RefIlmc()(b, mn, mx, air, layout, box, sweep_max=sweep_max, autobar=autobar)
    assert moved == counts[0] > 100
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    assert np.count_nonzero(a != f) > 500


@pytest.mark.parametrize("autobar", [False, True])
@pytest.mark.parametrize("s", [1, 2, 3])
def test_sweep_cells_brute_force(s, autobar):
    layout = (-1, 12, -1, 10, 8)
    minx, maxx, miny, maxy, nk = layout
    box = (1, 10, 2, 9, 2)
    nx, nxy = maxx - minx + 1, (maxx - minx + 1) * (maxy - miny + 1)
    for (i, j, k) in [(1, 2, 2), (5, 5, 5), (10, 9, 8), (2, 8, 3), (9, 3, 7)]:
        exp = []
        for jx in range(j - s, j + s + 1):
            for kx in ([k] if autobar else range(k - s, k + s + 1)):
                for ix in range(i - s, i + s + 1):
                    d = max(abs(ix - i), abs(jx - j), 0 if autobar else abs(kx - k))
                    if d != s or not (box[0] <= ix <= box[1] and box[2] <= jx <= box[3] and box[4] <= kx <= nk):
                        continue
                    exp.append((ix - minx) + nx * (jx - miny) + nxy * (kx - 1))
        got = O.ilmc_shell(layout, box, i, j, k, s, autobar=autobar)
        assert got.tolist() == exp


def test_isolated_overshoot_closed_form():
    ni, nj, nk = 9, 8, 7
    layout = (1 - MH, ni + MH, 1 - MH, nj + MH, nk)
    shape = (nk, nj + 2 * MH, ni + 2 * MH)
    n = int(np.prod(shape))
    rng = np.random.default_rng(5)
    f = np.full(n, 1.0, np.float32)
    mn = np.full(n, 0.5, np.float32)
    mx = rng.uniform(1.2, 1.6, n).astype(np.float32)
    air = rng.uniform(0.5, 2.0, n).astype(np.float32)
    box = (1, ni, 1, nj, 1)
    i, j, k = 5, 4, 4
    nx, nxy = shape[2], shape[1] * shape[2]
    c = (i - layout[0]) + nx * (j - layout[2]) + nxy * (k - 1)
    f[c] = np.float32(1.5)
    mx[c] = np.float32(1.25)
    f0 = f.copy()
    counts = O.ilmc(f, mn, mx, air, layout, box, sweep_max=2, min_max=True)
    sh = O.ilmc_shell(layout, box, i, j, k, 1)
    room = (mx[sh].astype(np.float64) - 1.0)
    msw = float(np.sum(room * air[sh]))
    o = 0.25 * float(air[c])
    assert msw > o                                         # sweep 1 takes all of it
    assert counts.tolist() == [1, 0, 0]
    assert f[c] == pytest.approx(1.25, rel=1e-6)
    np.testing.assert_allclose(f[sh], 1.0 + room * (o / msw), rtol=1e-6)
    others = np.ones(n, bool)
    others[sh] = False
    others[c] = False
    assert np.array_equal(f[others], f0[others])


@pytest.mark.parametrize("sweep_max", [1, 2])
def test_bounds_kept_and_mass_moved_only_by_rounding(sweep_max):
    layout, shape, f, mn, mx, air = _case(30, 24, 10, 7 + sweep_max, spread=0.15)
    box = (1, 30, 1, 24, 1)
    inbox = np.zeros(shape, bool)
    inbox[:, MH:MH + 24, MH:MH + 30] = True
    inbox = inbox.reshape(-1)
    inside = (f >= mn) & (f <= mx)
    g = f.copy()
    O.ilmc(g, mn, mx, air, layout, box, sweep_max=sweep_max, min_max=False)
    assert np.array_equal(g[~inbox], f[~inbox])
    assert np.all((g[inside] >= mn[inside]) & (g[inside] <= mx[inside]))
    m0 = float(np.sum(f.astype(np.float64) * air))
    m1 = float(np.sum(g.astype(np.float64) * air))
    assert abs(m1 - m0) <= 1e-6 * m0
    viol0 = np.count_nonzero(~inside & inbox)
    viol1 = np.count_nonzero(((g < mn) | (g > mx)) & inbox)
    assert viol1 < viol0 // 2                              # most violations repaired
    h = f.copy()                                           # the same with the min/max reset
    c = O.ilmc(h, mn, mx, air, layout, box, sweep_max=sweep_max, min_max=True)
    lo, hi = (g < mn) & inbox, (g > mx) & inbox
    assert c[1] == np.count_nonzero(lo) and c[2] == np.count_nonzero(hi)
    exp = g.copy()
    exp[lo], exp[hi] = mn[lo], mx[hi]
    assert np.array_equal(h, exp)


@pytest.mark.parametrize("autobar", [False, True])
@pytest.mark.parametrize("S", [1, 2, 3, 4])
def test_hyperplane_schedule(S, autobar):
    """For every offset (di, dj, dk) with Chebyshev norm <= 2S: the later cell in the
    (k, j, i) visiting order has the larger t; equal t implies distance > 2S."""
    a = 2 * S + 1
    b = 0 if autobar else a * a
    d = 2 * S
    for di, dj, dk in itertools.product(range(-d, d + 1), repeat=3):
        if (di, dj, dk) == (0, 0, 0) or (autobar and dk != 0):
            continue
        later = (dk, dj, di) > (0, 0, 0)
        dt = di + a * dj + b * dk
        assert (dt > 0) == later and dt != 0
