"""CPU: pin the oracle's interpolation restatement (oracle/oracle_interp.h).

Pins, in order of strength:
  * golden vectors written by the reference's own turbo_tricublin.c (tests/golden);
  * the compiled reference itself (oracle/_ref), where the reference tree exists;
  * mathematics: Lagrange weights sum to one and reproduce cubics, the reference's
    SELF_TEST polynomial, the index step against a brute-force floor.
"""
import os

import numpy as np
import pytest

from oracle import oracle as O


def _bits_equal(a, b):
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    return np.array_equal(a.view(np.uint8), b.view(np.uint8))


def _load_small(golden):
    g = golden("interp_small.npz")
    fields = [g["f0"], g["f1"], g["f2"]]
    zt = O.Ztab(g["levels"], int(g["ni"]), int(g["nj"]), int(g["offi"]), int(g["offj"]))
    return g, fields, zt


@pytest.mark.parametrize("which,order", [("avx2", O.ORD_AVX2), ("scalar", O.ORD_SCALAR)])
def test_golden_zyx1_mono_zyx3(golden, which, order):
    g, fields, zt = _load_small(golden)
    base = int(g["base"])
    assert _bits_equal(O.zyx1_m_n(zt, fields, base, g["pos"], order), g[f"zyx1_{which}"])
    mono = np.stack(O.mono_m_n(zt, fields, base, g["pos"], order))
    assert _bits_equal(mono, g[f"mono_{which}"])
    assert _bits_equal(O.zyx3_n(zt, g["f3"], 3 * base, g["pos"], order), g[f"zyx3_{which}"])


def test_golden_offsets(golden):
    g, fields, _ = _load_small(golden)
    zt = O.Ztab(g["levels"], int(g["ni"]), int(g["nj"]), -5, 2)
    assert _bits_equal(O.zyx1_m_n(zt, fields, int(g["base"]), g["pos_off"]), g["zyx1_avx2_off"])


def test_mono_footprints_differ_between_reference_paths(golden):
    """The reference's AVX2 min/max spans 4 y rows, the portable one 2 (documented)."""
    g, _, _ = _load_small(golden)
    a, s = g["mono_avx2"], g["mono_scalar"]
    assert np.all(a[2] <= s[2]) and np.all(a[3] >= s[3])
    assert np.any(a[2] < s[2]) and np.any(a[3] > s[3])


@pytest.mark.parametrize("which,order", [("avx2", O.ORD_AVX2), ("scalar", O.ORD_SCALAR)])
def test_selftest_polynomial(golden, which, order):
    """turbo_tricublin.c:1314-1505: cubic*cubic*linear field is reproduced exactly up to
    FP32 rounding, and the oracle matches the reference's outputs bit for bit."""
    g = golden("selftest_poly.npz")
    zt = O.Ztab(g["levels"], int(g["ni"]), int(g["nj"]))
    out = O.zyx1_m_n(zt, [g["a1"]], 0, g["pos"], order)[0]
    assert _bits_equal(out, g[f"zyx1_{which}"])
    rel = np.abs(out.astype(np.float64) / g["exact"] - 1.0)
    assert rel.max() < 1e-6


@pytest.mark.skipif(not os.path.exists(O.REF_AVX2), reason="reference build (oracle/_ref) not present")
@pytest.mark.parametrize("which,order", [("avx2", O.ORD_AVX2), ("scalar", O.ORD_SCALAR)])
def test_against_reference_build_random(which, order):
    rng = np.random.default_rng(11)
    ni, nj, nk, h = 37, 23, 17, 3
    nit, njt = ni + 2 * h, nj + 2 * h
    lev = np.cumsum(rng.uniform(0.3, 1.7, nk)) + 1.0
    fields = [rng.standard_normal(nit * njt * nk).astype(np.float32) for _ in range(4)]
    base = h + nit * h
    n = 5000
    pos = np.stack([rng.uniform(2, ni - 1.001, n), rng.uniform(2, nj - 1.001, n), rng.uniform(0.0, nk + 1.0, n)],
                   -1).astype(np.float32)
    R = O.RefKernel(which)
    lv = R.setup(lev, nit, njt, h, h)
    zt = O.Ztab(lev, nit, njt, h, h)
    assert _bits_equal(O.zyx1_m_n(zt, fields, base, pos, order), R.zyx1_m_n(lv, fields, base, pos))
    assert _bits_equal(np.stack(O.mono_m_n(zt, fields, base, pos, order)),
                       np.stack(R.mono_m_n(lv, fields, base, pos)))
    f3 = np.stack(fields[:3], -1).reshape(-1).copy()
    assert _bits_equal(O.zyx3_n(zt, f3, 3 * base, pos, order), R.zyx3_n(lv, f3, 3 * base, pos))


def test_weights_partition_of_unity_and_cubic_reproduction():
    rng = np.random.default_rng(3)
    nk = 12
    lev = np.cumsum(rng.uniform(0.2, 2.0, nk))
    zt = O.Ztab(lev, 20, 20)
    n = 2000
    pos = np.stack([rng.uniform(2, 18, n), rng.uniform(2, 18, n), rng.uniform(0.5, nk + 0.5, n)], -1)
    ijk, zl, c = O.vcoef(zt, pos)
    for s in (slice(0, 4), slice(4, 8), slice(8, 12), slice(12, 16), slice(16, 18), slice(18, 20)):
        assert np.allclose(c[:, s].sum(1), 1.0, atol=1e-12, rtol=0)
    # x weights reproduce cubics in the fractional offset (nodes -1, 0, 1, 2)
    dx = pos[:, 0] - np.trunc(pos[:, 0])
    nodes = np.array([-1.0, 0.0, 1.0, 2.0])
    for p in range(4):
        assert np.allclose((c[:, 0:4] * nodes ** p).sum(1), dx ** p, atol=1e-12)
    # cubic z weights reproduce cubics of the absolute level on non-uniform levels
    cub = zl == 0
    iz = np.clip(np.trunc(pos[:, 2]).astype(int), 1, nk - 1)
    dz = pos[:, 2] - iz
    zabs = dz * lev[iz] + (1 - dz) * lev[iz - 1]  # 0-based: lev[iz-1] is level iz
    izc = iz[cub]
    stencil = np.stack([lev[izc - 2 + q] for q in range(4)], -1)  # 0-based levels iz-2..iz+1
    for p in range(4):
        lhs = (c[cub, 8:12] * stencil ** p).sum(1)
        assert np.allclose(lhs, zabs[cub] ** p, rtol=1e-9, atol=1e-9)
    # linear-z flag exactly on the first and last intervals
    assert np.array_equal(zl != 0, (iz - 1 < 1) | (iz - 1 > nk - 3))


def test_index_step_bruteforce_faces():
    """Offsets of positions at exact integers and one FP32 ulp either side match a
    brute-force floor (truncation == floor on the positive, clamped domain)."""
    nk, ni, nj = 10, 30, 25
    zt = O.Ztab(np.arange(1.0, nk + 1.0), ni, nj)
    xs = np.arange(3, ni - 2).astype(np.float32)
    cand = np.concatenate([xs, np.nextafter(xs, np.float32(0)), np.nextafter(xs, np.float32(1e9))])
    n = cand.size
    pos = np.stack([cand, np.full(n, 5.5, np.float32), np.full(n, 4.25, np.float32)], -1)
    ijk, zl, _ = O.vcoef(zt, pos)
    ix = np.floor(cand.astype(np.float64)).astype(int)
    iz = 4
    expect = (ix - 2) + (5 - 2) * ni + (iz - 1) * ni * nj - ni * nj  # cubic: one plane down
    assert np.array_equal(ijk, expect)
    assert not zl.any()


def test_ztab_denominators():
    lev = np.array([1.0, 1.5, 2.7, 3.1, 4.9, 6.0, 7.7])
    z, ocz = O.Ztab(lev, 8, 8).tables()
    assert np.array_equal(z, lev)
    ocz = ocz.reshape(-1, 4)

    def den(a, b, c, d):
        return np.array([1 / ((a - b) * (a - c) * (a - d)), 1 / ((b - a) * (b - c) * (b - d)),
                         1 / ((c - a) * (c - b) * (c - d)), 1 / ((d - a) * (d - b) * (d - c))])

    nk = len(lev)
    assert np.allclose(ocz[0], den(*lev[0:4]), rtol=1e-15)
    for i in range(1, nk - 2):
        assert np.allclose(ocz[i], den(*lev[i - 1:i + 3]), rtol=1e-15)
    assert np.allclose(ocz[nk - 2], den(*lev[nk - 4:nk]), rtol=1e-15)
    assert np.allclose(ocz[nk - 1], den(*lev[nk - 4:nk]), rtol=1e-15)
