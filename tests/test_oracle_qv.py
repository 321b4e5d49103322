"""CPU: the oracle's bicubic-horizontal / quintic-vertical interpolation
(bicubHQV_lin_mono.F90:20-172, bicubHQV_lin_min_max_zyx.F90:20-99, tables of
adz_set.F90:181-190,286-307).

Pins:
- the evaluation equals the reference's own bicubHQV_lin_min_max_zyx, compiled from its
  source with amdflang (oracle/_ref), bit for bit, in the quintic, cubic and linear
  vertical modes (skipped where oracle/_ref is not built);
- the weights reproduce polynomials: cubic in x and y, quintic in z on non-uniform levels
  in the quintic mode, cubic in z next to the boundaries, linear in the end intervals;
  the linear result reproduces trilinear functions;
- min / max equal a brute-force search over the inner 2x2x2 box located by index
  arithmetic alone.
"""
import os

import numpy as np
import pytest

from oracle import oracle as O

H = 3
NI, NJ, NK = 20, 16, 12
LAYOUT = (1 - H, NI + H, 1 - H, NJ + H)
NX, NY = NI + 2 * H, NJ + 2 * H
IOFF = JOFF = -1  # l_i0 = l_j0 = 1


def _levels():
    # Ver_z_8%t(0:nk+1), non-uniform
    z = np.cumsum(np.concatenate([[0.0], np.linspace(0.6, 1.7, NK + 1)]))
    return z


def _positions(rng, n, kk_range=None):
    x = rng.uniform(2.0, NI - 2.0, n)
    y = rng.uniform(2.0, NJ - 2.0, n)
    if kk_range is None:
        z = rng.uniform(1.0, NK - 0.001, n)
    else:
        z = rng.integers(kk_range[0], kk_range[1] + 1, n) + rng.uniform(0.0, 0.999, n)
    return np.stack([x, y, z], axis=-1).astype(np.float32)


def _field(fn):
    k, j, i = np.meshgrid(np.arange(1, NK + 1), np.arange(LAYOUT[2], LAYOUT[3] + 1),
                          np.arange(LAYOUT[0], LAYOUT[1] + 1), indexing="ij")
    zt = _levels()
    return fn(i.astype(np.float64), j.astype(np.float64), zt[k]).astype(np.float32).reshape(-1)


def _zt_at(tab_z, pz):
    kk1 = int(pz)
    quint = 3 <= kk1 <= NK - 3
    cub = kk1 in (2, NK - 2)
    if not quint and not cub:
        kk1 = min(NK - 1, max(1, kk1))
    dz = tab_z[kk1 + 1] - tab_z[kk1]
    return tab_z[kk1] + (float(pz) - kk1) * dz, kk1, quint, cub


@pytest.mark.skipif(not os.path.exists(O.REF_QVZYX), reason="oracle/_ref not built (no reference tree)")
def test_evaluation_matches_reference_build():
    rng = np.random.default_rng(4)
    T = O.QvTab(_levels(), NK)
    ref = O.RefQvZyx()
    src = rng.normal(0.0, 1.0, NX * NY * NK).astype(np.float32)
    modes = set()
    for p in _positions(rng, 600):
        w, zq, zc, off = O.qv_coef(T, p, LAYOUT, IOFF, JOFF)
        modes.add((zq, zc))
        a = O.qv_zyx(src, off, w, NX, NX * NY, zq, zc)
        b = ref(src, off, w, NX, NX * NY, zq, zc)
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), (p, a, b)
    assert modes == {(True, False), (False, True), (False, False)}


@pytest.mark.parametrize("mode,kk,deg", [("quintic", (3, NK - 3), 5), ("cubic", (2, 2), 3), ("cubic", (NK - 2, NK - 2), 3),
                                         ("linear", (1, 1), 1), ("linear", (NK - 1, NK - 1), 1)])
def test_polynomial_reproduction(mode, kk, deg):
    rng = np.random.default_rng(7 + deg)
    T = O.QvTab(_levels(), NK)
    cz = rng.uniform(-1, 1, deg + 1)
    zc = 0.5 * (_levels()[1] + _levels()[NK])
    gx = lambda x: 1.5 + 0.3 * x - 0.02 * x ** 2 + 0.001 * x ** 3  # noqa: E731
    gy = lambda y: -0.7 + 0.2 * y + 0.03 * y ** 2 - 0.002 * y ** 3  # noqa: E731
    gz = lambda z: sum(c * ((z - zc) / 10.0) ** e for e, c in enumerate(cz))  # noqa: E731
    f = _field(lambda x, y, z: gx(x) + gy(y) + gz(z))
    pos = _positions(rng, 300, kk)
    qut, lin, mi, ma = O.qv_m_n(T, [f], LAYOUT, pos, IOFF, JOFF)
    zt = np.array([_zt_at(_levels(), p[2])[0] for p in pos])
    exp = gx(pos[:, 0].astype(np.float64)) + gy(pos[:, 1].astype(np.float64)) + gz(zt)
    assert np.abs(qut[0] - exp).max() <= 2e-5 * np.abs(f).max(), mode
    # the linear result reproduces a trilinear function
    fl = _field(lambda x, y, z: 0.3 + 0.05 * x - 0.04 * y + 0.1 * z)
    _, lin, _, _ = O.qv_m_n(T, [fl], LAYOUT, pos, IOFF, JOFF)
    expl = 0.3 + 0.05 * pos[:, 0] - 0.04 * pos[:, 1] + 0.1 * zt
    assert np.abs(lin[0] - expl).max() <= 2e-5 * np.abs(fl).max(), mode


def test_min_max_brute_force():
    rng = np.random.default_rng(9)
    T = O.QvTab(_levels(), NK)
    f = rng.normal(0.0, 1.0, NX * NY * NK).astype(np.float32)
    f3 = f.reshape(NK, NY, NX)
    pos = _positions(rng, 400)
    _, _, mi, ma = O.qv_m_n(T, [f], LAYOUT, pos, IOFF, JOFF)
    for n, p in enumerate(pos):
        i, j = int(p[0]), int(p[1])
        _, kk1, _, _ = _zt_at(_levels(), p[2])
        box = f3[kk1 - 1:kk1 + 1, j - LAYOUT[2]:j - LAYOUT[2] + 2, i - LAYOUT[0]:i - LAYOUT[0] + 2]
        assert mi[0, n] == box.min() and ma[0, n] == box.max(), n


def test_grid_points_are_reproduced():
    rng = np.random.default_rng(11)
    T = O.QvTab(_levels(), NK)
    f = rng.uniform(0.5, 2.0, NX * NY * NK).astype(np.float32)
    f3 = f.reshape(NK, NY, NX)
    pts = np.array([[i, j, k] for i in range(3, NI - 2, 3) for j in range(3, NJ - 2, 3) for k in range(1, NK)],
                   np.float32)
    qut, lin, _, _ = O.qv_m_n(T, [f], LAYOUT, pts, IOFF, JOFF)
    exp = np.array([f3[int(k) - 1, int(j) - LAYOUT[2], int(i) - LAYOUT[0]] for i, j, k in pts])
    assert np.array_equal(qut[0], exp)
    assert np.array_equal(lin[0], exp)
