"""CPU: pin the oracle's Bermejo-Conde restatement and its FP64 sums.

Bermejo_Conde.F90 cannot be built here (it needs MPI and GEM's modules), so the oracle is
pinned by properties a wrong term or sign would break:
  * almost_zero: the bit-pattern rule of almost_zero.F90 on hand-picked values;
  * exactly rounded sums: equal to Python's math.fsum on cancelling data;
  * without the min/max reset the additive correction restores the target mass exactly
    (M_m - lambda*W = M_target), and leaves untouched every point whose m - lin has the
    opposite sign of the deficit;
  * with the reset, the proportional fix restores the target mass;
  * the carried deficit identity (new epsilon = final mass - target);
  * a tracer with m == lin everywhere is skipped and keeps its deficit;
  * the shape of the correction: without the reset it is lambda * w, so corrections of
    two cells stand in the ratio of |m - lin|^p * factor (weights 1, 2, 3; p = 1, 2, 3),
    which a wrong exponent, a dropped max(0, m) or pr_k/pr_s, or a reversed sign rule
    breaks (each planted against the same check).
"""
import math

import numpy as np
import pytest

from oracle import oracle as O


def test_almost_zero_bit_rule():
    tiny = np.nextafter(0.0, 1.0)
    assert O.almost_zero(0.0) and O.almost_zero(tiny)
    assert not O.almost_zero(np.nextafter(tiny, 1.0))
    assert not O.almost_zero(-0.0) and not O.almost_zero(-tiny)
    assert not O.almost_zero(1e-300) and not O.almost_zero(-1.0)


def test_fsum_is_exactly_rounded():
    rng = np.random.default_rng(5)
    for _ in range(20):
        v = np.concatenate([rng.standard_normal(1000) * 10.0 ** rng.integers(-8, 9, 1000), [1e16, 1.0, -1e16]])
        rng.shuffle(v)
        assert O.fsum(v) == math.fsum(v)
    assert math.copysign(1.0, O.fsum(np.array([-0.0, -0.0]))) == 1.0  # starts from +0.0


def _case(rng, dt, nx=20, ny=16, nk=6, halo=2, ntr=3):
    layout = (1 - halo, nx + halo, 1 - halo, ny + halo, nk)
    scope = (3, nx - 2, 3, ny - 2, 1)
    shape = (nk, ny + 2 * halo, nx + 2 * halo)
    air0 = rng.uniform(1e8, 3e8, shape).astype(dt)
    air1 = (air0 * rng.uniform(0.99, 1.01, shape)).astype(dt)
    mask = np.ones((ny + 2 * halo, nx + 2 * halo))
    trs = []
    for n in range(ntr):
        p = (1e-6 * rng.lognormal(0, 1, shape)).astype(dt)
        lin = (p * rng.uniform(0.9, 1.1, shape)).astype(dt)
        m = (lin * rng.uniform(0.95, 1.07, shape)).astype(dt)
        mn = np.minimum(lin, m) * dt(0.97)
        mx = np.maximum(lin, m) * dt(1.03)
        trs.append(dict(p=p, m=m, lin=lin, min=mn.astype(dt), max=mx.astype(dt), mass=1))
    return layout, scope, air1, air0, mask, trs


def _mass(x, air, mask, layout, scope):
    return O.mass(x, air, mask, layout, scope)


@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_additive_correction_restores_target_mass(dt):
    rng = np.random.default_rng(7)
    layout, scope, a1, a0, mask, trs = _case(rng, dt)
    before = [t["m"].copy() for t in trs]
    de, diag = O.bermejo_conde(trs, layout, scope, a1, a0, mask, np.zeros(len(trs)), min_max=False)
    for n, t in enumerate(trs):
        target = diag[n, 0]
        got = _mass(t["m"], a0, mask, layout, scope)
        tol = 1e-6 if dt == np.float32 else 1e-13
        assert abs(got - target) <= tol * abs(target)
        assert de[n] == got - target                       # carried deficit identity
        D = diag[n, 1] - target
        s = np.sign(before[n].astype(np.float64) - t["lin"])
        opposite = (s != 0) & (s != np.sign(D))
        assert np.array_equal(t["m"][opposite], before[n][opposite])


@pytest.mark.parametrize("mass_code", [1, 112, 121, 131])
@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_min_max_then_proportional_restores_target(dt, mass_code):
    rng = np.random.default_rng(11)
    layout, scope, a1, a0, mask, trs = _case(rng, dt)
    for t in trs:
        t["mass"] = mass_code
    pkps = rng.uniform(0.1, 1.0, a0.shape).astype(dt)
    de, diag = O.bermejo_conde(trs, layout, scope, a1, a0, mask, np.full(len(trs), 1e-9), pkps=pkps,
                               pexp_list=(1.0, 2.0) + (1.0,) * 7)
    for n, t in enumerate(trs):
        target = diag[n, 0]
        got = _mass(t["m"], a0, mask, layout, scope)
        tol = 1e-6 if dt == np.float32 else 1e-13
        assert abs(got - target) <= tol * abs(target)
        assert de[n] == got - target
        assert np.isclose(target, _mass(t["p"], a1, mask, layout, scope) - 1e-9, rtol=1e-15)


def test_skip_when_weights_vanish():
    rng = np.random.default_rng(13)
    layout, scope, a1, a0, mask, trs = _case(rng, np.float32, ntr=1)
    trs[0]["m"] = trs[0]["lin"].copy()
    before = trs[0]["m"].copy()
    de, diag = O.bermejo_conde(trs, layout, scope, a1, a0, mask, np.array([3e-7]))
    assert diag[0, 4] == 1.0 and de[0] == 3e-7
    assert np.array_equal(trs[0]["m"], before)


def test_sequential_and_exact_sums_agree_closely():
    rng = np.random.default_rng(17)
    layout, scope, a1, a0, mask, trs = _case(rng, np.float32, ntr=1)
    x = trs[0]["m"]
    s0 = O.mass(x, a0, mask, layout, scope, sum_mode=0)
    s1 = O.mass(x, a0, mask, layout, scope, sum_mode=1)
    assert abs(s0 - s1) <= 1e-14 * abs(s1)


# ---------------------------------------------------------------- weight shape
def _expected_weights(m, lin, D, p, weight, pkps):
    """|m - lin|^p sign(D) times the weight kind's factor where sign(m - lin) = sign(D),
    zero elsewhere (Bermejo_Conde.F90:112-172), from the formula's meaning."""
    h = m.astype(np.float64) - lin.astype(np.float64)
    same = np.sign(h) == np.sign(D)
    w = np.where(same, np.abs(h) ** p * np.sign(D), 0.0)
    if weight == 2:
        w = w * np.maximum(0.0, m)
    if weight == 3:
        w = w * pkps
    return w


def _shape_matches(corr, w, scale):
    """The correction is lambda * w for one lambda (fitted on the largest weight): zero
    where w is zero, and within a few ulp of the stored value (`scale`, the larger of m
    before and after) of lambda * w elsewhere -- corr = m - (m - lambda w) cancels, so
    small corrections carry the rounding of m itself."""
    nz = w != 0.0
    if np.any(corr[~nz] != 0.0) or nz.sum() < 10:
        return False
    q = np.argmax(np.abs(w))
    lam = corr.flat[q] / w.flat[q]
    err = np.abs(corr - lam * w)
    return bool(np.all(err <= 4 * np.finfo(np.float64).eps * scale + 1e-12 * np.abs(lam * w)))


@pytest.mark.parametrize("weight", [1, 2, 3])
@pytest.mark.parametrize("pexp_n", [1, 2, 3])
@pytest.mark.parametrize("deficit_sign", [1, -1])
def test_correction_shape_is_the_weight(weight, pexp_n, deficit_sign):
    """Without the min/max reset the fixer moves m by lambda * w: the ratio of the
    corrections of two cells is the ratio of their weights |m - lin|^p * factor
    (factor 1, max(0, m) or pr_k/pr_s).  Checked in FP64 storage, where the stored
    weight and correction are not rounded to FP32.  A wrong exponent, a missing
    max(0, m) or a missing pr_k/pr_s fails the check (each planted below)."""
    rng = np.random.default_rng(100 + 10 * weight + pexp_n)
    layout, scope, a1, a0, mask, trs = _case(rng, np.float64, ntr=1)
    t = trs[0]
    # both signs of m - lin, and some negative m (weight 2 keeps those untouched)
    t["m"] = t["lin"] * rng.uniform(0.9, 1.1, t["lin"].shape)
    t["m"][rng.random(t["m"].shape) < 0.1] *= -1.0
    # a carried deficit that fixes the sign of D = M_m - (M_p - eps)
    mp = _mass(t["p"], a1, mask, layout, scope)
    mm = _mass(t["m"], a0, mask, layout, scope)
    eps = (mp - mm) + deficit_sign * 0.05 * abs(mm)
    t["mass"] = 100 + 10 * weight + pexp_n
    pkps = rng.uniform(0.1, 1.0, a0.shape)
    before = t["m"].copy()
    de, diag = O.bermejo_conde(trs, layout, scope, a1, a0, mask, np.array([eps]), pkps=pkps, min_max=False,
                               pexp_list=(1.0, 2.0, 3.0) + (1.0,) * 6)
    D = diag[0, 1] - diag[0, 0]
    assert np.sign(D) == deficit_sign and diag[0, 4] == 0.0
    i0, in_, j0, jn, k0 = scope
    minx, maxx, miny, maxy, nk = layout
    sl = (slice(k0 - 1, nk), slice(j0 - miny, jn - miny + 1), slice(i0 - minx, in_ - minx + 1))
    corr = (before - t["m"])[sl]
    scale = np.maximum(np.abs(before), np.abs(t["m"]))[sl]
    w = _expected_weights(before, t["lin"], D, float(pexp_n), weight, pkps)[sl]
    assert _shape_matches(corr, w, scale)
    # lambda = D / sum(w * air0 * mask) (Bermejo_Conde.F90:200)
    q = np.argmax(np.abs(w))
    lam = corr.flat[q] / w.flat[q]
    W = math.fsum((w * a0[sl]).ravel())
    assert abs(lam - D / W) <= 1e-9 * abs(D / W)
    # planted mistakes fail the same check
    assert not _shape_matches(corr, _expected_weights(before, t["lin"], D, float(pexp_n) + 1.0, weight, pkps)[sl], scale)
    if weight == 2 and deficit_sign < 0:  # cells with m < 0 < lin are selected only when D < 0
        planted = _expected_weights(before, t["lin"], D, float(pexp_n), 1, pkps) * before
        assert not _shape_matches(corr, planted[sl], scale)   # max(0, m) dropped
    if weight == 3:
        assert not _shape_matches(corr, _expected_weights(before, t["lin"], D, float(pexp_n), 1, pkps)[sl], scale)
    if weight == 1:
        flipped = _expected_weights(before, t["lin"], -D, float(pexp_n), 1, pkps)
        assert not _shape_matches(corr, flipped[sl], scale)   # sign rule reversed


@pytest.mark.parametrize("min_max", [False, True])
def test_autobar_masses_are_level_one_area_sums(min_max):
    """Schm_autobar_L (mass_tr_PM.F90:79-85, mass_tr_X.F90:73-79): every mass is the sum over
    level 1 of tracer * geomh_area_mask_8, with no air mass.  Checked by its consequences:
    the air masses and the levels above 1 do not change any mass (scrambling them leaves
    the diagnostics unchanged), the fixer restores the target mass measured that way, the
    carried-deficit identity holds, and the target is the level-1 area sum of p minus the
    deficit (a scope or weight taken from the 3-D form fails these)."""
    rng = np.random.default_rng(29)
    layout, scope, a1, a0, mask, trs = _case(rng, np.float64)
    area = rng.uniform(1e6, 4e6, mask.shape)
    base = [{k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in t.items()} for t in trs]
    de, diag = O.bermejo_conde(trs, layout, scope, a1, a0, area, np.full(len(trs), 2e-9), min_max=min_max,
                               autobar=True)
    nk = layout[4]
    level1 = lambda x: x.reshape(nk, -1)[0].reshape(mask.shape)
    i0, in_, j0, jn, _ = scope
    sl = (slice(j0 - layout[2], jn - layout[2] + 1), slice(i0 - layout[0], in_ - layout[0] + 1))
    for n, t in enumerate(trs):
        target = diag[n, 0]
        assert target == O.fsum((level1(t["p"]) * area)[sl].ravel()) - 2e-9
        got = O.fsum((level1(t["m"]) * area)[sl].ravel())
        assert abs(got - target) <= 1e-13 * abs(target)
        assert de[n] == got - target
    # air masses and the upper levels' values of p do not enter any mass
    trs2 = [{k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in t.items()} for t in base]
    for t in trs2:
        t["p"].reshape(nk, -1)[1:] *= 3.0
    de2, diag2 = O.bermejo_conde(trs2, layout, scope, a1 * 2.0, a0 * 0.5, area, np.full(len(trs), 2e-9),
                                 min_max=min_max, autobar=True)
    assert np.array_equal(diag2[:, 0], diag[:, 0]) and np.array_equal(diag2[:, 1], diag[:, 1])
    # the 3-D form of the same inputs gives other masses
    _, diag3 = O.bermejo_conde([{k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in t.items()}
                                for t in base], layout, scope, a1, a0, area, np.full(len(trs), 2e-9),
                               min_max=min_max)
    assert not np.array_equal(diag3[:, 0], diag[:, 0])
