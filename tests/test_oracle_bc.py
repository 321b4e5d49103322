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
  * a tracer with m == lin everywhere is skipped and keeps its deficit.
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
