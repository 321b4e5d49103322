"""CPU: the oracle's LAM boundary flux (adz_BC_LAM_Aranami_hlt.F90:18-236,
adz_BC_LAM_setup.F90:18-92, mass_tr_PM.F90:101-127).

Pins:
- integer displacements: every target's departure point is a grid point, where the
  tricubic weights are (0, 1, 0, 0) along x and y and the vertical weight of the level is
  1 to within a few FP64 ulps.  FLUX_out then equals the tracer at the departure point
  when the target is outside the core and the departure point inside it, and zero
  otherwise; FLUX_in the reverse, with "outside" meaning the tile minus the core.  The
  expected arrays come from index arithmetic alone, so a shifted pilot zone, a wrong top
  level, swapped masks or a misplaced zeroing strip fails;
- the reference's own compiled Tricublin_zyx1_m_n on the same masked fields gives the same
  bits as the oracle's interpolation (skipped where oracle/_ref is not built);
- the fixer's target mass moves by exactly mass_fi - mass_fo.
"""
import os

import numpy as np
import pytest

from oracle import oracle as O

MH = 2


def _case(ni=40, nj=30, nk=12, pil=7, kc0=3, seed=5):
    layout = (1 - MH, ni + MH, 1 - MH, nj + MH, nk)
    core = (1 + pil, ni - pil, 1 + pil, nj - pil, kc0)
    bbox = (5, ni - 4, 5, nj - 4)
    levels = np.cumsum(np.linspace(1.0, 2.0, nk))
    rng = np.random.default_rng(seed)
    nx, ny = ni + 2 * MH, nj + 2 * MH
    tr = [rng.uniform(0.5, 2.0, nx * ny * nk).astype(np.float32) for _ in range(2)]
    tr[1][rng.random(tr[1].size) < 0.1] *= -1.0  # some negative values
    return dict(ni=ni, nj=nj, nk=nk, layout=layout, core=core, bbox=bbox, levels=levels, tr=tr, nx=nx, ny=ny)


def _shift_positions(C, a, b, c):
    """Departure point of target (i, j, k) at (i - a, j - b, max(k - c, 1))."""
    i0b, inb, j0b, jnb = C["bbox"]
    k, j, i = np.meshgrid(np.arange(1, C["nk"] + 1), np.arange(j0b, jnb + 1), np.arange(i0b, inb + 1),
                          indexing="ij")
    pos = np.stack([i - a, j - b, np.maximum(k - c, 1)], axis=-1).astype(np.float32)
    return pos.reshape(-1), (i, j, k)


def _expected(C, f, a, b, c):
    """FLUX_out and FLUX_in by index arithmetic (no interpolation)."""
    ni, nj, nk, nx, ny = C["ni"], C["nj"], C["nk"], C["nx"], C["ny"]
    ic0, ic1, jc0, jc1, kc0 = C["core"]
    i0b, inb, j0b, jnb = C["bbox"]
    f3 = f.reshape(nk, ny, nx)
    fo = np.zeros((nk, ny, nx), np.float32)
    fi = np.zeros((nk, ny, nx), np.float32)

    def core(i, j, k):
        return ic0 <= i <= ic1 and jc0 <= j <= jc1 and k >= kc0

    for k in range(1, nk + 1):
        for j in range(j0b, jnb + 1):
            for i in range(i0b, inb + 1):
                si, sj, sk = i - a, j - b, max(k - c, 1)
                v = f3[sk - 1, sj - 1 + MH, si - 1 + MH]
                tile = 1 <= si <= ni and 1 <= sj <= nj
                if core(i, j, k):
                    fi[k - 1, j - 1 + MH, i - 1 + MH] = v if (tile and not core(si, sj, sk)) else 0.0
                else:
                    fo[k - 1, j - 1 + MH, i - 1 + MH] = v if core(si, sj, sk) else 0.0
    return fo.reshape(-1), fi.reshape(-1)


@pytest.mark.parametrize("shift", [(2, -1, 0), (-2, 2, 1), (1, 1, 2)])
def test_integer_displacement_fluxes(shift):
    C = _case()
    zt = O.Ztab(C["levels"], C["nx"], C["ny"], 0, 0)
    pos, _ = _shift_positions(C, *shift)
    fo, fi = O.bc_lam_aranami(C["tr"], C["layout"], (C["ni"], C["nj"]), C["bbox"], C["core"], zt, pos)
    for n, f in enumerate(C["tr"]):
        eo, ei = _expected(C, f, *shift)
        assert np.array_equal(fo[n], eo), (shift, n, np.flatnonzero(fo[n] != eo)[:5])
        assert np.array_equal(fi[n], ei), (shift, n, np.flatnonzero(fi[n] != ei)[:5])
        assert np.count_nonzero(eo) > 0 and np.count_nonzero(ei) > 0


def test_flux_masses_against_direct_sums():
    C = _case()
    zt = O.Ztab(C["levels"], C["nx"], C["ny"], 0, 0)
    pos, _ = _shift_positions(C, 2, -1, 1)
    fo, fi = O.bc_lam_aranami(C["tr"], C["layout"], (C["ni"], C["nj"]), C["bbox"], C["core"], zt, pos)
    rng = np.random.default_rng(9)
    air = rng.uniform(0.5, 1.5, fo[0].size).astype(np.float32)
    mask = rng.uniform(0.8, 1.0, C["nx"] * C["ny"])
    m = O.flux_masses(fo, fi, air, mask, C["layout"], C["bbox"])
    i0b, inb, j0b, jnb = C["bbox"]
    sl = (slice(None), slice(j0b - 1 + MH, jnb + MH), slice(i0b - 1 + MH, inb + MH))
    a3, m2 = air.reshape(C["nk"], C["ny"], C["nx"]), mask.reshape(C["ny"], C["nx"])
    for n in range(len(fo)):
        for q, x in enumerate((fo[n], fi[n])):
            terms = ((x.reshape(a3.shape) * a3).astype(np.float64) * m2[None])[sl]
            assert m[2 * n + q] == pytest.approx(float(np.sum(terms)), rel=1e-13, abs=1e-300)


@pytest.mark.skipif(not os.path.exists(O.REF_AVX2), reason="oracle/_ref not built (no reference tree)")
def test_masked_interpolation_matches_reference_build():
    C = _case(seed=11)
    zt = O.Ztab(C["levels"], C["nx"], C["ny"], 0, 0)
    rng = np.random.default_rng(3)
    i0b, inb, j0b, jnb = C["bbox"]
    pos, (i, j, k) = _shift_positions(C, 0, 0, 0)
    p3 = pos.reshape(-1, 3).astype(np.float64)
    p3[:, 0] += rng.uniform(-2.4, 2.4, len(p3))
    p3[:, 1] += rng.uniform(-2.4, 2.4, len(p3))
    p3[:, 2] = np.clip(p3[:, 2] + rng.uniform(-1.5, 1.5, len(p3)), 1.0, C["nk"])
    p3[:, 0] = np.clip(p3[:, 0], 2.0, C["ni"] - 2.0)
    p3[:, 1] = np.clip(p3[:, 1], 2.0, C["nj"] - 2.0)
    pos = p3.astype(np.float32).reshape(-1)
    ref = O.RefKernel("avx2")
    lv = ref.setup(C["levels"], C["nx"], C["ny"], 0, 0)
    got = O.bc_lam_aranami(C["tr"], C["layout"], (C["ni"], C["nj"]), C["bbox"], C["core"], zt, pos)
    exp = O.bc_lam_aranami(C["tr"], C["layout"], (C["ni"], C["nj"]), C["bbox"], C["core"], zt, pos,
                           kernel=(ref, lv))
    for g, e in zip(got[0] + got[1], exp[0] + exp[1]):
        assert np.array_equal(g.view(np.uint32), e.view(np.uint32))
    assert any(np.count_nonzero(x) for x in got[1])


def test_fixer_target_moves_by_the_net_flux():
    rng = np.random.default_rng(21)
    ni, nj, nk = 24, 20, 8
    layout = (1 - MH, ni + MH, 1 - MH, nj + MH, nk)
    scope = (8, ni - 7, 8, nj - 7, 1)
    n = (ni + 2 * MH) * (nj + 2 * MH) * nk

    def tracer():
        lin = rng.uniform(1.0, 2.0, n).astype(np.float32)
        m = (lin * rng.uniform(0.97, 1.03, n)).astype(np.float32)
        return dict(p=rng.uniform(1.0, 2.0, n).astype(np.float32), m=m, lin=lin, min=np.minimum(lin, m) * 0.98,
                    max=np.maximum(lin, m) * 1.02, mass=1)

    air1 = rng.uniform(0.9, 1.1, n).astype(np.float32)
    air0 = rng.uniform(0.9, 1.1, n).astype(np.float32)
    mask = np.ones((nj + 2 * MH) * (ni + 2 * MH))
    t = [tracer() for _ in range(2)]
    flux = np.array([3.0, 7.5, 0.25, 0.0])
    _, d0 = O.bermejo_conde([dict(x, m=x["m"].copy()) for x in t], layout, scope, air1, air0, mask, np.zeros(2))
    _, d1 = O.bermejo_conde([dict(x, m=x["m"].copy()) for x in t], layout, scope, air1, air0, mask, np.zeros(2),
                            flux=flux)
    for q in range(2):
        net = flux[2 * q + 1] - flux[2 * q]
        assert d1[q, 0] == d0[q, 0] + net            # target = (mass_p - eps) + (fi - fo)
        assert abs(d1[q, 3] - d1[q, 0]) <= 1e-6 * abs(d1[q, 0])  # the fixer reaches the new target
