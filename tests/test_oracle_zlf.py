"""CPU: the oracle's ZLF steps (adz_BC_LAM_zlf_hlt.F90:18-319), pinned by the sizes of the
regions each step touches (closed forms of the tile geometry) and by the round trip of
steps 2 and 4 (the piloting values saved before the interpolation come back outside the
core, and only there).  nint(k0/2) rounds halves away from zero (Fortran NINT)."""
import numpy as np
import pytest

from oracle import oracle as O

MH = 2
NI, NJ, NK = 30, 24, 10
PIL = (7, 6, 5, 8)  # w, e, s, n
LAYOUT = (1 - MH, NI + MH, 1 - MH, NJ + MH, NK)
SHAPE = (NK, NJ + 2 * MH, NI + 2 * MH)
TILE = (slice(None), slice(MH, NJ + MH), slice(MH, NI + MH))


def _ones():
    return np.ones(int(np.prod(SHAPE)), np.float32)


@pytest.mark.parametrize("k0,k0d2", [(1, 1), (3, 2), (4, 2), (5, 3)])
def test_setup1_zeroes_outside_the_extension(k0, k0d2):
    ext = 3
    f = _ones()
    O.bc_lam_zlf(1, f, None, LAYOUT, (NI, NJ), PIL, ext, k0)
    f3 = f.reshape(SHAPE)
    ne = (NI - PIL[0] - PIL[1] + 2 * ext) * (NJ - PIL[2] - PIL[3] + 2 * ext)
    assert np.count_nonzero(f3[TILE] == 0) == NI * NJ * NK - ne * (NK - k0d2 + 1)
    assert np.all(f3[:, :MH, :] == 1) and np.all(f3[:, :, :MH] == 1)  # halo untouched
    assert f3[k0d2 - 1, MH + NJ // 2, MH + NI // 2] == 1
    if k0d2 > 1:
        assert f3[k0d2 - 2, MH + NJ // 2, MH + NI // 2] == 0


def test_setup3_keeps_the_bcs_box():
    f = _ones()
    O.bc_lam_zlf(3, f, None, LAYOUT, (NI, NJ), PIL, 3, 3)
    f3 = f.reshape(SHAPE)
    assert np.count_nonzero(f3[TILE] == 0) == (NI * NJ - (NI - 8) * (NJ - 8)) * NK
    assert np.all(f3[:, MH + 4:MH + NJ - 4, MH + 4:MH + NI - 4] == 1)


@pytest.mark.parametrize("k0", [1, 3])
def test_setups_2_and_4_round_trip(k0):
    rng = np.random.default_rng(2)
    dst = rng.normal(size=int(np.prod(SHAPE))).astype(np.float32)
    orig = dst.copy()
    pil = np.zeros_like(dst)
    O.bc_lam_zlf(2, dst, pil, LAYOUT, (NI, NJ), PIL, 3, k0)
    ncore = (NI - PIL[0] - PIL[1]) * (NJ - PIL[2] - PIL[3]) * (NK - k0 + 1)
    assert np.count_nonzero(pil.reshape(SHAPE)[TILE]) == NI * NJ * NK - ncore
    new = rng.normal(size=dst.size).astype(np.float32)   # the interpolation and fixer overwrite dst
    dst[:] = new
    O.bc_lam_zlf(4, dst, pil, LAYOUT, (NI, NJ), PIL, 3, k0)
    d3, o3, n3 = dst.reshape(SHAPE), orig.reshape(SHAPE), new.reshape(SHAPE)
    outside = np.zeros(SHAPE, bool)
    outside[TILE] = True
    outside[k0 - 1:, MH + PIL[2]:MH + NJ - PIL[3], MH + PIL[0]:MH + NI - PIL[1]] = False
    assert np.array_equal(d3[outside], o3[outside])
    assert np.array_equal(d3[~outside], n3[~outside])
