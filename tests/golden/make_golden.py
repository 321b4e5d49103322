"""Write the golden vectors for the interpolation kernel from the reference's own
turbo_tricublin.c, compiled from its source by `make -C oracle ref` (oracle/_ref/).

Runs only where the reference tree is present (this container).  The fixtures are data
-- inputs and the outputs of the reference build -- and travel with the repository so
that the oracle and the HIP kernels are pinned everywhere else.

Fixtures:
  interp_small.npz   a 14x12x9 extended grid (halo 3, offsets 3/3) with three FP32 fields
                     and 600 targets covering the interior, the linear-z bottom and top
                     intervals, extrapolation below level 1 and above level nk, positions at
                     exact integers and one FP32 ulp either side, and zyx3 on the
                     interleaved fields; outputs of the AVX2+FMA and the portable builds.
  selftest_poly.npz  the reference's SELF_TEST field f=fx(x)fy(y)fz(z) (cubic, cubic,
                     linear; turbo_tricublin.c:1352-1355) on a 30x20x12 grid with unit
                     levels, its five probe positions (:1390-1466) and a bulk sweep.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
from oracle import oracle as O  # noqa: E402


def interp_small(rng):
    ni, nj, nk, h = 8, 6, 9, 3
    nit, njt = ni + 2 * h, nj + 2 * h
    levels = np.cumsum(rng.uniform(0.4, 1.6, nk)) + 2.0
    fields = [rng.standard_normal(nit * njt * nk).astype(np.float32) * s for s in (1.0, 30.0, 1e-3)]
    base = h + nit * h  # element offset of local (1,1,1)
    n = 600
    pos = np.empty((n, 3), np.float32)
    # local index space is [1-h, ni+h]; the stencil needs ix-1 >= 1-h and ix+2 <= ni+h
    lo_x, hi_x = 2 - h, ni + h - 1
    lo_y, hi_y = 2 - h, nj + h - 1
    pos[:, 0] = rng.uniform(lo_x, hi_x - 1e-3, n)
    pos[:, 1] = rng.uniform(lo_y, hi_y - 1e-3, n)
    pos[:, 2] = rng.uniform(1.0, nk, n)
    # vertical special cases
    pos[0:60, 2] = rng.uniform(1.0, 2.0, 60)          # linear bottom interval
    pos[60:120, 2] = rng.uniform(nk - 1.0, nk, 60)    # linear top interval
    pos[120:160, 2] = rng.uniform(0.0, 1.0, 40)       # extrapolation below
    pos[160:200, 2] = rng.uniform(nk, nk + 1.0, 40)   # extrapolation above
    pos[200:240, 2] = rng.integers(1, nk + 1, 40).astype(np.float32)  # exact levels
    # horizontal exact integers and +-1 ulp
    xi = rng.integers(lo_x + 1, hi_x - 1, 120).astype(np.float32)
    pos[240:280, 0] = xi[:40]
    pos[280:320, 0] = np.nextafter(xi[40:80], np.float32(-1e9))
    pos[320:360, 0] = np.nextafter(xi[80:120], np.float32(1e9))
    yi = rng.integers(lo_y + 1, hi_y - 1, 40).astype(np.float32)
    pos[360:400, 1] = np.nextafter(yi, np.float32(-1e9))
    pos[400:420, 2] = np.nextafter(rng.integers(2, nk, 20).astype(np.float32), np.float32(-1e9))
    # positions are in global index space: global = local - offset (offset = 0 here since
    # the setup passes offi/offj = h relative to an array whose (1,1,1) is at `base`)
    f3 = np.stack(fields, -1).reshape(-1).copy()
    out = dict(levels=levels, ni=nit, nj=njt, nk=nk, offi=0, offj=0, base=base,
               f0=fields[0], f1=fields[1], f2=fields[2], f3=f3, pos=pos)
    for which in ("avx2", "scalar"):
        R = O.RefKernel(which)
        lv = R.setup(levels, nit, njt, 0, 0)
        out[f"zyx1_{which}"] = R.zyx1_m_n(lv, fields, base, pos)
        out[f"mono_{which}"] = np.stack(R.mono_m_n(lv, fields, base, pos))
        out[f"zyx3_{which}"] = R.zyx3_n(lv, f3, 3 * base, pos)
    # non-zero offi/offj: the same source addressed from global positions shifted by (5,-2)
    R = O.RefKernel("avx2")
    lv = R.setup(levels, nit, njt, -5, 2)
    posg = pos.copy()
    posg[:, 0] += 5.0
    posg[:, 1] -= 2.0
    out["pos_off"] = posg
    out["zyx1_avx2_off"] = R.zyx1_m_n(lv, fields, base, posg)
    return out


def selftest_poly():
    NI, NJ, NK = 30, 20, 12

    def fx(x):
        x = np.float32(x) * np.float32(0.125)
        return x * x * x * np.float32(1.011) + x * x * np.float32(1.022) + x * np.float32(1.01) + np.float32(0.123)

    def fy(y):
        y = np.float32(y) * np.float32(0.125)
        return y * y * y * np.float32(1.044) + y * y * np.float32(1.055) + y * np.float32(1.03) + np.float32(0.234)

    def fz(z):
        return np.float32(z) * np.float32(1.123) + np.float32(0.345)

    k, j, i = np.meshgrid(np.arange(NK), np.arange(NJ), np.arange(NI), indexing="ij")
    a1 = (fx(i + 1.0) * fy(j + 1.0) * fz(k + 1.0)).astype(np.float32).reshape(-1)
    levels = np.arange(1, NK + 1, dtype=np.float64)
    II, JJ, dx, dy = NI // 2, NJ // 2, 0.555, 0.777
    probes = [(NK // 2, 0.333), (1, 0.25), (0, 0.75), (NK - 1, 0.75), (NK, 0.25)]
    pos = []
    for KK, dz in probes:
        pos.append([np.float32(dx) + II, np.float32(dy) + JJ, np.float32(dz) + KK])
    rng = np.random.default_rng(7)
    m = 400
    bulk = np.stack([rng.uniform(2, NI - 1.01, m), rng.uniform(2, NJ - 1.01, m), rng.uniform(0.5, NK + 0.5, m)], -1)
    pos = np.concatenate([np.array(pos, np.float32), bulk.astype(np.float32)])
    exact = (fx(pos[:, 0]).astype(np.float64) * fy(pos[:, 1]) * fz(pos[:, 2]))
    out = dict(levels=levels, ni=NI, nj=NJ, nk=NK, a1=a1, pos=pos, exact=exact)
    for which in ("avx2", "scalar"):
        R = O.RefKernel(which)
        lv = R.setup(levels, NI, NJ, 0, 0)
        out[f"zyx1_{which}"] = R.zyx1_m_n(lv, [a1], 0, pos)[0]
    return out


def main():
    if not (os.path.exists(O.REF_AVX2) and os.path.exists(O.REF_SCALAR)):
        raise SystemExit("oracle/_ref not built: run `make -C oracle ref` where the reference tree exists")
    rng = np.random.default_rng(20261202)
    np.savez_compressed(os.path.join(HERE, "interp_small.npz"), **interp_small(rng))
    np.savez_compressed(os.path.join(HERE, "selftest_poly.npz"), **selftest_poly())
    print("wrote", sorted(f for f in os.listdir(HERE) if f.endswith(".npz")))


if __name__ == "__main__":
    main()
