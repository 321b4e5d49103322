"""GPU parity of the interpolation entry points (adz_tricublin_*, adz_vcoef) against the
oracle and the reference's golden vectors.

Bar: bit-exact for indices and for ADZ_MATH_REF (FP32 fields: the reference's own output
bits; FP64 fields: the oracle's FP64 restatement in the same operation order).  The FP32
arithmetic form is held to |err| <= 5e-6 * max|f| against the FP64 oracle.
"""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _bits_equal(a, b):
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    return a.shape == b.shape and np.array_equal(a.view(np.uint8), b.view(np.uint8))


@pytest.fixture(scope="module")
def T(cuda):
    import torch

    from gem_amd import _lib as L
    from gem_amd import tricublin as TC

    return torch, L, TC


def _dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def test_golden_vectors_on_gpu(T, golden):
    torch, L, TC = T
    g = golden("interp_small.npz")
    ni, nj, nk, base = int(g["ni"]), int(g["nj"]), int(g["nk"]), int(g["base"])
    h = 3
    lv = TC.Ztab(g["levels"], ni, nj, 0, 0).set_bounds(1 - h, ni - h, 1 - h, nj - h)
    fs = [_dev(torch, g[k]) for k in ("f0", "f1", "f2")]
    pos = _dev(torch, g["pos"])
    n = pos.numel() // 3
    d = torch.empty(3 * n, dtype=torch.float32, device="cuda")
    TC.Tricublin_zyx1_m_n(d, fs, pos, lv, n, 3, base_off=base)
    assert _bits_equal(d.cpu().numpy().reshape(3, n), g["zyx1_avx2"])
    outs = [torch.empty(3 * n, dtype=torch.float32, device="cuda") for _ in range(4)]
    TC.Tricublin_mono_zyx_m_n(*outs, fs, pos, lv, n, 3, base_off=base)
    assert _bits_equal(np.stack([o.cpu().numpy().reshape(3, n) for o in outs]), g["mono_avx2"])
    d3 = torch.empty(3 * n, dtype=torch.float32, device="cuda")
    TC.Tricublin_zyx3_n(d3, _dev(torch, g["f3"]), pos, lv, n, base_off=3 * base)
    assert _bits_equal(d3.cpu().numpy().reshape(n, 3), g["zyx3_avx2"])
    lv2 = TC.Ztab(g["levels"], ni, nj, -5, 2).set_bounds(1 - h, ni - h, 1 - h, nj - h)
    posg = _dev(torch, g["pos_off"])
    TC.Tricublin_zyx1_m_n(d, fs, posg, lv2, n, 3, base_off=base)
    assert _bits_equal(d.cpu().numpy().reshape(3, n), g["zyx1_avx2_off"])


def _case(seed, ni=61, nj=47, nk=23, h=4, n=40000, fdt=np.float32, pdt=np.float32):
    rng = np.random.default_rng(seed)
    nit, njt = ni + 2 * h, nj + 2 * h
    lev = np.cumsum(rng.uniform(0.3, 1.7, nk)) + 1.0
    # smooth fields plus noise, values O(1..100)
    x = np.arange(nit)[None, None, :]
    y = np.arange(njt)[None, :, None]
    z = np.arange(nk)[:, None, None]
    fields = []
    for q in range(5):
        f = (50 * np.sin(0.11 * (q + 1) * x + 0.07 * y + 0.2 * z + q) + 3 * rng.standard_normal((nk, njt, nit)))
        fields.append(f.reshape(-1).astype(fdt))
    base = h + nit * h  # f(1,1,1) = element (h, h, 0); local index range [1-h, ni+h]
    # stencil ix-1..ix+2 must stay in [1-h, ni+h]: ix in [2-h, ni+h-2]
    pos = np.stack([rng.uniform(3 - h, ni + h - 2.0001, n), rng.uniform(3 - h, nj + h - 2.0001, n),
                    rng.uniform(0.0, nk + 1.0, n)], -1).astype(pdt)
    pos[: n // 20, 2] = rng.integers(1, nk + 1, n // 20)  # exact levels
    return lev, nit, njt, nk, h, base, fields, pos


@pytest.mark.parametrize("fdt,pdt", [(np.float32, np.float32), (np.float32, np.float64),
                                     (np.float64, np.float32), (np.float64, np.float64)])
def test_ref_math_bit_exact_vs_oracle(T, fdt, pdt):
    torch, L, TC = T
    lev, nit, njt, nk, h, base, fields, pos = _case(5, fdt=fdt, pdt=pdt)
    n, m = pos.shape[0], len(fields)
    zt = O.Ztab(lev, nit, njt, 0, 0)
    lv = TC.Ztab(lev, nit, njt, 0, 0).set_bounds(1 - h, nit - h, 1 - h, njt - h)
    fs = [_dev(torch, f) for f in fields]
    dpos = _dev(torch, pos)
    tdt = torch.float32 if fdt == np.float32 else torch.float64
    d = torch.empty(m * n, dtype=tdt, device="cuda")
    TC.Tricublin_zyx1_m_n(d, fs, dpos, lv, n, m, base_off=base)
    ref = O.zyx1_m_n(zt, fields, base, pos, O.ORD_AVX2)
    assert _bits_equal(d.cpu().numpy().reshape(m, n), ref)
    outs = [torch.empty(m * n, dtype=tdt, device="cuda") for _ in range(4)]
    TC.Tricublin_mono_zyx_m_n(*outs, fs, dpos, lv, n, m, base_off=base)
    refm = O.mono_m_n(zt, fields, base, pos, O.ORD_AVX2)
    for a, b in zip(outs, refm):
        assert _bits_equal(a.cpu().numpy().reshape(m, n), b)
    f3 = np.stack(fields[:3], -1).reshape(-1).copy()
    d3 = torch.empty(3 * n, dtype=tdt, device="cuda")
    TC.Tricublin_zyx3_n(d3, _dev(torch, f3), dpos, lv, n, base_off=3 * base)
    assert _bits_equal(d3.cpu().numpy().reshape(n, 3), O.zyx3_n(zt, f3, 3 * base, pos, O.ORD_AVX2))


@pytest.mark.parametrize("pdt", [np.float32, np.float64])
def test_indices_and_weights_bit_exact(T, pdt):
    torch, L, TC = T
    lev, nit, njt, nk, h, base, fields, pos = _case(9, pdt=pdt)
    zt = O.Ztab(lev, nit, njt, 0, 0)
    lv = TC.Ztab(lev, nit, njt, 0, 0).set_bounds(1 - h, nit - h, 1 - h, njt - h)
    ijk, zl, coef = TC.vcoef(_dev(torch, pos), lv)
    rijk, rzl, rcoef = O.vcoef(zt, pos, zabs_fma=1)
    assert np.array_equal(ijk.cpu().numpy(), rijk)
    assert np.array_equal(zl.cpu().numpy() != 0, rzl != 0)
    assert _bits_equal(coef.cpu().numpy(), rcoef)


@pytest.mark.parametrize("sdt", ["f32", "bf16"])
def test_f32_math_against_fp64_oracle(T, sdt):
    torch, L, TC = T
    lev, nit, njt, nk, h, base, fields64, pos = _case(13, fdt=np.float64, pdt=np.float64)
    n, m = pos.shape[0], len(fields64)
    zt = O.Ztab(lev, nit, njt, 0, 0)
    ref = O.zyx1_m_n(zt, fields64, base, pos, O.ORD_AVX2)
    lv = TC.Ztab(lev, nit, njt, 0, 0).set_bounds(1 - h, nit - h, 1 - h, njt - h)
    tdt = torch.float32 if sdt == "f32" else torch.bfloat16
    fs = [torch.from_numpy(f).cuda().to(tdt) for f in fields64]
    d = torch.empty(m * n, dtype=tdt, device="cuda")
    TC.Tricublin_zyx1_m_n(d, fs, _dev(torch, pos), lv, n, m, base_off=base, math=L.ADZ_MATH_F32)
    got = d.float().cpu().numpy().astype(np.float64).reshape(m, n)
    scale = np.abs(np.stack(fields64)).max()
    err = np.abs(got - ref).max()
    tol = 5e-6 if sdt == "f32" else 1.5e-2
    assert err <= tol * scale, (sdt, err / scale)
    # the linear / min / max outputs: min and max are exact copies of stored values
    outs = [torch.empty(m * n, dtype=tdt, device="cuda") for _ in range(4)]
    TC.Tricublin_mono_zyx_m_n(*outs, fs, _dev(torch, pos), lv, n, m, base_off=base, math=L.ADZ_MATH_F32)
    stored = [f.cpu().float().numpy().astype(np.float64) for f in fs]
    _, _, rmi, rma = O.mono_m_n(zt, stored, base, pos, O.ORD_AVX2)
    assert np.array_equal(outs[2].float().cpu().numpy().reshape(m, n), rmi.astype(np.float32))
    assert np.array_equal(outs[3].float().cpu().numpy().reshape(m, n), rma.astype(np.float32))


def test_out_of_range_positions_give_nan_not_faults(T):
    torch, L, TC = T
    lev, nit, njt, nk, h, base, fields, pos = _case(17, n=4096)
    lv = TC.Ztab(lev, nit, njt, 0, 0).set_bounds(1 - h, nit - h, 1 - h, njt - h)
    bad = pos.copy()
    bad[0::4, 0] = -1.0e6        # far left
    bad[1::4, 1] = njt + 50.0    # far north
    bad[2::4, 0] = np.nan
    fs = [_dev(torch, f) for f in fields]
    n, m = bad.shape[0], len(fields)
    d = torch.empty(m * n, dtype=torch.float32, device="cuda")
    TC.Tricublin_zyx1_m_n(d, fs, _dev(torch, bad), lv, n, m, base_off=base)
    out = d.cpu().numpy().reshape(m, n)
    assert np.isnan(out[:, 0::4]).all() and np.isnan(out[:, 1::4]).all() and np.isnan(out[:, 2::4]).all()
    assert np.isfinite(out[:, 3::4]).all()
