"""GPU: adz_bermejo_conde_static, the fixer that allocates nothing (its tracer table,
partials and scalars in a fixed device array), against the fixer's plan entry points
(adz_bc_create + adz_bermejo_conde / adz_bermejo_conde_lam) on identical copies of the
inputs, through the C ABI: corrected tracers, carried deficits and diagnostics bit for bit,
on two calls in a row, the second with fewer tracers over a smaller domain (the array keeps
the first call's contents), with and without the LAM flux masses.  Also: more than 64
tracers are refused before any device work."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def T(cuda):
    import torch

    from gem_amd import _lib as L
    from gem_amd import adz as A

    return torch, L, A


def _case(rng, nx, ny, nk, ntr, halo=2):
    layout = (1 - halo, nx + halo, 1 - halo, ny + halo, nk)
    scope = (4, nx - 3, 4, ny - 3, 2)
    shape = (nk, ny + 2 * halo, nx + 2 * halo)
    air0 = rng.uniform(1e8, 3e8, shape).astype(np.float32)
    air1 = (air0 * rng.uniform(0.99, 1.01, shape)).astype(np.float32)
    mask = np.ones((ny + 2 * halo, nx + 2 * halo))
    mask[:, : halo + 6] = 0.5
    pkps = rng.uniform(0.1, 1.0, shape).astype(np.float32)
    codes = [1, 112, 121, 131, 111, 122, 1, 113][:ntr]
    trs = []
    for c in codes:
        p = (1e-6 * rng.lognormal(0, 1, shape)).astype(np.float32)
        lin = (p * rng.uniform(0.9, 1.1, shape)).astype(np.float32)
        m = (lin * rng.uniform(0.95, 1.07, shape)).astype(np.float32)
        trs.append(dict(p=p, m=m, lin=lin, min=np.minimum(lin, m) * np.float32(0.97),
                        max=np.maximum(lin, m) * np.float32(1.03), mass=c))
    return layout, scope, air1, air0, pkps, mask, trs


def _static(A, L, torch, dev, layout, scope, a1, a0, pk, mask, deficit, diag, flux, min_max, pl):
    minx, maxx, miny, maxy, nk = layout
    off = (1 - minx) + (maxx - minx + 1) * (1 - miny)
    arr = (A.AdzBcTracer * len(dev))()
    for n, t in enumerate(dev):
        legacy, weight, pexp_n = A.bc_settings(t["mass"])
        a = arr[n]
        a.p, a.m = A._ptr_at(t["p"], off), A._ptr_at(t["m"], off)
        a.lin, a.mn, a.mx = A._ptr_at(t["lin"], off), A._ptr_at(t["min"], off), A._ptr_at(t["max"], off)
        a.legacy, a.weight, a.pexp_n = legacy, weight, pexp_n
    dom = A.AdzBcDomain(minx, maxx, miny, maxy, nk, *scope, 1 if min_max else 0, (C.c_float * 9)(*pl))
    # every array at element (1,1,1) -- the 2-D mask at (1,1), the same offset
    return L.lib().adz_bermejo_conde_static(arr, len(dev), C.byref(dom), L.ADZ_F32, A._ptr_at(a1, off),
                                            A._ptr_at(a0, off), A._ptr_at(pk, off), A._ptr_at(mask, off),
                                            None if flux is None else flux.data_ptr(), deficit.data_ptr(),
                                            diag.data_ptr(), A._stream())


def test_static_fixer_matches_plan_fixer_on_two_calls(T):
    torch, L, A = T
    rng = np.random.default_rng(77)
    pl = (1.0, 2.0, 3.0) + (1.0,) * 6
    cases = [(_case(rng, 300, 90, 12, 8), True, True), (_case(rng, 70, 40, 6, 3), False, False)]
    for call, ((layout, scope, a1, a0, pk, mask, trs), with_flux, min_max) in enumerate(cases):
        print(f"[gpu-case] static fixer call {call}: {len(trs)} tracers, layout {layout}", flush=True)
        minx, maxx, miny, maxy, nk = layout
        off = (1 - minx) + (maxx - minx + 1) * (1 - miny)
        up = lambda x: torch.from_numpy(np.ascontiguousarray(x).reshape(-1)).cuda()
        dev_s = [{k: (up(v) if isinstance(v, np.ndarray) else v) for k, v in t.items()} for t in trs]
        dev_p = [{k: (up(v) if isinstance(v, np.ndarray) else v) for k, v in t.items()} for t in trs]
        d_a1, d_a0, d_pk, d_mask = up(a1), up(a0), up(pk), up(mask)
        ntr = len(trs)
        start = torch.tensor(rng.uniform(-1e-9, 1e-9, ntr), dtype=torch.float64, device="cuda")
        def_s, def_p = start.clone(), start.clone()
        diag_s = torch.full((6 * ntr,), -3.0, dtype=torch.float64, device="cuda")
        diag_p = diag_s.clone()
        flux = torch.tensor(rng.uniform(-1e-7, 1e-7, 2 * ntr), dtype=torch.float64, device="cuda") \
            if with_flux else None
        rc = _static(A, L, torch, dev_s, layout, scope, d_a1, d_a0, d_pk, d_mask, def_s, diag_s, flux, min_max, pl)
        L.check(rc, "adz_bermejo_conde_static")
        bc = A.BermejoConde(dev_p, layout, scope, off, min_max=min_max, pexp_list=pl)
        bc(d_a1, d_a0, d_mask, def_p, pkps=d_pk, diag=diag_p, flux=flux)
        torch.cuda.synchronize()
        bc.free()
        for n in range(ntr):
            assert torch.equal(dev_s[n]["m"].view(torch.int32), dev_p[n]["m"].view(torch.int32)), (call, n)
        assert torch.equal(def_s.view(torch.int64), def_p.view(torch.int64)), call
        assert torch.equal(diag_s.view(torch.int64), diag_p.view(torch.int64)), call
        assert not torch.equal(def_s, start), call  # the fixer ran


def test_static_fixer_refuses_more_than_64_tracers(T):
    torch, L, A = T
    t = torch.zeros(16, device="cuda")
    arr = (A.AdzBcTracer * 65)()
    for a in arr:
        a.p = a.m = a.lin = a.mn = a.mx = t.data_ptr()
        a.legacy, a.weight = 1, 1
    dom = A.AdzBcDomain(1, 4, 1, 2, 2, 1, 4, 1, 2, 1, 1, (C.c_float * 9)(*([1.0] * 9)))
    rc = L.lib().adz_bermejo_conde_static(arr, 65, C.byref(dom), L.ADZ_F32, t.data_ptr(), t.data_ptr(), None,
                                          t.data_ptr(), None, t.data_ptr(), None, A._stream())
    assert rc == L.ADZ_EINVAL
    assert "64" in L.lib().adz_last_error().decode()
