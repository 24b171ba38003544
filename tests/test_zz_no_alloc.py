"""The allocation-free surrogate entry point
(pcfd_chem_surrogate_wdot_device_noalloc) against the allocating one
(pcfd_chem_surrogate_wdot_device) through the C ABI: identical rates and
fallback counts on two calls in a row, the second on fewer cells, and a
clear refusal of more cells than its fixed scratch holds."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _call(L, fn, ctx, net, d, n, rr, tt, ww):
    nf, ms = C.c_int(-1), C.c_double()
    rc = getattr(L, fn)(ctx.h, net.h, C.byref(d), n, rr.ptr, tt.ptr, ww.ptr, C.byref(nf), C.byref(ms))
    assert rc == 0, (fn, rc, L.pcfd_mlp_last_error(net.h))
    return nf.value


def test_noalloc_matches_allocating_entry_point():
    from proteuscfd_amd import chem, gpu, surrogate as S
    L = gpu._mlp_lib()
    ctx = gpu.ChemContext(chem.load_air5())
    print("[noalloc] short surrogate training (200k states, 600 steps)", flush=True)
    net, tr, _ = S.train_surrogate(ctx, n=200_000, steps=600, batch=4096, lr=3e-3, seed=40)
    d = S.desc(tr, tol=0.1, floor=10.0)
    # states partly outside the training box, so both branches of the rule run
    rhoi, T = S.sample_states(300_000, seed=41)
    T[::97] *= 3.0
    for n in (300_000, 70_001):   # second call smaller, ragged last pass
        print(f"[noalloc] n={n}", flush=True)
        rr, tt = gpu.DeviceArray(rhoi[:n]), gpu.DeviceArray(T[:n])
        wa, wb = gpu.DeviceArray(np.full((n, 5), np.nan)), gpu.DeviceArray(np.full((n, 5), np.nan))
        nfa = _call(L, "pcfd_chem_surrogate_wdot_device", ctx, net, d, n, rr, tt, wa)
        nfb = _call(L, "pcfd_chem_surrogate_wdot_device_noalloc", ctx, net, d, n, rr, tt, wb)
        a, b = wa.numpy(), wb.numpy()
        assert nfa == nfb and 0 < nfb < n, (nfa, nfb)
        assert np.isfinite(b).all()
        assert np.array_equal(a, b)
        # the fallback cells carry the exact rates
        outside = T[:n] > S.T_MAX
        we = ctx.wdot(rhoi[:n][outside], T[:n][outside])
        np.testing.assert_allclose(b[outside], we, rtol=1e-13, atol=0)
    # more cells than the scratch holds: refused before any launch
    nbig = gpu.SURROGATE_NOALLOC_MAX_CELLS + 1
    rc = L.pcfd_chem_surrogate_wdot_device_noalloc(ctx.h, net.h, C.byref(d), nbig, rr.ptr, tt.ptr, wb.ptr, None,
                                                   None)
    assert rc == -1
    assert b"PCFD_SURROGATE_NOALLOC_MAX_CELLS" in L.pcfd_mlp_last_error(net.h)
    for a in (rr, tt, wa, wb):
        a.close()
    net.close()
    ctx.close()
