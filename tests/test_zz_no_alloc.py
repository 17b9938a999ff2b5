"""hh_cf_to_sf_noalloc (the CF -> SF transform with its work sequences in the library's static
device array) against hh_cf_to_sf on the same inputs, through the C ABI: two calls in a row,
the second with a smaller transform, each bit-identical to the allocating entry point (same
kernels, same order), and the input apodised in place as hh_cf_to_sf does."""
import ctypes as C

import numpy as np
import pytest

import hawaiihybrid_amd as H

pytestmark = pytest.mark.gpu


def _run(fn, eng, cf_host, dt, noalloc):
    import torch
    cf = torch.from_numpy(cf_host.copy()).to("cuda:0")
    nu = torch.full_like(cf, np.nan)
    sf = torch.full_like(cf, np.nan)
    torch.cuda.synchronize()
    n = cf.numel()
    if noalloc:
        rc = fn(eng.h, C.c_void_p(cf.data_ptr()), n, dt, C.c_void_p(nu.data_ptr()), C.c_void_p(sf.data_ptr()))
    else:
        ms = C.c_double(0)
        rc = fn(eng.h, C.c_void_p(cf.data_ptr()), n, dt, C.c_void_p(nu.data_ptr()), C.c_void_p(sf.data_ptr()),
                C.byref(ms))
    assert rc == 0, H.lib().hh_last_error()
    return cf.cpu().numpy(), nu.cpu().numpy(), sf.cpu().numpy()


def test_noalloc_cf_to_sf_matches_allocating_entry_point():
    eng = H.Engine(0)
    lib = H.lib()
    rng = np.random.default_rng(3)
    # a damped CF as the engine produces; 65536 points run the global-memory transform, 2048 the
    # LDS-resident one, and the second call is the smaller
    for n in (65536, 2048):
        t = 200.0 * np.arange(n)
        cf = np.exp(-t / 4.0e5) * (1.0 + 0.1 * rng.standard_normal(n)) * 1e-10
        ref = _run(lib.hh_cf_to_sf, eng, cf, 200.0, noalloc=False)
        got = _run(lib.hh_cf_to_sf_noalloc, eng, cf, 200.0, noalloc=True)
        for a, b in zip(got, ref):
            assert np.array_equal(a, b)
        assert np.all(np.isfinite(got[2]))
    eng.close()


def test_noalloc_cf_to_sf_rejects_a_transform_above_its_array():
    eng = H.Engine(0)
    import torch
    n = 131072
    cf = torch.zeros(n, dtype=torch.float64, device="cuda:0")
    rc = H.lib().hh_cf_to_sf_noalloc(eng.h, C.c_void_p(cf.data_ptr()), n, 200.0, C.c_void_p(cf.data_ptr()),
                                     C.c_void_p(cf.data_ptr()))
    assert rc == -1   # HH_E_INVALID
    assert b"static work array" in H.lib().hh_last_error()
    eng.close()
