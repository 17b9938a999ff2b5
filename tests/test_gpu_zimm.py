"""Zimmermann's trick on the device (the dipole series of both halves, FFT cross-correlation over
every time origin) against the oracle's direct O(L^2) restatement of
correlation_eval_zimmerman_trick on the same bound-state initial conditions.

Tolerance: the device and the oracle integrate the same FP64 DOP853 trajectories (same tolerance),
so their dipole series agree to ~1e-10 relative; the ensemble CFs agree to 1e-8 relative L2
(FFT rounding adds ~1e-15).  The plain and batched engines run the same per-item code and the
same reduction: their CFs are equal bit for bit.
"""
import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu


def bound_ics(system_h, system_o, pesmin, n, T, seed):
    import hawaiihybrid_amd as H
    smp = H.Sampler(H.IC_REJECTION, H.PAIR_BOUND, T, 4.0, 40.0, pesmin, 0.0)
    ic, _, _ = H.sample_ics(system_h, smp, seed, n)
    osmp = O.Sampler(O.IC_REJECTION, O.PAIR_BOUND, T, 4.0, 40.0, pesmin, 0.0)
    ref, _, _ = O.sample_ics(system_o, osmp, seed, n)
    assert np.array_equal(ic, ref)
    return ic


@pytest.mark.parametrize("which,L,n", [("hear", 128, 300), ("hear", 100, 200), ("co2ar", 64, 96)])
def test_zimmermann_matches_oracle(which, L, n):
    import hawaiihybrid_amd as H
    print(f"zimmermann {which} L={L} n={n}", flush=True)
    if which == "hear":
        sh, so, pesmin, T = H.he_ar(), O.hear(), H.PESMIN_HEAR, 60.0
    else:
        sh, so, pesmin, T = H.co2_ar(), O.co2ar(), H.PESMIN_CO2AR, 150.0
    ic = bound_ics(sh, so, pesmin, n, T, 5)
    eng = H.Engine(0)
    x = eng.to_device_soa(ic)
    outs = {}
    for engine in (H.ENGINE_PLAIN, H.ENGINE_BATCHED):
        p = H.cf_params(L, rtol=1e-12, halves=H.HALVES_BOTH, engine=engine, zimmermann=True)
        out, st = eng.cf_run(sh, p, [x])
        outs[engine] = out[0].cpu().numpy()
        items = eng.items(0, 2 * n, L, rows=True, zimmermann=True)
        assert items["rows"].shape == (2 * n, 3, L)
    assert np.array_equal(outs[H.ENGINE_PLAIN], outs[H.ENGINE_BATCHED])
    ref = O.cf_rows(so, O.CFParams(1e-12, 1e-12, 200.0, L, 40.0, O.HALVES_BOTH, 50000, 1), ic, 8)
    cf_ref = O.ensemble_from_rows(ref["crln"])
    err = np.linalg.norm(outs[H.ENGINE_BATCHED] - cf_ref) / np.linalg.norm(cf_ref)
    print(f"  rel L2 vs oracle {err:.3e}", flush=True)
    assert err < 1e-8
    # the dipole series the device recorded: mu(0) of both halves is the same point
    assert np.allclose(items["rows"][0::2, :, 0], items["rows"][1::2, :, 0], rtol=0, atol=0)


def test_zimmermann_rejects_forward_only_and_long_series():
    import hawaiihybrid_amd as H
    eng = H.Engine(0)
    ic = np.zeros((4, 6))
    ic[:, 4] = 10.0
    x = eng.to_device_soa(ic)
    with pytest.raises(H.HHError):
        eng.cf_run(H.he_ar(), H.cf_params(64, halves=H.HALVES_FORWARD, zimmermann=True), [x])
    with pytest.raises(H.HHError):
        eng.cf_run(H.he_ar(), H.cf_params(8192, halves=H.HALVES_BOTH, zimmermann=True), [x])
