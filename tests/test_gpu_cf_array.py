"""CORRELATION_ARRAY on the device (hh_cf_array) against the oracle: the same ensemble's rows
integrated by the oracle (FP64 DOP853, rtol 1e-12), weighted with the oracle's restatement of
src/hawaii.c:2639-2674 at energies from the oracle's Hamiltonian.

Tolerances: the FP64 engines agree with the oracle to ~1e-10 per trajectory (He-Ar); the
weights differ only by the rounding of the two Hamiltonians (~1e-16 relative).  At the base
temperature with ratio 1 every weight is exactly 1, so the array's CF is the plain CF's bits.
"""
import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eng():
    import hawaiihybrid_amd as H
    return H.Engine(0)


def rel_l2(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("engine", [0, 1])
def test_hear_bound_array_matches_oracle(eng, engine):
    import hawaiihybrid_amd as H
    print(f"cf_array he-ar bound engine={engine}", flush=True)
    T, n, L = 100.0, 240, 256
    temps = [50.0, 70.0, 100.0, 150.0]
    ratios = [1.5, 1.2, 1.0, 0.8]
    smp = H.Sampler(H.IC_REJECTION, H.PAIR_BOUND, T, 4.0, 40.0, H.PESMIN_HEAR, 0.0)
    ic, _, _ = H.sample_ics(H.he_ar(), smp, 11, n)
    x = eng.to_device_soa(ic)
    p = H.cf_params(L, rtol=1e-12, halves=H.HALVES_BOTH, engine=engine)
    cf, nstar, st = eng.cf_array(H.he_ar(), p, x, T, H.PAIR_BOUND, H.PESMIN_HEAR, temps, ratios)
    cf = cf.cpu().numpy()
    ref = O.cf_rows(O.hear(), O.CFParams(1e-12, 1e-12, 200.0, L, 40.0, O.HALVES_BOTH, 50000, 0), ic, 8)
    w, ns = O.cf_array_weights(O.hamiltonian(O.hear(), ic), T, 2, H.PESMIN_HEAR, temps, ratios,
                               counted=O.array_counted(ref["status"]))
    want = O.ensemble_array_from_rows(ref["crln"], w, ns)
    assert np.allclose(nstar, ns, rtol=1e-12, atol=0)
    for k in range(len(temps)):
        err = rel_l2(cf[k], want[k])
        print(f"  T={temps[k]}: nstar {nstar[k]:.4f} rel L2 {err:.2e}", flush=True)
        assert err < 1e-8


def test_array_at_base_temperature_is_the_plain_cf(eng):
    import hawaiihybrid_amd as H
    smp = H.Sampler(H.IC_R0_INJECT, H.PAIR_FREE, 200.0, 4.0, 40.0, H.PESMIN_CO2AR, 40.0)
    x = eng.to_device_soa(H.sample_ics(H.co2_ar(), smp, 12, 300)[0])
    p = H.cf_params(512, engine=H.ENGINE_BATCHED, precision=H.PREC_MIXED, rtol=1e-10)
    one, _ = eng.cf_run(H.co2_ar(), p, [x])
    arr, nstar, _ = eng.cf_array(H.co2_ar(), p, x, 200.0, H.PAIR_FREE, H.PESMIN_CO2AR, [200.0, 300.0])
    arr = arr.cpu().numpy()
    assert nstar[0] == 300.0 and nstar[1] == 0.0
    assert np.array_equal(arr[0], one[0].cpu().numpy())
    assert np.all(arr[1] == 0.0)


def test_array_rejects_bad_input(eng):
    import hawaiihybrid_amd as H
    x = eng.to_device_soa(np.tile([0.0, 0.0, 1.0, 0.0, 10.0, -1.0], (4, 1)))
    p = H.cf_params(64)
    with pytest.raises(H.HHError):
        eng.cf_array(H.he_ar(), p, x, 100.0, H.PAIR_ALL, H.PESMIN_HEAR, [50.0])
    with pytest.raises(H.HHError):
        eng.cf_array(H.he_ar(), p, x, 100.0, H.PAIR_FREE, H.PESMIN_HEAR, [])
    with pytest.raises(H.HHError):
        eng.cf_array(H.he_ar(), H.cf_params(64, zimmermann=True), x, 100.0, H.PAIR_BOUND, H.PESMIN_HEAR, [50.0])
