"""VEGAS phase-space integrals on the device (hh_vegas) against the reference's own runs.

Known answers (statistical): the reference logs of example_configs/*.conf.out hold hep-mc VEGAS
estimates over 12 iterations x 10^6 points of the same integrands at T = 100 K:
  CO2-Ar FREE_AND_METASTABLE, R in [4, 40]: PF 1.27517e12 (+-0.036%), M0 integral 1.01506e7
      (+-0.070%), i.e. PPF ratio 2.73233e5 and M0 6.25651e-4 (CF-F-CO2-Ar.conf.out:138-189)
  CO2-Ar BOUND:                             PF 1.53022e10 (+-0.068%), M0 integral 6.6114e6
      (+-0.085%), M0 4.07505e-4 (CF-B-CO2-Ar.conf.out:138-189)
  He-Ar FREE_AND_METASTABLE, R in [3, 40]:  M0 integral 0.347622 (+-0.018%), M0 2.08597e-6
      (CF-F-He-Ar.conf.out:186-190)
and of integrand_M2 (the same runs, lines 190-239 of each log): CO2-Ar FREE 0.148636 (+-0.076%),
CO2-Ar BOUND 0.0338833 (+-0.117%), He-Ar FREE 4.65447e-8 (+-0.018%).
The device estimate must agree within 4 combined standard errors.  A second pin ties VEGAS to the
rejection sampler (bit-exact with the reference's draws): M0-integral / PF-integral = <mu^2>,
which hh_m0 estimates independently; likewise M2-integral / PF-integral = <|d mu/dt|^2> of hh_m2.
"""
import math

import pytest

import hawaiihybrid_amd as H

pytestmark = pytest.mark.gpu


def close(val, err, ref, ref_rel, k=4.0):
    sig = math.hypot(err, ref * ref_rel)
    return abs(val - ref) < k * sig, (val - ref) / sig


@pytest.fixture(scope="module")
def eng():
    return H.Engine(0)


@pytest.mark.parametrize("case", [
    ("co2ar", H.PAIR_FREE, 4.0, H.INTEGRAND_PF, 1.27517e12, 3.59572e-4),
    ("co2ar", H.PAIR_FREE, 4.0, H.INTEGRAND_M0, 1.01506e7, 7.04919e-4),
    ("co2ar", H.PAIR_BOUND, 4.0, H.INTEGRAND_PF, 1.53022e10, 6.82507e-4),
    ("co2ar", H.PAIR_BOUND, 4.0, H.INTEGRAND_M0, 6.6114e6, 8.46597e-4),
    ("hear", H.PAIR_FREE, 3.0, H.INTEGRAND_M0, 0.347622, 1.79001e-4),
    ("co2ar", H.PAIR_FREE, 4.0, H.INTEGRAND_M2, 0.148636, 7.61738e-4),
    ("co2ar", H.PAIR_BOUND, 4.0, H.INTEGRAND_M2, 0.0338833, 1.16809e-3),
    ("hear", H.PAIR_FREE, 3.0, H.INTEGRAND_M2, 4.65447e-08, 1.76591e-4),
])
def test_vegas_matches_reference_logs(eng, case):
    which, pair, rmin, integrand, ref, ref_rel = case
    sysd = H.co2_ar() if which == "co2ar" else H.he_ar()
    print(f"vegas {which} pair {pair} integrand {integrand}", flush=True)
    r = eng.vegas(sysd, integrand, 100.0, rmin, 40.0, pair_state=pair, niterations=12, npoints=1000000, seed=7)
    ok, z = close(r.value, r.error, ref, ref_rel)
    print(f"  {r.value:.6e} +- {r.error:.2e} (ref {ref:.6e}); z = {z:+.2f}; chi2/dof {r.chi2_dof:.2f}; "
          f"nonfinite {r.nonfinite}; {r.ms:.1f} ms", flush=True)
    assert ok
    assert r.nonfinite == 0


def test_vegas_ratio_matches_rejection_sampler(eng):
    """<mu^2> two ways for He-Ar at 300 K FREE: VEGAS M0/PF, and the rejection sampler."""
    sysd = H.he_ar()
    T, rmin, rmax = 300.0, 4.0, 40.0
    m0 = eng.vegas(sysd, H.INTEGRAND_M0, T, rmin, rmax, niterations=12, npoints=1000000, seed=3)
    pf = eng.vegas(sysd, H.INTEGRAND_PF, T, rmin, rmax, niterations=12, npoints=1000000, seed=4)
    ratio = m0.value / pf.value
    rerr = ratio * math.hypot(m0.error / m0.value, pf.error / pf.value)
    smp = H.Sampler(H.IC_REJECTION, H.PAIR_FREE, T, rmin, rmax, H.PESMIN_HEAR, 0.0)
    nstreams = 1024
    m, q, _ = eng.m0(sysd, smp, 11, 2_000_000, nstreams, 1.0)
    # q is mpi_calculate_M0's error, the average of the streams' own errors (src/hawaii.c:1701-1702):
    # the error of the mean over the streams is q / sqrt(nstreams)
    mu2, mu2_err = m / H.ZERO_COEFF, q / H.ZERO_COEFF / math.sqrt(nstreams)
    z = (ratio - mu2) / math.hypot(rerr, mu2_err)
    print(f"VEGAS <mu^2> {ratio:.6e} +- {rerr:.1e}; rejection {mu2:.6e} +- {mu2_err:.1e}; z = {z:+.2f}", flush=True)
    assert abs(z) < 4.0


def test_vegas_m2_ratio_matches_rejection_sampler(eng):
    """<|d mu/dt|^2> two ways for He-Ar at 300 K FREE: VEGAS M2/PF, and hh_m2."""
    sysd = H.he_ar()
    T, rmin, rmax = 300.0, 4.0, 40.0
    m2 = eng.vegas(sysd, H.INTEGRAND_M2, T, rmin, rmax, niterations=12, npoints=1000000, seed=5)
    pf = eng.vegas(sysd, H.INTEGRAND_PF, T, rmin, rmax, niterations=12, npoints=1000000, seed=4)
    ratio = m2.value / pf.value
    rerr = ratio * math.hypot(m2.error / m2.value, pf.error / pf.value)
    smp = H.Sampler(H.IC_REJECTION, H.PAIR_FREE, T, rmin, rmax, H.PESMIN_HEAR, 0.0)
    nstreams = 1024
    m, q, _ = eng.m2(sysd, smp, 13, 2_000_000, nstreams, 1.0)
    rate, rate_err = m / H.SECOND_COEFF, q / H.SECOND_COEFF / math.sqrt(nstreams)
    z = (ratio - rate) / math.hypot(rerr, rate_err)
    print(f"VEGAS <|mudot|^2> {ratio:.6e} +- {rerr:.1e}; rejection {rate:.6e} +- {rate_err:.1e}; z = {z:+.2f}",
          flush=True)
    assert abs(z) < 4.0


def test_vegas_rejects_bad_input(eng):
    with pytest.raises(H.HHError):
        eng.vegas(H.he_ar(), 7, 100.0, 3.0, 40.0)
    with pytest.raises(H.HHError):
        eng.vegas(H.he_ar(), H.INTEGRAND_PF, 100.0, 40.0, 3.0)
