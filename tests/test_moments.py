"""Spectral moments from a CF, the analytic partition function, and the phase-space sampler mode
(host code of the engine: no GPU needed).

Pins: the stencils reproduce the textbook 9-point second difference and are exact on even
polynomials up to their order; M2 of a Gaussian CF has a closed form; the partition functions
equal the Boltzmann integrals of the kinetic energy done by quadrature; the engine and the
oracle (two different constructions of the stencil weights) agree; N2-Ar phase points are the
oracle's draws bit for bit."""
import ctypes as C
import math

import numpy as np
import pytest

import hawaiihybrid_amd as H
import oracle_lib as O

ALU = 5.29177210903e-11
ZERO = 0.00361479637 / (4.0 * math.pi)
SECOND = 13856114.29344114
HKT = 4.3597447222071e-18 / 1.380649e-23


def orc_moment(cf, dt, n, ntraj=0.0):
    out = C.c_double()
    cf = np.ascontiguousarray(cf, dtype=np.float64)
    rc = O.lib().orc_cf_moment(O.ptr(cf), len(cf), dt, ntraj, n, C.byref(out))
    assert rc == 0
    return out.value


def test_m0_is_cf0_times_zero_coeff():
    cf = np.array([2.5e-33, 1e-33, 0.0, 0.0, 0.0])
    assert H.cf_moment(cf, 200.0, 0) == pytest.approx(2.5e-33 * ZERO / ALU**3, rel=1e-15)
    assert H.cf_moment(cf, 200.0, 0, ntraj=4.0) == pytest.approx(2.5e-33 * ZERO / ALU**3 / 4.0, rel=1e-15)


@pytest.mark.parametrize("length", [5, 11, 64])
def test_m2_exact_on_even_polynomials(length):
    """f(t) = sum_m a_m t^(2m) up to the stencil's order: f''(0) = 2 a_1 exactly (|t| <= 1 on the
    stencil, so rounding does not swamp the difference)."""
    dt = 0.1
    N = 10 if length >= 11 else 4
    rng = np.random.default_rng(length)
    a = rng.normal(size=N + 1)
    t = dt * np.arange(length)
    f = sum(a[m] * t ** (2 * m) for m in range(N + 1)) * 1e-30
    expect = -SECOND * (2 * a[1] * 1e-30) / ALU**3
    for got in (H.cf_moment(f, dt, 2), orc_moment(f, dt, 2)):
        assert got == pytest.approx(expect, rel=1e-9)


def test_nine_point_stencil_is_the_textbook_one():
    """Weights of the 9-point f''(0): [-1/560, 8/315, -1/5, 8/5, -205/72, ...]; probe each weight
    with a unit impulse at sample j (even CF: the impulse sits at +-j)."""
    textbook = [-205.0 / 72.0, 8.0 / 5.0, -1.0 / 5.0, 8.0 / 315.0, -1.0 / 560.0]
    for j in range(5):
        f = np.zeros(5)
        f[j] = 1.0
        w = H.cf_moment(f, 1.0, 2) / (-SECOND / ALU**3)
        assert w == pytest.approx(textbook[j] * (1 if j == 0 else 2), rel=1e-14)


def test_m2_of_a_gaussian_cf():
    sigma, dt = 3000.0, 50.0
    t = dt * np.arange(4096)
    cf = np.exp(-t**2 / (2 * sigma**2))
    expect = SECOND / sigma**2 / ALU**3
    assert H.cf_moment(cf, dt, 2) == pytest.approx(expect, rel=1e-10)
    assert H.cf_moment(cf[:7], dt, 2) == pytest.approx(expect, rel=1e-6)


def test_engine_and_oracle_stencils_agree():
    rng = np.random.default_rng(1)
    for length in (5, 8, 11, 100):
        cf = np.cumsum(rng.normal(size=length))[::-1] * 1e-33
        a, b = H.cf_moment(cf, 200.0, 2, ntraj=3.0), orc_moment(cf, 200.0, 2, ntraj=3.0)
        assert a == pytest.approx(b, rel=1e-12, abs=1e-12 * np.max(np.abs(cf)) / ALU**3 * SECOND / 200**2)


def test_moment_rejects_short_or_unknown():
    with pytest.raises(H.HHError):
        H.cf_moment(np.ones(4), 1.0, 2)
    with pytest.raises(H.HHError):
        H.cf_moment(np.ones(20), 1.0, 4)


def test_partition_function_equals_kinetic_boltzmann_integral():
    from scipy import integrate
    T = 250.0
    kT = T / HKT
    # atom-atom: int exp(-p^2/2 mu kT) d^3p
    mu = H.MU_HEAR
    g = math.sqrt(2 * math.pi * mu * kT)
    assert H.pf_analytic(H.he_ar(), T) == pytest.approx(g**3, rel=1e-12)
    # linear molecule + atom: translational x rotor int dtheta dphi dptheta dpphi
    # exp(-(ptheta^2 + pphi^2/sin^2 theta)/2IkT) = 8 pi^2 I kT
    s = H.co2_ar()
    rot, _ = integrate.quad(lambda th: 2 * math.pi * math.sqrt(2 * math.pi * s.I * kT)
                            * math.sqrt(2 * math.pi * s.I * kT) * math.sin(th), 0, math.pi)
    expect = math.sqrt(2 * math.pi * s.mu * kT) ** 3 * rot
    assert H.pf_analytic(s, T) == pytest.approx(expect, rel=1e-10)
    orc = O.co2ar()
    assert O.lib().orc_pf_analytic(C.byref(orc), T) == pytest.approx(expect, rel=1e-10)
    # rotor + linear molecule: the classical rigid asymmetric top 8 pi^2 (2 pi kT)^(3/2) sqrt(I1 I2 I3)
    # times the linear molecule 8 pi^2 I kT times the translation (2 pi mu kT)^(3/2)
    I3, I_lin, mu = (2.0e4, 2.1e4, 2.2e4), 3.0e5, 2.0e4
    r = H.System(H.SYS_CH4CO2, mu, I_lin, (C.c_double * 3)(*I3))
    classical = (8 * math.pi**2 * (2 * math.pi * kT) ** 1.5 * math.sqrt(I3[0] * I3[1] * I3[2])
                 * 8 * math.pi**2 * I_lin * kT * (2 * math.pi * mu * kT) ** 1.5)
    assert H.pf_analytic(r, T) == pytest.approx(classical, rel=1e-12)


def test_n2ar_phase_points_bit_exact_with_oracle():
    n = 3000
    from hawaiihybrid_amd import n2ar
    ic = n2ar.phase_points(n, 300.0, seed=7)
    osys = O.System(O.SYS_CO2AR, H.MU_N2AR, (C.c_double * 2)(H.I_N2, H.I_N2))
    osmp = O.Sampler(O.IC_PHASE_SPACE, O.PAIR_ALL, 300.0, 4.51, 40.0, 0.0, 0.0)
    ref, ncand, _ = O.sample_ics(osys, osmp, 7, n)
    assert ncand == n
    assert np.array_equal(ic, ref)
    # R is drawn uniformly in R^3 over [Rmin, Rmax]
    R3 = ic[:, 4] ** 3
    u = (R3 - 4.51**3) / (40.0**3 - 4.51**3)
    assert abs(u.mean() - 0.5) < 0.02 and u.min() >= 0 and u.max() <= 1


def test_n2ar_rejection_needs_plugin():
    smp = H.Sampler(H.IC_REJECTION, H.PAIR_FREE, 300.0, 4.51, 40.0, H.PESMIN_N2AR, 0.0)
    with pytest.raises(H.HHError):
        H.sample_ics(H.n2_ar(), smp, 1, 10)
