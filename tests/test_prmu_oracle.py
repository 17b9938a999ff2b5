"""Pins for the oracle's PR/mu pieces (src/hawaii.c:3148-3977): the per-trajectory spectrum
against numpy's real FFT and Parseval's theorem, and the dipole series against an independent
integration of Hamilton's equations (scipy's DOP853).  CPU only."""
import ctypes as C

import numpy as np
import pytest
from scipy.integrate import solve_ivp

import oracle_lib as O


def connes(x, dt):
    """(1 - (t/tmax)^2)^2 with tmax = dt * len (connes_apodization, src/hawaii.c:5401-5410)."""
    n = x.shape[-1]
    r = (dt * np.arange(n)) / (dt * n)
    return x * (1.0 - r * r) ** 2


@pytest.mark.parametrize("L,used", [(64, 40), (1024, 1024), (4096, 700)])
def test_prmu_power_matches_numpy_rfft(L, used):
    rng = np.random.default_rng(L)
    x = np.zeros((3, L))
    x[:, :used] = rng.normal(size=(3, used))
    nfreq = L // 2 + 1
    got = O.prmu_power(x, 200.0, nfreq)
    want = sum(np.abs(np.fft.rfft(connes(x[c], 200.0))) ** 2 for c in range(3))
    np.testing.assert_allclose(got, want, rtol=1e-11, atol=1e-12 * want.max())


def test_prmu_power_parseval():
    """sum over all L bins of |X_k|^2 = L sum |x_i|^2 (the half spectrum counted twice inside)."""
    L = 2048
    x = np.random.default_rng(3).normal(size=(3, L))
    P = O.prmu_power(x, 100.0, L // 2 + 1)
    total = P[0] + P[-1] + 2 * P[1:-1].sum()
    assert total == pytest.approx(L * (connes(x, 100.0) ** 2).sum(), rel=1e-12)


def test_prmu_series_matches_independent_integration():
    """He-Ar trajectories injected at R0 = 15: the oracle's series (t = k dt, up to and
    including the first sample beyond R0) against scipy's DOP853 on the oracle's Hamilton
    equations (pinned against the Hamiltonian elsewhere) and the oracle's dipole."""
    s = O.hear()
    L = O.lib()
    smp = O.Sampler(O.IC_R0_INJECT, O.PAIR_FREE, 300.0, 4.0, 40.0, O.PESMIN_HEAR, 15.0)
    ic, _, _ = O.sample_ics(s, smp, 5, 4)
    dt, R0 = 200.0, 15.0

    def f(t, y):
        out = np.zeros(6)
        L.orc_rhs(C.byref(s), O.ptr(np.ascontiguousarray(y)), O.ptr(out))
        return out

    for y0 in ic:
        status, n, vec = O.prmu_series(s, y0, dt, 4096, R0)
        assert status == 0 and 2 < n < 4096
        t_end = (n + 1) * dt
        sol = solve_ivp(f, (0.0, t_end), y0, method="DOP853", rtol=1e-13, atol=1e-13, dense_output=True)
        ys = sol.sol(dt * np.arange(n + 1))
        R = ys[4]
        # the series ends at the first sample beyond R0
        assert np.all(R[1:n - 1] <= R0) and R[n - 1] > R0
        for k in range(n):
            mu = np.zeros(3)
            L.orc_dipole(C.byref(s), O.ptr(np.ascontiguousarray(ys[0::2, k])), O.ptr(mu))
            np.testing.assert_allclose(vec[k], mu, rtol=1e-6, atol=1e-9 * np.abs(vec).max())


def test_sf_moment_simpson_exact_on_cubics():
    """integrate_composite_simpson is Simpson's 3/8 rule: exact for nu^2 (and nu^3) when the grid
    has 3m + 1 points; the engine's and the oracle's restatements agree, odd orders give 0."""
    import hawaiihybrid_amd as H
    coeff = (8.0 * np.pi ** 3) * 2.686780111e25 ** 2 / 3.0 / 1.054571817e-34
    nu = np.linspace(0.0, 150.0, 3 * 40 + 1)
    sf = np.ones_like(nu)
    m2 = H.sf_moment(nu, sf, 2)
    assert m2 == pytest.approx(2 * coeff * 150.0 ** 3 / 3, rel=1e-13)
    m0 = H.sf_moment(nu, nu.copy(), 0, ntraj=4.0)      # int nu d nu / 4
    assert m0 == pytest.approx(2 * coeff * 150.0 ** 2 / 2 / 4.0, rel=1e-13)
    assert H.sf_moment(nu, sf, 1) == 0.0
    rng = np.random.default_rng(2)
    y = rng.random(301)
    x = np.linspace(0.0, 200.0, 301)
    L = O.lib()
    for n in (0, 2):
        assert H.sf_moment(x, y, n, 7.0) == pytest.approx(L.orc_sf_moment(O.ptr(x), O.ptr(y), 301, 7.0, n), rel=1e-14)


def test_sample_ics_from_continues_the_stream():
    """hh_sample_ics_from(skip = the draws of a first call) returns what one longer call returns
    after the first call's points (R0 injection draws a fixed number of words per point)."""
    import hawaiihybrid_amd as H
    sh = H.co2_ar()
    smp = H.Sampler(H.IC_R0_INJECT, H.PAIR_FREE, 300.0, 4.0, 40.0, H.PESMIN_CO2AR, 20.0)
    whole, _, _ = H.sample_ics(sh, smp, 77, 50)
    first, _, d1 = H.sample_ics_from(sh, smp, 77, 0, 20)
    rest, _, d2 = H.sample_ics_from(sh, smp, 77, d1, 30)
    assert np.array_equal(first, whole[:20]) and np.array_equal(rest, whole[20:])
    rej = H.Sampler(H.IC_REJECTION, H.PAIR_FREE, 300.0, 4.0, 40.0, H.PESMIN_CO2AR, 0.0)
    a, _, da = H.sample_ics(sh, rej, 5, 40)
    b, _, db = H.sample_ics_from(sh, rej, 5, 0, 25)
    c, _, dc = H.sample_ics_from(sh, rej, 5, db, 15)
    assert np.array_equal(np.vstack([b, c]), a) and dc == da
