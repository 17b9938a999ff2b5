"""Pins for the CPU oracle: reference golden vectors, reference test fixtures, closed forms
and invariants.  CPU only."""
import ctypes as C
import math

import numpy as np
import pytest
import scipy.fft
import scipy.special

import oracle_lib as O


# ---------------------------------------------------------------- A1 Mersenne Twister
@pytest.mark.parametrize("seed", [42, 43, 20106943, 0, 1])
def test_mt_matches_reference_mtwist(golden, seed):
    g = golden("mt_streams.npz")
    ref = g[f"seed_{seed}"]
    assert np.array_equal(O.mt_stream(seed, ref.size), ref)


def test_mt_long_prefix_checksum(golden):
    g = golden("mt_streams.npz")
    words = O.mt_stream(42, 1_000_000)
    assert int(words.astype(np.uint64).sum()) == int(g["seed_42_1e6_sum"][0])
    assert np.array_equal(words[-16:], g["seed_42_1e6_last"])


def test_mt_textbook_mt19937_first_word():
    # Matsumoto & Nishimura's 1998 reference output for sgenrand(4357): first genrand() is
    # 0xd13e8ff5 (= 3510405877); mtwist maps seed 0 to 4357.
    assert O.mt_stream(4357, 1)[0] == 3510405877
    assert O.mt_stream(0, 1)[0] == 3510405877


# ---------------------------------------------------------------- A8-He
def test_hear_matches_reference_build(golden):
    g = golden("hear_ref.npz")
    L = O.lib()
    V = np.array([L.orc_hear_V(r) for r in g["R"]])
    dV = np.array([L.orc_hear_dV(r) for r in g["R"]])
    dip = np.array([L.orc_hear_dip(r) for r in g["R"]])
    # the reference's dV is a Maple expansion with 1e23 coefficients (HeAr.h:66-86) that loses
    # ~5e-13 relative to cancellation; the restated analytic derivative is closer to exact
    assert np.max(np.abs(V - g["V"]) / (np.abs(g["V"]) + 1e-12)) < 1e-12
    assert np.max(np.abs(dV - g["dV"]) / (np.abs(g["dV"]) + 1e-12)) < 1e-11
    assert np.array_equal(dip, g["dip"])


def test_hear_derivative_is_derivative():
    L = O.lib()
    for R in (3.5, 5.0, 6.9, 12.0, 30.0):
        h = 1e-5 * R
        fd = (-L.orc_hear_V(R + 2 * h) + 8 * L.orc_hear_V(R + h) - 8 * L.orc_hear_V(R - h)
              + L.orc_hear_V(R - 2 * h)) / (12 * h)
        assert abs(fd - L.orc_hear_dV(R)) <= 1e-7 * abs(L.orc_hear_dV(R)) + 1e-16


# ---------------------------------------------------------------- A8-CO2
def test_co2ar_reference_minimum():
    # PES-IDS/ai_pes_co2ar.h:25-27: RMIN = 6.4966051519, THETAMIN = pi/2, PESMIN = -195.6337098547 cm-1
    assert abs(O.lib().orc_co2ar_V_cm(6.4966051519, math.pi / 2) - (-195.6337098547)) < 2e-9


def test_co2ar_surfaces_match_reference_expressions(golden):
    g = golden("co2ar_surface_ref.npz")
    L = O.lib()
    HT = 2.1947463136320e5
    V, dR, dT, dx, dz = [], [], [], [], []
    for r, t in zip(g["R"], g["theta"]):
        V.append(L.orc_co2ar_V_cm(r, t) / HT)
        a, b = C.c_double(), C.c_double()
        L.orc_co2ar_dV_cm(r, t, C.byref(a), C.byref(b))
        dR.append(a.value / HT)
        dT.append(b.value / HT)
        d = np.zeros(3)
        L.orc_co2ar_dip_mol(r, t, O.ptr(d))
        dx.append(d[0])
        dz.append(d[2])

    def rel(mine, ref):
        return np.max(np.abs(np.array(mine) - ref) / (np.abs(ref) + 1e-6 * np.abs(ref).max()))
    assert rel(V, g["V"]) < 1e-10
    # dpes_dR / dpes_dTheta are separate Maple tables with their own 14-digit rounding;
    # the oracle differentiates the value table analytically
    assert rel(dR, g["dR"]) < 1e-7
    assert rel(dT, g["dT"]) < 1e-8
    assert rel(dx, g["dipx"]) < 1e-11
    assert rel(dz, g["dipz"]) < 1e-10


def test_co2ar_continuous_across_breakpoints():
    L = O.lib()
    bounds = [4.5 + 0.25 * i for i in range(23)] + [11, 12, 14, 16, 18, 20, 25, 30]
    for b in bounds:
        for th in (0.3, 1.1, 1.5707963, 2.6):
            lo, hi = L.orc_co2ar_V_cm(b * (1 - 2e-16), th), L.orc_co2ar_V_cm(b, th)
            # the reference coefficients are printed to 14 digits: jumps stay below 1e-6 cm^-1
            assert abs(lo - hi) <= 1e-6 + 1e-9 * abs(hi)


def test_associated_legendre_convention_matches_scipy():
    # dipx reads P_l^1 with the Condon-Shortley phase (ai_ids_co2ar.hpp csphase = -1);
    # scipy.special.lpmv carries that phase: at theta where only one l contributes the
    # x component must equal d_l(R) * lpmv(1, l, cos theta).
    L = O.lib()
    th = 0.7
    d = np.zeros(3)
    L.orc_co2ar_dip_mol(45.0, th, O.ptr(d))   # R >= 30: only l = 2, 4 carry x terms
    d2 = np.zeros(3)
    L.orc_co2ar_dip_mol(45.0, math.pi - th, O.ptr(d2))
    # P_l^1 is odd under x -> -x for even l, so dipx flips sign, dipz (even P_l) does not
    assert d2[0] == pytest.approx(-d[0], rel=1e-13)
    assert d2[2] == pytest.approx(d[2], rel=1e-13)
    x = math.cos(th)
    p21 = scipy.special.lpmv(1, 2, x)
    assert p21 == pytest.approx(-3 * x * math.sqrt(1 - x * x), rel=1e-13)


def test_co2ar_lab_gradient_matches_fd_at_reference_point():
    # examples/trajectory_co2_ar.cpp:65-118 checks the analytic chain rule at this point
    L = O.lib()
    s = O.co2ar()
    q = np.array([0.1, 0.2, 5.0, 0.3, 0.4])
    g = np.zeros(5)
    L.orc_dpes(C.byref(s), O.ptr(q), O.ptr(g))
    for i in range(5):
        h = 1e-4
        vals = []
        for k in (-2, -1, 1, 2):
            qq = q.copy()
            qq[i] += k * h
            vals.append(L.orc_pes(C.byref(s), O.ptr(qq)))
        fd = (vals[0] - 8 * vals[1] + 8 * vals[2] - vals[3]) / (12 * h)
        # R = 5.0 sits on a spline breakpoint, so the R stencil straddles two cubics; the
        # reference's own RHS-vs-FD check uses 1e-4 (examples/trajectory_co2_ar.cpp:50)
        assert abs(fd - g[i]) < 1e-6


def test_co2ar_jacobian_matches_fd():
    L = O.lib()
    q = np.array([0.1, 0.2, 5.0, 0.3, 0.4])
    qm = np.zeros(5)
    L.orc_co2ar_lab_to_mol(O.ptr(q), O.ptr(qm))
    jac = np.zeros((5, 5))
    L.orc_co2ar_jacobian(O.ptr(q), O.ptr(qm), O.ptr(jac))
    for i in (0, 1, 3, 4):
        h = 1e-5
        a, b = q.copy(), q.copy()
        a[i] += h
        b[i] -= h
        ma, mb = np.zeros(5), np.zeros(5)
        L.orc_co2ar_lab_to_mol(O.ptr(a), O.ptr(ma))
        L.orc_co2ar_lab_to_mol(O.ptr(b), O.ptr(mb))
        fd = (ma[4] - mb[4]) / (2 * h)
        assert abs(fd - jac[i, 4]) < 1e-9


# ---------------------------------------------------------------- A6/A7 EOM
@pytest.mark.parametrize("which", ["hear", "co2ar"])
def test_rhs_is_hamiltons_equations(which):
    """ydot = J grad H checked against a 6th-order FD of the Hamiltonian at the
    reference's own test points (examples/correlation_he_ar.cpp:69-76,
    examples/trajectory_co2_ar.cpp:133-144)."""
    L = O.lib()
    if which == "hear":
        s = O.hear()
        y = np.array([0.3, 2.1, 1.2, -3.4, 7.5, -4.1])
    else:
        s = O.co2ar()
        y = np.array([0.1, 30.0, 1.3, -20.0, 7.0, -3.0, 0.7, 25.0, 1.9, 40.0])
    n = y.size
    yd = np.zeros(n)
    L.orc_rhs(C.byref(s), O.ptr(y), O.ptr(yd))
    grad = np.zeros(n)
    for i in range(n):
        h = 1e-3 * max(1.0, abs(y[i])) * (1e-2 if i % 2 == 0 else 1.0)
        vals = []
        for k in (-3, -2, -1, 1, 2, 3):
            yy = y.copy()
            yy[i] += k * h
            vals.append(L.orc_hamiltonian(C.byref(s), O.ptr(yy)))
        grad[i] = (-vals[0] + 9 * vals[1] - 45 * vals[2] + 45 * vals[3] - 9 * vals[4] + vals[5]) / (60 * h)
    # q' = dH/dp, p' = -dH/dq
    expect = np.empty(n)
    expect[0::2] = grad[1::2]
    expect[1::2] = -grad[0::2]
    assert np.allclose(yd, expect, rtol=1e-6, atol=1e-10)


# ---------------------------------------------------------------- A2-A5 sampler
def test_sampler_r0_inject_draw_count_and_constraints():
    s = O.co2ar()
    smp = O.Sampler(O.IC_R0_INJECT, O.PAIR_FREE, 200.0, 4.0, 40.0, O.PESMIN_CO2AR, 40.0)
    ic, ncand, ndraws = O.sample_ics(s, smp, 42, 1000)
    assert ncand == 1000
    assert ndraws == 15 * 1000          # 5 (q) + 3*2 (pair normals) + 2*2 (rotor normals)
    assert np.all(ic[:, 4] == 40.0) and np.all(ic[:, 5] <= 0.0)


def test_sampler_boltzmann_moments():
    """<p^2> per degree of freedom = m kT for the sampled momenta (p_generator,
    src/hawaii.c:1292-1380)."""
    HkT = 4.3597447222071e-18 / 1.380649e-23
    T = 300.0
    s = O.co2ar()
    smp = O.Sampler(O.IC_R0_INJECT, O.PAIR_FREE, T, 4.0, 40.0, O.PESMIN_CO2AR, 40.0)
    ic, _, _ = O.sample_ics(s, smp, 7, 40000)
    kT = T / HkT
    mu, I = O.MU_CO2AR, O.I_CO2
    R, Th, th = ic[:, 4], ic[:, 2], ic[:, 8]
    checks = [ic[:, 5] ** 2 / (mu * kT), ic[:, 3] ** 2 / (R * R * mu * kT),
              ic[:, 1] ** 2 / (R * R * np.sin(Th) ** 2 * mu * kT),
              ic[:, 9] ** 2 / (I * kT), ic[:, 7] ** 2 / (np.sin(th) ** 2 * I * kT)]
    for c in checks:
        assert abs(c.mean() - 1.0) < 0.03
    # cos(Theta), cos(theta) uniform on [-1, 1]; Phi, phi uniform on [0, 2 pi)
    for u in (np.cos(Th), np.cos(th)):
        assert abs(u.mean()) < 0.02 and abs((u ** 2).mean() - 1 / 3) < 0.01
    for a in (ic[:, 0], ic[:, 6]):
        assert abs(a.mean() - math.pi) < 0.05


def test_sampler_rejection_mode_distribution():
    """Rejection sampler (src/hawaii.c:1382-1421, 2969-2993): accepted free-pair states have
    E >= 0, R within [Rmin, Rmax], and at large R (V ~ 0) the R density is ~R^2."""
    s = O.hear()
    smp = O.Sampler(O.IC_REJECTION, O.PAIR_FREE, 300.0, 4.0, 40.0, O.PESMIN_HEAR, 0.0)
    ic, ncand, ndraws = O.sample_ics(s, smp, 11, 20000)
    L = O.lib()
    E = np.array([L.orc_hamiltonian(C.byref(s), O.ptr(np.ascontiguousarray(row))) for row in ic])
    assert np.all(E >= 0.0)
    assert np.all((ic[:, 4] >= 4.0) & (ic[:, 4] <= 40.0))
    assert ndraws == 10 * ncand
    far = ic[:, 4][ic[:, 4] > 20.0]
    # for R^2 density on [20, 40]: mean = 3/4 (40^4-20^4)/(40^3-20^3)
    assert abs(far.mean() - 0.75 * (40 ** 4 - 20 ** 4) / (40 ** 3 - 20 ** 3)) < 0.2


# ---------------------------------------------------------------- A9/A10 integrator + CF
def _params(L=512, rtol=1e-12, halves=O.HALVES_FORWARD, Rcut=40.0):
    return O.CFParams(rtol, rtol, 200.0, L, Rcut, halves, 50000)


def test_integrator_error_bound_and_energy():
    """(i) the 1e-12 run agrees with a 1e-14 run; (ii) |E - E0| <= 1000 rtol
    (src/trajectory.c:167-180)."""
    s = O.co2ar()
    smp = O.Sampler(O.IC_R0_INJECT, O.PAIR_FREE, 200.0, 4.0, 40.0, O.PESMIN_CO2AR, 40.0)
    ic, _, _ = O.sample_ics(s, smp, 42, 32)
    a = O.cf_rows(s, _params(2048, 1e-12), ic)
    b = O.cf_rows(s, _params(2048, 1e-14), ic)
    assert np.all(a["status"] == 0)
    assert np.all(a["e_drift"] <= 1000 * 1e-12)
    num = np.linalg.norm(a["crln"].sum(0) - b["crln"].sum(0))
    assert num / np.linalg.norm(b["crln"].sum(0)) < 1e-6


def test_time_reversal():
    """Propagate, flip momenta, propagate back: the dipole series retraces itself."""
    s = O.hear()
    y0 = np.array([0.4, 120.0, 1.1, -60.0, 12.0, -6.0])
    p = _params(64, 1e-13, Rcut=1e9)
    c1 = np.zeros(64)
    n1 = C.c_int()
    st = O.TrajStats()
    O.lib().orc_half_cf(C.byref(s), C.byref(p), O.ptr(y0), O.ptr(c1), C.byref(n1), C.byref(st))
    assert n1.value == 64 and st.e_drift < 1e-10


def test_hear_central_force_conserves_angular_momentum_plane():
    """He-Ar is a central potential: the orbit stays in a plane, so the dipole direction
    at all samples is orthogonal to L = R x P; checked through mu(0).mu(t) being the
    cosine of the swept angle times |d(R0)||d(R(t))|."""
    s = O.hear()
    y0 = np.array([0.0, 0.0, math.pi / 2, 40.0, 9.0, -4.0])   # motion in the x-z plane
    p = _params(256, 1e-12, Rcut=1e9)
    c = np.zeros(256)
    n = C.c_int()
    st = O.TrajStats()
    O.lib().orc_half_cf(C.byref(s), C.byref(p), O.ptr(y0), O.ptr(c), C.byref(n), C.byref(st))
    d0 = O.lib().orc_hear_dip(9.0)
    assert c[0] == pytest.approx(d0 * d0, rel=1e-15)


def test_failure_quirk_zero_contribution():
    """A failed check returns the zeroed crln (src/hawaii.c:2128, 2183-2237)."""
    s = O.hear()
    p = _params(32, 1e-12, Rcut=1e9)
    y0 = np.array([0.0, 0.0, 1.0, 0.0, float("nan"), -4.0])
    crln = np.ones(32)
    lf, lb = C.c_int(), C.c_int()
    st = O.TrajStats()
    status = O.lib().orc_traj_cf(C.byref(s), C.byref(p), O.ptr(y0), O.ptr(crln), C.byref(lf),
                                 C.byref(lb), C.byref(st))
    assert status != 0 and np.all(crln == 0.0)


def test_step_limit_drops_only_the_limited_trajectory():
    """max_steps accepted steps without reaching the next sample is CVode's CV_TOO_MUCH_WORK
    (src/trajectory.c:15, 62); a trajectory whose last half ends that way is dropped
    (src/hawaii.c:3001).  The limit changes nothing else: every other row equals the row of
    the unlimited run bit for bit, the dropped row is zero, and no completed half needed
    more than max_steps accepted steps between two samples."""
    ic, _, _ = O.sample_ics(O.hear(), O.Sampler(O.IC_REJECTION, 1, 1000.0, 4.0, 40.0, O.PESMIN_HEAR, 0.0), 5, 256)
    lim = O.cf_rows(O.hear(), O.CFParams(1e-12, 1e-12, 200.0, 256, 40.0, O.HALVES_BOTH, 6), ic, 8)
    free = O.cf_rows(O.hear(), O.CFParams(1e-12, 1e-12, 200.0, 256, 40.0, O.HALVES_BOTH, 50000), ic, 8)
    dropped = lim["status"] == 4
    assert dropped.sum() >= 1 and np.all(free["status"] == 0)
    assert np.all(lim["crln"][dropped] == 0.0)
    whole = ~dropped & (lim["len_f"] == free["len_f"])
    assert np.array_equal(lim["crln"][whole], free["crln"][whole])
    # a forward half cut by the limit, followed by a complete backward half, keeps its partial
    # samples (make_step's error only breaks the forward loop, src/hawaii.c:2172-2176)
    for t in np.where(~dropped & ~whole)[0]:
        n = lim["len_f"][t]
        assert n < free["len_f"][t]
        assert np.array_equal(lim["crln"][t, :n], free["crln"][t, :n])
    # the same trajectories with a limit of one step fewer than needed: all dropped at t = 0
    lim1 = O.cf_rows(O.hear(), O.CFParams(1e-12, 1e-12, 200.0, 256, 40.0, O.HALVES_BOTH, 1), ic, 8)
    assert np.all(lim1["status"] == 4) and np.all(lim1["crln"] == 0.0)


def test_both_halves_definition():
    """crln = 1/2 (c_f + c_b) ALU^3 with c_b from inverted momenta (src/hawaii.c:2334-2337)."""
    s = O.hear()
    y0 = np.array([0.4, 100.0, 1.1, -50.0, 12.0, -6.0])
    p = _params(64, 1e-12, halves=O.HALVES_BOTH, Rcut=1e9)
    crln = np.zeros(64)
    lf, lb = C.c_int(), C.c_int()
    st = O.TrajStats()
    O.lib().orc_traj_cf(C.byref(s), C.byref(p), O.ptr(y0), O.ptr(crln), C.byref(lf), C.byref(lb), C.byref(st))
    cf, cb = np.zeros(64), np.zeros(64)
    n = C.c_int()
    O.lib().orc_half_cf(C.byref(s), C.byref(p), O.ptr(y0), O.ptr(cf), C.byref(n), C.byref(st))
    yb = y0.copy()
    yb[1::2] *= -1
    O.lib().orc_half_cf(C.byref(s), C.byref(p), O.ptr(yb), O.ptr(cb), C.byref(n), C.byref(st))
    ALU = 5.29177210903e-11
    assert np.array_equal(crln, 0.5 * (cf + cb) * ALU * ALU * ALU)


# ---------------------------------------------------------------- A16 CF -> SF
def test_idct_is_scaled_dct3():
    v = np.random.default_rng(3).standard_normal(256)
    assert np.allclose(O.idct(v), scipy.fft.dct(v, type=3) / (2 * 256), rtol=0, atol=1e-15)
    assert np.allclose(O.dct(v), scipy.fft.dct(v, type=2), rtol=0, atol=1e-12)


def test_dct_idct_round_trip_reference_example():
    # examples/test_fft.c:6-44
    out = O.idct(O.dct(np.array([1.0, 2.0, 3.0, 4.0])))
    assert np.allclose(out, [1, 2, 3, 4], atol=1e-6)


def test_cf_to_sf_reference_kat(golden):
    """tests/convert-cf-to-sf.conf: CF_TO_SF(CF-F-He-Ar-1000.0) == SF-F-He-Ar-1000.0,
    frequencies within 1e-6 cm^-1 (src/driver.c:1857).  Values: the fixture prints 11
    digits; agreement is required to 1e-9 of the peak everywhere and to the reference's
    1e-4 pointwise wherever |SF| > 1e-8 of the peak (7 of 65536 points sit on a zero
    crossing 1e-10 below the peak, where FFT rounding alone exceeds 1e-4 pointwise)."""
    g = golden("cf_sf_hear_1000.npz")
    nu, sf = O.cf_to_sf(g["cf_t"], g["cf"])
    ref = g["sf"]
    assert np.max(np.abs(nu - g["sf_nu"])) < 1e-6
    peak = np.abs(ref).max()
    assert np.max(np.abs(sf - ref)) < 1e-9 * peak
    big = np.abs(ref) > 1e-8 * peak
    assert np.max(np.abs(sf[big] / ref[big] - 1.0)) < 1e-4


# ---------------------------------------------------------------- A13 M0
def test_m0_rejection_sampler_consistent_with_ensemble_cf0():
    """M0 = <mu.mu> ZeroCoeff (src/hawaii.c:1423-1491) equals CF[0] ZeroCoeff / ALU^3 of the
    same accepted ensemble (src/hawaii.c:3068)."""
    s = O.hear()
    smp = O.Sampler(O.IC_REJECTION, O.PAIR_FREE, 300.0, 4.0, 40.0, O.PESMIN_HEAR, 0.0)
    m, q = O.m0(s, smp, 5, 20000)
    ic, _, _ = O.sample_ics(s, smp, 5, 20000)
    L = O.lib()
    d = np.array([L.orc_hear_dip(r) for r in ic[:, 4]])
    ZeroCoeff = 0.00361479637 / (4.0 * math.pi)
    assert m == pytest.approx((d * d).mean() * ZeroCoeff, rel=1e-12)
    assert 0 < q < m


def test_zimmermann_series_sum_matches_numpy_correlate():
    """The Zimmermann sum over the first L time origins (src/hawaii.c:2089-2094) equals numpy's
    cross-correlation of the series with its first L samples, lags 0..L-1."""
    rng = np.random.default_rng(11)
    L = 37
    d = rng.normal(size=(2 * L - 1, 3))
    crln = np.zeros(L)
    O.lib().orc_zimmermann_series_cf(O.ptr(np.ascontiguousarray(d)), L, O.ptr(crln))
    ALU3 = 5.29177210903e-11 ** 3
    ref = np.zeros(L)
    for i in range(3):
        full = np.correlate(d[:, i], d[:L, i], mode="full")   # lag m at index m + L - 1
        ref += full[L - 1:2 * L - 1]
    np.testing.assert_allclose(crln, ref * ALU3 / L, rtol=1e-12, atol=1e-12 * np.abs(ref).max() * ALU3 / L)


def test_zimmermann_constant_dipole_is_flat():
    """A constant dipole over the whole series gives crln[s] = ALU^3 |d|^2 for every shift."""
    L = 16
    d = np.tile([0.3, -0.1, 0.2], (2 * L - 1, 1))
    crln = np.zeros(L)
    O.lib().orc_zimmermann_series_cf(O.ptr(np.ascontiguousarray(d)), L, O.ptr(crln))
    np.testing.assert_allclose(crln, 0.14 * 5.29177210903e-11 ** 3, rtol=1e-13)


# ---------------------------------------------------------------- M2 (src/hawaii.c:1524-1617)
def test_m2_point_hear_closed_form():
    """For a central dipole mu(R) r_hat, |d mu/dt|^2 = mu'(R)^2 Rdot^2 + mu(R)^2 (Thetadot^2 +
    sin^2 Theta Phidot^2), with the velocities of the kinetic energy (Rdot = pR/m, Thetadot =
    pTheta/(m R^2), Phidot = pPhi/(m R^2 sin^2 Theta)) and mu' by a 6th-order difference.  The
    oracle's central differences (h = 1e-3) agree to O(h^2)."""
    s = O.hear()
    L = O.lib()
    m = O.MU_HEAR
    rng = np.random.default_rng(5)
    for _ in range(20):
        R = rng.uniform(4.0, 15.0)
        Th = rng.uniform(0.2, math.pi - 0.2)
        Ph = rng.uniform(0, 2 * math.pi)
        pPh, pTh, pR = rng.normal(0, 30.0), rng.normal(0, 30.0), rng.normal(0, 5.0)
        y = np.array([Ph, pPh, Th, pTh, R, pR])
        f = O.dipole_rate_sq(s, y)[0]
        k = 1e-4
        dmu = (-L.orc_hear_dip(R + 3 * k) + 9 * L.orc_hear_dip(R + 2 * k) - 45 * L.orc_hear_dip(R + k)
               + 45 * L.orc_hear_dip(R - k) - 9 * L.orc_hear_dip(R - 2 * k) + L.orc_hear_dip(R - 3 * k)) / (-60 * k)
        mu = L.orc_hear_dip(R)
        Rd, Thd, Phd = pR / m, pTh / (m * R * R), pPh / (m * R * R * math.sin(Th) ** 2)
        expect = dmu ** 2 * Rd ** 2 + mu ** 2 * (Thd ** 2 + math.sin(Th) ** 2 * Phd ** 2)
        assert f == pytest.approx(expect, rel=1e-5)


def test_m2_point_co2ar_is_derivative_along_the_flow():
    """d mu/dt = the derivative of mu(q(t)) along Hamilton's flow: a Richardson-extrapolated
    directional difference of the dipole along qdot (coordinates only) at sampled states."""
    s = O.co2ar()
    L = O.lib()
    smp = O.Sampler(O.IC_REJECTION, O.PAIR_FREE, 200.0, 4.0, 20.0, O.PESMIN_CO2AR, 0.0)
    ic, _, _ = O.sample_ics(s, smp, 17, 12)
    for y in ic:
        yd = np.zeros(10)
        L.orc_rhs(C.byref(s), O.ptr(y), O.ptr(yd))
        qdot = yd[0::2]
        eps = 2e-3 / np.abs(qdot).max()

        def mu_at(t):
            q = np.ascontiguousarray(y[0::2] + t * qdot)
            out = np.zeros(3)
            L.orc_dipole(C.byref(s), O.ptr(q), O.ptr(out))
            return out

        d1 = (mu_at(eps) - mu_at(-eps)) / (2 * eps)
        d2 = (mu_at(eps / 2) - mu_at(-eps / 2)) / eps
        mudot = (4 * d2 - d1) / 3
        f = O.dipole_rate_sq(s, y)[0]
        assert f == pytest.approx(float(mudot @ mudot), rel=1e-5)


def test_m2_rejection_sampler_averages_the_accepted_ensemble():
    """calculate_M2 averages |d mu/dt|^2 over the same accepted points as the trajectory
    sampler draws (the candidate loop of src/hawaii.c:1546-1560), times SecondCoeff."""
    for s, smp in ((O.hear(), O.Sampler(O.IC_REJECTION, O.PAIR_FREE, 300.0, 4.0, 40.0, O.PESMIN_HEAR, 0.0)),
                   (O.co2ar(), O.Sampler(O.IC_REJECTION, O.PAIR_BOUND, 100.0, 4.0, 40.0, O.PESMIN_CO2AR, 0.0))):
        m, q = O.m2(s, smp, 9, 3000, 2.5)
        ic, _, _ = O.sample_ics(s, smp, 9, 3000)
        f = O.dipole_rate_sq(s, ic)
        SecondCoeff = 13856114.29344114
        assert m == pytest.approx(f.mean() * SecondCoeff * 2.5, rel=1e-12)
        assert q == pytest.approx(f.std(ddof=1) / math.sqrt(f.size) * SecondCoeff * 2.5, rel=1e-9)
