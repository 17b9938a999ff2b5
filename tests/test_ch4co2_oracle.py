"""CH4-CO2 (ROTOR + LINEAR_MOLECULE) in the oracle, pinned by what the reference and the
mathematics fix:

* the dipole in the CH4 frame (PES-IDS/ai_ids_ch4_co2.cpp:672-1066) against the reference's own
  expressions evaluated term by term at 200 points with scipy's associated Legendre functions
  (tests/golden/ch4co2_dipole_ref.npz, written by hawaiihybrid_amd/data/gen_ch4co2_tables.py):
  a dropped term, a wrong radial interval or a Legendre sign convention fails it;
* the lab dipole rotates with the whole complex: rotating the lab frame by G maps
  mu_lab(q) to G mu_lab(q) (the CH4-frame angles and the S1^T back-rotation of
  PES-IDS/ai_ids_ch4_co2_lib.cpp:8-34 both enter);
* the CH4-frame transform (src/angles_handler.cpp:357-419) puts the intermolecular vector and the
  CO2 axis at the right angles (closed form with CH4 unrotated), and its 8 x 5 Jacobian
  (src/angles_handler.cpp:421-555) equals central differences of the transform;
* the stand-in PES's five derivatives equal differences of its value, and its minimum is the
  documented -320 cm^-1;
* Hamilton's equations of the rotor (src/hawaii.c:505-531) and the linear molecule are
  J grad H with H = kinetic_energy (src/hawaii.c:1041-1137) + V, by 6th-order differences;
* the rotor momenta of p_generator_rotor (src/hawaii.c:1305-1327) have the Boltzmann moments of
  a rigid rotor: <K_rot> = 3/2 kT for the rotor and kT for the linear molecule.
"""
import ctypes as C
import os

import numpy as np
import pytest

import oracle_lib as O

HERE = os.path.dirname(os.path.abspath(__file__))
HTOCM = 2.1947463136320e5


def arr(*v):
    return np.ascontiguousarray(np.array(v, dtype=np.float64))


@pytest.fixture(scope="module")
def golden():
    return np.load(os.path.join(HERE, "golden", "ch4co2_dipole_ref.npz"))


def test_dipole_in_ch4_frame_matches_reference_expressions(golden):
    L = O.lib()
    d = np.zeros(3)
    got = []
    for i in range(len(golden["R"])):
        q = arr(golden["R"][i], golden["phi1"][i], golden["theta1"][i], golden["phi2"][i], golden["theta2"][i])
        L.orc_ch4co2_dipole_kal(O.ptr(q), O.ptr(d))
        got.append(d.copy())
    got = np.array(got)
    ref = golden["dipole"]
    scale = np.max(np.abs(ref))
    assert np.max(np.abs(got - ref)) < 1e-10 * scale
    # the tail (R >= 30) is a separate radial form: it is sampled and matches too
    tail = golden["R"] >= 30.0
    assert tail.sum() >= 10
    assert np.max(np.abs(got[tail] - ref[tail])) < 1e-10 * np.max(np.abs(ref[tail]))


def lab_vectors(q):
    """R direction, CO2 axis and the CH4 rotation S1 in the lab, with the reference's matrices
    (Sz, Sx, the non-standard Sy of src/angles_handler.cpp:45-104)."""
    def Sz(a):
        s, c = np.sin(a), np.cos(a)
        return np.array([[c, s, 0], [-s, c, 0], [0, 0, 1.0]])

    def Sx(a):
        s, c = np.sin(a), np.cos(a)
        return np.array([[1.0, 0, 0], [0, c, s], [0, -s, c]])

    def Sy(a):
        s, c = np.sin(a), np.cos(a)
        return np.array([[c, 0, s], [0, 1.0, 0], [-s, 0, c]])

    z = np.array([0, 0, 1.0])
    Rh = Sz(q[0]).T @ Sy(q[1]).T @ z
    u = Sz(q[6]).T @ Sy(q[7]).T @ z
    S1 = Sz(q[5]) @ Sx(q[4]) @ Sz(q[3])
    return Rh, u, S1


def angles_from(Rh, u, S1):
    """Lab coordinates q (R set separately) whose vectors are Rh, u and whose CH4 rotation is S1."""
    # Rh = (-sin T cos P, -sin T sin P, cos T); u likewise
    T = np.arccos(np.clip(Rh[2], -1, 1))
    P = np.arctan2(-Rh[1], -Rh[0])
    t2 = np.arccos(np.clip(u[2], -1, 1))
    p2 = np.arctan2(-u[1], -u[0])
    # S1 = Sz(psi) Sx(theta) Sz(phi): S1[2] = (sin th sin ph, -sin th cos ph, cos th),
    # S1[:, 2] = (sin psi sin th, cos psi sin th, cos th)
    th = np.arccos(np.clip(S1[2, 2], -1, 1))
    ph = np.arctan2(S1[2, 0], -S1[2, 1])
    ps = np.arctan2(S1[0, 2], S1[1, 2])
    return P, T, ph, th, ps, p2, t2


def test_lab_dipole_rotates_with_the_complex():
    rng = np.random.default_rng(5)
    s = O.ch4co2()
    mu1, mu2 = np.zeros(3), np.zeros(3)
    for _ in range(40):
        q = arr(rng.uniform(0, 2 * np.pi), np.arccos(rng.uniform(-0.95, 0.95)), rng.uniform(5.0, 12.0),
                rng.uniform(0, 2 * np.pi), np.arccos(rng.uniform(-0.95, 0.95)), rng.uniform(0, 2 * np.pi),
                rng.uniform(0, 2 * np.pi), np.arccos(rng.uniform(-0.95, 0.95)))
        Rh, u, S1 = lab_vectors(q)
        # a random proper rotation G of the lab frame
        A = rng.standard_normal((3, 3))
        G, _ = np.linalg.qr(A)
        if np.linalg.det(G) < 0:
            G[:, 0] *= -1
        P, T, ph, th, ps, p2, t2 = angles_from(G @ Rh, G @ u, S1 @ G.T)
        q2 = arr(P, T, q[2], ph, th, ps, p2, t2)
        # the rotated configuration's vectors are the rotated vectors
        Rh2, u2, S12 = lab_vectors(q2)
        assert np.allclose(Rh2, G @ Rh, atol=1e-12) and np.allclose(S12, S1 @ G.T, atol=1e-12)
        O.lib().orc_dipole(C.byref(s), O.ptr(q), O.ptr(mu1))
        O.lib().orc_dipole(C.byref(s), O.ptr(q2), O.ptr(mu2))
        assert np.allclose(mu2, G @ mu1, atol=1e-11 * max(1.0, np.abs(mu1).max()))
        # and the potential is invariant
        v1 = O.lib().orc_pes(C.byref(s), O.ptr(q))
        v2 = O.lib().orc_pes(C.byref(s), O.ptr(q2))
        assert abs(v1 - v2) < 1e-12 * max(1e-6, abs(v1))


def test_kal_frame_closed_form_and_jacobian_by_differences():
    L = O.lib()
    rng = np.random.default_rng(2)
    k = np.zeros(5)
    # CH4 unrotated (S1 = 1): the kal angles are the lab directions' polar angles
    q = arr(0.7, 1.1, 6.5, 0.0, 0.0, 0.0, 2.1, 0.4)
    L.orc_ch4co2_lab_to_kal(O.ptr(q), O.ptr(k))
    Rh, u, _ = lab_vectors(q)
    assert abs(k[0] - 6.5) == 0.0
    assert abs(np.cos(k[2]) - Rh[2]) < 1e-14 and abs(np.arctan2(Rh[1], Rh[0]) - k[1]) % (2 * np.pi) < 1e-12
    assert abs(np.cos(k[4]) - u[2]) < 1e-14
    jac = np.zeros((8, 5))
    for _ in range(20):
        q = arr(rng.uniform(0, 2 * np.pi), np.arccos(rng.uniform(-0.9, 0.9)), rng.uniform(5, 10),
                rng.uniform(0, 2 * np.pi), np.arccos(rng.uniform(-0.9, 0.9)), rng.uniform(0, 2 * np.pi),
                rng.uniform(0, 2 * np.pi), np.arccos(rng.uniform(-0.9, 0.9)))
        L.orc_ch4co2_lab_to_kal(O.ptr(q), O.ptr(k))
        L.orc_ch4co2_jacobian(O.ptr(q), O.ptr(k), O.ptr(jac))
        fd = np.zeros((8, 5))
        h = 1e-6
        for i in range(8):
            qp, qm = q.copy(), q.copy()
            qp[i] += h
            qm[i] -= h
            kp, km = np.zeros(5), np.zeros(5)
            L.orc_ch4co2_lab_to_kal(O.ptr(qp), O.ptr(kp))
            L.orc_ch4co2_lab_to_kal(O.ptr(qm), O.ptr(km))
            dk = kp - km
            dk[[1, 3]] = (dk[[1, 3]] + np.pi) % (2 * np.pi) - np.pi   # azimuths wrap
            fd[i] = dk / (2 * h)
        assert np.max(np.abs(jac - fd)) < 1e-6


def test_standin_pes_derivatives_and_well():
    L = O.lib()
    rng = np.random.default_rng(8)
    d = np.zeros(5)
    for _ in range(30):
        x = [rng.uniform(5.5, 12.0), rng.uniform(0, 2 * np.pi), np.arccos(rng.uniform(-0.95, 0.95)),
             rng.uniform(0, 2 * np.pi), np.arccos(rng.uniform(-0.95, 0.95))]
        L.orc_ch4co2_dpes_kal(*x, O.ptr(d))
        for i in range(5):
            h = 1e-5
            xp, xm = list(x), list(x)
            xp[i] += h
            xm[i] -= h
            fd = (L.orc_ch4co2_pes_kal(*xp) - L.orc_ch4co2_pes_kal(*xm)) / (2 * h)
            assert abs(fd - d[i]) < 1e-7 * max(1e-5, abs(d[i]))
    # the minimum over a grid refined by a local search is the documented well depth, above the
    # reference's sampling bound pesmin
    from scipy.optimize import minimize
    best = None
    for R in (6.0, 6.5, 7.0, 7.5):
        for t1 in np.linspace(0.2, 2.9, 6):
            for t2 in np.linspace(0.2, 2.9, 6):
                for p1 in np.linspace(0, 2 * np.pi, 6, endpoint=False):
                    v = L.orc_ch4co2_pes_kal(R, p1, t1, 0.3, t2)
                    if best is None or v < best[0]:
                        best = (v, [R, p1, t1, 0.3, t2])
    res = minimize(lambda x: L.orc_ch4co2_pes_kal(*x) * HTOCM, best[1], method="Nelder-Mead",
                   options={"xatol": 1e-9, "fatol": 1e-9, "maxiter": 20000})
    assert abs(res.fun - (-320.0)) < 0.5
    assert res.fun * 1.0 / HTOCM > O.PESMIN_CH4CO2


def test_hamilton_equations_are_the_gradient_of_H():
    L = O.lib()
    s = O.ch4co2()
    rng = np.random.default_rng(3)
    for _ in range(10):
        y = np.zeros(16)
        y[0::2] = [rng.uniform(0, 6), np.arccos(rng.uniform(-0.9, 0.9)), rng.uniform(6, 10), rng.uniform(0, 6),
                   np.arccos(rng.uniform(-0.9, 0.9)), rng.uniform(0, 6), rng.uniform(0, 6),
                   np.arccos(rng.uniform(-0.9, 0.9))]
        y[1::2] = rng.standard_normal(8) * np.array([50, 50, 20, 5, 5, 5, 20, 20])
        yd = np.zeros(16)
        L.orc_rhs(C.byref(s), O.ptr(y), O.ptr(yd))
        grad = np.zeros(16)
        for i in range(16):
            h = 1e-4 * max(1.0, abs(y[i]))
            fs = []
            for m in (-3, -2, -1, 1, 2, 3):
                yy = y.copy()
                yy[i] += m * h
                fs.append(L.orc_hamiltonian(C.byref(s), O.ptr(yy)))
            grad[i] = (-fs[0] + 9 * fs[1] - 45 * fs[2] + 45 * fs[3] - 9 * fs[4] + fs[5]) / (60 * h)
        # dq/dt = dH/dp, dp/dt = -dH/dq
        scale = np.abs(grad).max()
        assert np.max(np.abs(yd[0::2] - grad[1::2])) < 1e-6 * scale
        assert np.max(np.abs(yd[1::2] + grad[0::2])) < 1e-6 * scale


def test_rotor_sampler_boltzmann_moments():
    """R0 injection at T: <K_rotor> = 3/2 kT (three rotational degrees of freedom), <K_linear> = kT;
    the sampled Euler angles cover their ranges uniformly (cos theta uniform)."""
    s = O.ch4co2()
    smp = O.Sampler(O.IC_R0_INJECT, O.PAIR_FREE, 300.0, 4.75, 40.0, O.PESMIN_CH4CO2, 40.0)
    n = 20000
    ic = np.zeros((n, 16))
    nd = C.c_long(0)
    O.lib().orc_sample_ics(C.byref(s), C.byref(smp), 7, n, O.ptr(ic), C.byref(nd))
    # draws: 3 + 3 + 2 (q) + 3 + 3 + 2 normals of two draws each (+ Box-Muller retries, rare)
    assert nd.value >= n * (8 + 2 * 8)
    kT = 300.0 / (4.3597447222071e-18 / 1.380649e-23)
    I = O.I_CH4
    th, ps = ic[:, 8], ic[:, 10]
    t1 = (ic[:, 7] - ic[:, 11] * np.cos(th)) * np.sin(ps) + ic[:, 9] * np.sin(th) * np.cos(ps)
    t2 = (ic[:, 7] - ic[:, 11] * np.cos(th)) * np.cos(ps) - ic[:, 9] * np.sin(th) * np.sin(ps)
    krot = (t1 ** 2 + t2 ** 2) / (2 * I * np.sin(th) ** 2) + ic[:, 11] ** 2 / (2 * I)
    klin = ic[:, 15] ** 2 / (2 * O.I_CO2) + ic[:, 13] ** 2 / (2 * O.I_CO2 * np.sin(ic[:, 14]) ** 2)
    assert abs(krot.mean() / (1.5 * kT) - 1) < 0.03
    assert abs(klin.mean() / kT - 1) < 0.03
    for a in (np.cos(ic[:, 8]), np.cos(ic[:, 14])):
        h, _ = np.histogram(a, bins=10, range=(-1, 1))
        assert np.all(np.abs(h / (n / 10) - 1) < 0.1)
    assert np.all(ic[:, 5] <= 0) and np.all(ic[:, 4] == 40.0)


def test_host_sampler_matches_oracle_bit_for_bit():
    """The engine's host sampler (hh_sample_ics) for CH4-CO2 against the oracle on 20000 R0-injected
    initial conditions: the same draws in the same order, and p_generator_rotor's sincos (whose
    cosine differs from cos in the last bit for some angles, src/hawaii.c:1317-1318)."""
    import hawaiihybrid_amd as H
    n = 20000
    smp = O.Sampler(O.IC_R0_INJECT, O.PAIR_FREE, 300.0, 4.75, 40.0, O.PESMIN_CH4CO2, 40.0)
    ic = np.zeros((n, 16))
    nd = C.c_long(0)
    O.lib().orc_sample_ics(C.byref(O.ch4co2()), C.byref(smp), 42, n, O.ptr(ic), C.byref(nd))
    sm = H.Sampler(H.IC_R0_INJECT, 1, 300.0, 4.75, 40.0, H.PESMIN_CH4CO2, 40.0)
    ic2, _, nd2 = H.sample_ics(H.ch4_co2(), sm, 42, n)
    assert np.array_equal(ic, ic2)
    assert nd2 == nd.value
