"""ctypes access to the CPU oracle (oracle/_build/liboracle.so) and to the reference-built
checkers (oracle/_ref/*.so).  Test infrastructure: the oracle is the checker, never the
thing under test."""
import ctypes as C
import os

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(REPO, "oracle", "_build", "liboracle.so")
REF_DIR = os.path.join(REPO, "oracle", "_ref")

SYS_HEAR, SYS_CO2AR, SYS_CH4CO2 = 0, 1, 3
IC_REJECTION, IC_R0_INJECT, IC_PHASE_SPACE = 0, 1, 2
PAIR_FREE, PAIR_BOUND, PAIR_ALL = 1, 2, 3
HALVES_BOTH, HALVES_FORWARD = 0, 1

# physical inputs of the example configs (example_configs/CF-F-He-Ar.conf:5,
# example_configs/CF-F-CO2-Ar.conf:5,42)
MU_HEAR = 6.632039160324215e3
MU_CO2AR = 3.817083069753919e4
I_CO2 = 2.819827114123505e5
PESMIN_HEAR = -9.424412560638486e-05
PESMIN_CO2AR = -8.913727688689150e-04
# example_configs/CF-F-CH4-CO2.conf:5, :27, :47, :52
MU_CH4CO2 = 2.141790360592430e4
I_CH4 = 2.093840965876434e4
PESMIN_CH4CO2 = -1.562522273622102e-03


class System(C.Structure):
    _fields_ = [("system", C.c_int), ("mu", C.c_double), ("II", C.c_double * 2), ("IR", C.c_double * 3)]


class Sampler(C.Structure):
    _fields_ = [("mode", C.c_int), ("pair_state", C.c_int), ("T", C.c_double), ("Rmin", C.c_double),
                ("Rmax", C.c_double), ("pesmin", C.c_double), ("R0", C.c_double)]


class CFParams(C.Structure):
    _fields_ = [("rtol", C.c_double), ("atol", C.c_double), ("dt", C.c_double), ("L", C.c_int),
                ("Rcut", C.c_double), ("halves", C.c_int), ("max_steps", C.c_long), ("zimmermann", C.c_int)]


class TrajStats(C.Structure):
    _fields_ = [("nrhs", C.c_long), ("ndip", C.c_long), ("naccept", C.c_long), ("nreject", C.c_long),
                ("e_drift", C.c_double)]


class OrcMlp(C.Structure):
    _fields_ = [("n_layers", C.c_int), ("dims", C.c_int * 5), ("act", C.c_int),
                ("W", C.c_void_p * 4), ("b", C.c_void_p * 4)]


class OrcN2ar(C.Structure):
    _fields_ = [("pes", OrcMlp), ("pes_xmean", C.c_void_p), ("pes_xscale", C.c_void_p),
                ("pes_ymean", C.c_double), ("pes_yscale", C.c_double), ("pes_en_inf", C.c_double),
                ("dip", (OrcMlp * 3) * 2), ("dip_xmean", C.c_void_p * 2), ("dip_xscale", C.c_void_p * 2)]


ACT_SILU, ACT_TANH = 0, 1


def orc_mlp(layers, act, keep):
    m = OrcMlp()
    m.n_layers = len(layers)
    m.dims[0] = layers[0][0].shape[0]
    for k, (W, b) in enumerate(layers):
        W = np.ascontiguousarray(W, dtype=np.float64)
        b = np.ascontiguousarray(b, dtype=np.float64)
        keep += [W, b]
        m.dims[k + 1] = W.shape[1]
        m.W[k] = W.ctypes.data
        m.b[k] = b.ctypes.data
    m.act = act
    return m


def n2ar_model(weights, en_inf):
    """The oracle's N2-Ar model over the weight arrays of hawaiihybrid_amd.n2ar.load_weights()
    (the arrays are data; the oracle does its own arithmetic).  Returns (model, keep-alive)."""
    keep = []
    m = OrcN2ar()
    layers, xm, xs, ym, ys = weights["pes"]
    m.pes = orc_mlp(layers, ACT_SILU, keep)
    xm, xs = np.ascontiguousarray(xm), np.ascontiguousarray(xs)
    keep += [xm, xs]
    m.pes_xmean, m.pes_xscale = xm.ctypes.data, xs.ctypes.data
    m.pes_ymean, m.pes_yscale, m.pes_en_inf = ym, ys, en_inf
    for k, (subnets, dxm, dxs) in enumerate(weights["dip"]):
        for s, sub in enumerate(subnets):
            m.dip[k][s] = orc_mlp(sub, ACT_SILU, keep)
        dxm, dxs = np.ascontiguousarray(dxm), np.ascontiguousarray(dxs)
        keep += [dxm, dxs]
        m.dip_xmean[k], m.dip_xscale[k] = dxm.ctypes.data, dxs.ctypes.data
    return m, keep


def hear():
    return System(SYS_HEAR, MU_HEAR, (C.c_double * 2)(0.0, 0.0))


def co2ar():
    return System(SYS_CO2AR, MU_CO2AR, (C.c_double * 2)(I_CO2, I_CO2))


def ch4co2():
    return System(SYS_CH4CO2, MU_CH4CO2, (C.c_double * 2)(I_CO2, I_CO2), (C.c_double * 3)(I_CH4, I_CH4, I_CH4))


def qp_size(sysid):
    return {SYS_CO2AR: 10, SYS_CH4CO2: 16}.get(sysid, 6)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        L = C.CDLL(ORACLE_SO)
        d, i, v, u32, sz = C.c_double, C.c_int, C.c_void_p, C.c_uint32, C.c_size_t
        P = C.POINTER
        sig = {
            "orc_mt_fill": (None, [u32, v, sz]),
            "orc_pes": (d, [P(System), v]),
            "orc_dpes": (None, [P(System), v, v]),
            "orc_dipole": (None, [P(System), v, v]),
            "orc_rhs": (None, [P(System), v, v]),
            "orc_kinetic": (d, [P(System), v]),
            "orc_hamiltonian": (d, [P(System), v]),
            "orc_hear_V": (d, [d]), "orc_hear_dV": (d, [d]), "orc_hear_dip": (d, [d]),
            "orc_co2ar_V_cm": (d, [d, d]),
            "orc_co2ar_dV_cm": (None, [d, d, v, v]),
            "orc_co2ar_dip_mol": (None, [d, d, v]),
            "orc_co2ar_lab_to_mol": (None, [v, v]),
            "orc_co2ar_jacobian": (None, [v, v, v]),
            "orc_ch4co2_lab_to_kal": (None, [v, v]),
            "orc_ch4co2_jacobian": (None, [v, v, v]),
            "orc_ch4co2_pes_kal": (d, [d, d, d, d, d]),
            "orc_ch4co2_dpes_kal": (None, [d, d, d, d, d, v]),
            "orc_ch4co2_well_scale": (d, []),
            "orc_ch4co2_dipole_kal": (None, [v, v]),
            "orc_sample_ics": (C.c_long, [P(System), P(Sampler), u32, C.c_long, v, P(C.c_long)]),
            "orc_half_cf": (i, [P(System), P(CFParams), v, v, P(C.c_int), P(TrajStats)]),
            "orc_traj_cf": (i, [P(System), P(CFParams), v, v, P(C.c_int), P(C.c_int), P(TrajStats)]),
            "orc_cf_rows": (None, [P(System), P(CFParams), v, C.c_long, i, v, v, v, v, v, v, v]),
            "orc_cf_ensemble": (None, [P(System), P(CFParams), v, C.c_long, i, d, v, P(C.c_long)]),
            "orc_fft_radix2": (None, [v, v, sz, i]),
            "orc_idct": (None, [v, sz, v]),
            "orc_dct": (None, [v, sz, v]),
            "orc_connes": (None, [v, sz, d]),
            "orc_cf_to_sf": (None, [v, v, sz, v, v]),
            "orc_m0": (None, [P(System), P(Sampler), u32, C.c_long, d, P(C.c_double), P(C.c_double)]),
            "orc_m2": (None, [P(System), P(Sampler), u32, C.c_long, d, P(C.c_double), P(C.c_double)]),
            "orc_dipole_rate_sq": (d, [P(System), v]),
            "orc_prmu_series": (i, [P(System), P(CFParams), v, v, P(C.c_int), P(TrajStats)]),
            "orc_prmu_power": (None, [v, sz, d, sz, v]),
            "orc_sf_moment": (d, [v, v, sz, d, i]),
            "orc_pf_analytic": (d, [P(System), d]),
            "orc_cf_moment": (i, [v, sz, d, d, i, P(C.c_double)]),
            "orc_cf_array_weights": (None, [v, C.c_long, d, i, d, i, v, v, v, v, v]),
            "orc_zimmermann_series_cf": (None, [v, i, v]),
            "orc_mlp_forward": (d, [P(OrcMlp), v, v]),
            "orc_pip_n2ar_pes": (None, [v, v]),
            "orc_pip_n2ar_dip": (None, [v, v]),
            "orc_n2ar_pes_xyz": (d, [P(OrcN2ar), v]),
            "orc_n2ar_dpes_xyz": (None, [P(OrcN2ar), v, v]),
            "orc_n2ar_pes": (d, [P(OrcN2ar), v]),
            "orc_n2ar_dpes": (None, [P(OrcN2ar), v, v]),
            "orc_n2ar_dipole_bf": (None, [P(OrcN2ar), d, d, v]),
            "orc_n2ar_dipole_model_bf": (None, [P(OrcN2ar), C.c_int, d, d, v]),
            "orc_n2ar_dipole": (None, [P(OrcN2ar), v, v]),
            "orc_n2ar_mlp_on_pes_features": (d, [P(OrcMlp), v, v, v]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def mt_stream(seed, n):
    out = np.zeros(n, dtype=np.uint32)
    lib().orc_mt_fill(seed, ptr(out), n)
    return out


def sample_ics(system, sampler, seed, ntraj):
    qp = qp_size(system.system)
    ic = np.zeros((ntraj, qp), dtype=np.float64)
    nd = C.c_long(0)
    ncand = lib().orc_sample_ics(C.byref(system), C.byref(sampler), seed, ntraj, ptr(ic), C.byref(nd))
    return ic, ncand, nd.value


def cf_rows(system, params, ic, nthreads=8):
    ic = np.ascontiguousarray(ic, dtype=np.float64)
    n = ic.shape[0]
    crln = np.zeros((n, params.L))
    status = np.zeros(n, dtype=np.int32)
    lf = np.zeros(n, dtype=np.int32)
    lb = np.zeros(n, dtype=np.int32)
    nrhs = np.zeros(n, dtype=np.int64)
    ndip = np.zeros(n, dtype=np.int64)
    ed = np.zeros(n)
    lib().orc_cf_rows(C.byref(system), C.byref(params), ptr(ic), n, nthreads, ptr(crln), ptr(status),
                      ptr(lf), ptr(lb), ptr(nrhs), ptr(ndip), ptr(ed))
    return dict(crln=crln, status=status, len_f=lf, len_b=lb, nrhs=nrhs, ndip=ndip, e_drift=ed)


def ensemble_from_rows(crln, ppf=1.0):
    """Sequential sum in trajectory order, then / ntraj (src/hawaii.c:3007-3009, 3111-3114)."""
    acc = np.zeros(crln.shape[1])
    for row in crln:
        acc += ppf * row
    return acc / crln.shape[0]


def hamiltonian(system, ic):
    """H at each initial point (rows of ic)."""
    L = lib()
    return np.array([L.orc_hamiltonian(C.byref(system), ptr(np.ascontiguousarray(r, dtype=np.float64))) for r in ic])


def cf_array_weights(E, T_base, pair_state, pesmin, temps, ratios=None, counted=None):
    """CORRELATION_ARRAY weights w [n][M] (weight x ratio) and nstar [M] (oracle/moments.c);
    counted[t] == 0 marks a trajectory the array loop skips (step limit)."""
    E = np.ascontiguousarray(E, dtype=np.float64)
    t = np.ascontiguousarray(temps, dtype=np.float64)
    r = None if ratios is None else np.ascontiguousarray(ratios, dtype=np.float64)
    c = None if counted is None else np.ascontiguousarray(counted, dtype=np.int32)
    w = np.zeros((len(E), len(t)))
    ns = np.zeros(len(t))
    lib().orc_cf_array_weights(ptr(E), len(E), T_base, pair_state, pesmin, len(t), ptr(t),
                               ptr(r) if r is not None else None, ptr(c) if c is not None else None, ptr(w), ptr(ns))
    return w, ns


def array_counted(status, halves_both=True):
    """1 where the array loop counts a trajectory: all but those whose last integrator call
    ended on the step limit without an earlier dipole failure (oracle orc_traj_cf's status)."""
    return (np.asarray(status) != 4).astype(np.int32)


def ensemble_array_from_rows(crln, w, nstar):
    """Per temperature: sum over trajectories, in order, of w[t][st] crln[t], / nstar[st]
    (src/hawaii.c:2666-2670, the writer divides by the weight sum)."""
    M = w.shape[1]
    acc = np.zeros((M, crln.shape[1]))
    for t, row in enumerate(crln):
        for st in range(M):
            acc[st] += w[t, st] * row
    return np.array([acc[st] / nstar[st] if nstar[st] > 0 else np.zeros(crln.shape[1]) for st in range(M)])


def idct(v):
    v = np.ascontiguousarray(v, dtype=np.float64)
    out = np.zeros_like(v)
    lib().orc_idct(ptr(v), v.size, ptr(out))
    return out


def dct(v):
    v = np.ascontiguousarray(v, dtype=np.float64)
    out = np.zeros_like(v)
    lib().orc_dct(ptr(v), v.size, ptr(out))
    return out


def cf_to_sf(t, cf):
    t = np.ascontiguousarray(t, dtype=np.float64)
    cf = np.array(cf, dtype=np.float64, copy=True)
    nu = np.zeros_like(cf)
    sf = np.zeros_like(cf)
    lib().orc_cf_to_sf(ptr(t), ptr(cf), cf.size, ptr(nu), ptr(sf))
    return nu, sf


def m0(system, sampler, seed, npoints, ppf=1.0):
    m = C.c_double(0)
    q = C.c_double(0)
    lib().orc_m0(C.byref(system), C.byref(sampler), seed, npoints, ppf, C.byref(m), C.byref(q))
    return m.value, q.value


def m2(system, sampler, seed, npoints, ppf=1.0):
    m = C.c_double(0)
    q = C.c_double(0)
    lib().orc_m2(C.byref(system), C.byref(sampler), seed, npoints, ppf, C.byref(m), C.byref(q))
    return m.value, q.value


def dipole_rate_sq(system, y):
    """|d mu/dt|^2 at each full state (rows of y)."""
    L = lib()
    return np.array([L.orc_dipole_rate_sq(C.byref(system), ptr(np.ascontiguousarray(r, dtype=np.float64)))
                     for r in np.atleast_2d(y)])


def prmu_series(system, y0, dt, L, R0, rtol=1e-12):
    """The PR/mu dipole series of one trajectory: (status, len, vec [len][3]) with vec[k] the
    dipole at t = k dt, up to the first sample beyond R0, at most L samples."""
    p = CFParams(rtol, rtol, dt, L, R0, HALVES_FORWARD, 50000, 0)
    vec = np.zeros(3 * L)
    n = C.c_int(0)
    st = TrajStats()
    status = lib().orc_prmu_series(C.byref(system), C.byref(p), ptr(np.ascontiguousarray(y0, dtype=np.float64)),
                                   ptr(vec), C.byref(n), C.byref(st))
    return status, n.value, vec.reshape(L, 3)[:n.value].copy()


def prmu_power(x, dt, nfreq):
    """sum_c |FFT(Connes(x_c))_k|^2 for x [3][L]."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.zeros(nfreq)
    lib().orc_prmu_power(ptr(x), x.shape[1], dt, nfreq, ptr(out))
    return out


def ref_lib(name):
    """A reference-built checker from oracle/_ref (None where it has not been built)."""
    path = os.path.join(REF_DIR, name)
    return C.CDLL(path) if os.path.exists(path) else None
