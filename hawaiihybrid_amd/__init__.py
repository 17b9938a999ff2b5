"""hawaiihybrid_amd: MI355X trajectory-ensemble engine (Python side of libhawaii_gpu.so).

The compute path is the C ABI of include/hawaii_gpu.h; this module only binds it with
ctypes and uses torch for device memory.  There is no CPU fallback: if the library is not
built, or no gfx950 device is visible, the calls raise.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libhawaii_gpu.so")

SYS_HEAR, SYS_CO2AR, SYS_N2AR, SYS_CH4CO2 = 0, 1, 2, 3
IC_REJECTION, IC_R0_INJECT, IC_PHASE_SPACE = 0, 1, 2
PAIR_FREE, PAIR_BOUND, PAIR_ALL = 1, 2, 3
ENGINE_PLAIN, ENGINE_BATCHED = 0, 1
PREC_FP64, PREC_MIXED = 0, 1
HALVES_BOTH, HALVES_FORWARD = 0, 1
EVAL_PES, EVAL_DIPOLE, EVAL_HAMILTONIAN, EVAL_RHS, EVAL_DIPOLE_RATE = 0, 1, 2, 3, 4

# example_configs/CF-F-He-Ar.conf:5, example_configs/CF-F-CO2-Ar.conf:5,27,42
MU_HEAR = 6.632039160324215e3
MU_CO2AR = 3.817083069753919e4
I_CO2 = 2.819827114123505e5
PESMIN_HEAR = -9.424412560638486e-05
PESMIN_CO2AR = -8.913727688689150e-04
# src/constants.h:36-98: RAMTOAMU, m_N, m_Ar, L_N2, II_N2 = m_N/2 L_N2^2
_RAMTOAMU = 1822.888485332
_M_N = 14.003074004460 * _RAMTOAMU
_M_AR = 39.9623831237 * _RAMTOAMU
L_N2 = 2.07856749
MU_N2AR = 2.0 * _M_N * _M_AR / (2.0 * _M_N + _M_AR)
I_N2 = _M_N / 2.0 * L_N2 * L_N2
PESMIN_N2AR = -97.3 / 2.1947463136320e5     # examples/correlation_n2_ar.cpp:214


class System(C.Structure):
    _fields_ = [("system", C.c_int), ("mu", C.c_double), ("I", C.c_double), ("I_rotor", C.c_double * 3)]


class Sampler(C.Structure):
    _fields_ = [("mode", C.c_int), ("pair_state", C.c_int), ("T", C.c_double), ("Rmin", C.c_double),
                ("Rmax", C.c_double), ("pesmin", C.c_double), ("R0", C.c_double)]


class CFParams(C.Structure):
    _fields_ = [("rtol", C.c_double), ("atol", C.c_double), ("dt", C.c_double), ("Rcut", C.c_double),
                ("ppf", C.c_double), ("L", C.c_int), ("halves", C.c_int), ("max_steps", C.c_int),
                ("engine", C.c_int), ("precision", C.c_int), ("block", C.c_int), ("grid", C.c_int),
                ("zimmermann", C.c_int), ("orbit_rtol", C.c_double)]


class CFArrayParams(C.Structure):
    """hh_cf_array_params: satellite temperatures of a CORRELATION_ARRAY run."""
    _fields_ = [("T_base", C.c_double), ("pair_state", C.c_int), ("ntemp", C.c_int), ("pesmin", C.c_double),
                ("temps", C.c_void_p), ("ratios", C.c_void_p)]


class PrmuParams(C.Structure):
    """hh_prmu_params: the PR/mu spectral function's trajectory and spectrum settings."""
    _fields_ = [("rtol", C.c_double), ("atol", C.c_double), ("dt", C.c_double), ("R0", C.c_double),
                ("L", C.c_int), ("nfreq", C.c_int), ("precision", C.c_int), ("max_steps", C.c_int),
                ("allow_truncating", C.c_int), ("orbit_rtol", C.c_double)]


class CFStats(C.Structure):
    _fields_ = [("integrate_ms", C.c_double), ("reduce_ms", C.c_double), ("n_items", C.c_longlong),
                ("nrhs", C.c_longlong), ("nsamples", C.c_longlong), ("n_failed", C.c_longlong),
                ("n_stepfail", C.c_longlong), ("max_e_drift", C.c_double), ("block", C.c_int), ("grid", C.c_int),
                ("launches", C.c_int), ("n_orbit", C.c_longlong), ("nrhs_orbit", C.c_longlong),
                ("nsamples_orbit", C.c_longlong), ("pool_launches", C.c_int), ("pool_kernel_ms", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class SessionOpts(C.Structure):
    _fields_ = [("slice_ms", C.c_double), ("max_device_gib", C.c_double), ("block", C.c_int)]


class RequestInfo(C.Structure):
    _fields_ = [("kind", C.c_int), ("state", C.c_int), ("submit_ms", C.c_double), ("admit_ms", C.c_double),
                ("done_ms", C.c_double), ("latency_ms", C.c_double), ("trajectories", C.c_longlong),
                ("cf", CFStats), ("m0", C.c_double), ("m0_err", C.c_double), ("error", C.c_char * 160)]


class VegasParams(C.Structure):
    _fields_ = [("integrand", C.c_int), ("pair_state", C.c_int), ("T", C.c_double), ("Rmin", C.c_double),
                ("Rmax", C.c_double), ("niterations", C.c_int), ("npoints", C.c_long), ("seed", C.c_uint32)]


class VegasResult(C.Structure):
    _fields_ = [("value", C.c_double), ("error", C.c_double), ("chi2_dof", C.c_double), ("iterations", C.c_int),
                ("points", C.c_longlong), ("nonfinite", C.c_longlong), ("ms", C.c_double),
                ("iter_value", C.c_double * 32), ("iter_error", C.c_double * 32)]


INTEGRAND_M0, INTEGRAND_M2, INTEGRAND_PF = 0, 1, 2
ZERO_COEFF = 0.00361479637 / (4.0 * 3.141592653589793)   # src/constants.h:29
SECOND_COEFF = 13856114.29344114                          # src/constants.h:30

REQ_WAITING, REQ_RUNNING, REQ_DONE, REQ_FAILED = 0, 1, 2, 3
REQ_CF, REQ_M0 = 0, 1


def he_ar():
    return System(SYS_HEAR, MU_HEAR, 0.0)


def co2_ar():
    return System(SYS_CO2AR, MU_CO2AR, I_CO2)


# example_configs/CF-F-CH4-CO2.conf:5, :27, :47 (CH4, a spherical top), :52 (CO2)
MU_CH4CO2 = 2.141790360592430e4
I_CH4 = 2.093840965876434e4
PESMIN_CH4CO2 = -1.562522273622102e-03


def ch4_co2():
    """CH4-CO2 (ROTOR + LINEAR_MOLECULE): mu, I of CO2 and the principal moments of CH4."""
    return System(SYS_CH4CO2, MU_CH4CO2, I_CO2, (C.c_double * 3)(I_CH4, I_CH4, I_CH4))


def n2_ar():
    """N2-Ar (examples/correlation_n2_ar.cpp:184-186): mu = m_N2 m_Ar / (m_N2 + m_Ar), I = II_N2."""
    return System(SYS_N2AR, MU_N2AR, I_N2)


def qp_size(system):
    return {SYS_HEAR: 6, SYS_CO2AR: 10, SYS_N2AR: 10, SYS_CH4CO2: 16}[system]


class HHError(RuntimeError):
    pass


_lib = None


def lib():
    """The loaded engine library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise HHError("libhawaii_gpu.so is not built: run `python -m hawaiihybrid_amd.build`")
        # PyTorch ships its own HIP/HSA runtime under a different soname; loaded after ours it
        # would start a second runtime that finds no GPU.  Loading torch first makes the process
        # share torch's runtime (the engine only needs one).
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(LIB_PATH)
        v, i, d, P = C.c_void_p, C.c_int, C.c_double, C.POINTER
        sigs = {
            "hh_last_error": (C.c_char_p, []),
            "hh_qp_size": (i, [i]),
            "hh_mt_fill": (None, [C.c_uint32, v, C.c_size_t]),
            "hh_sample_ics": (C.c_long, [P(System), P(Sampler), C.c_uint32, C.c_long, v, v, P(C.c_long)]),
            "hh_sample_ics_from": (C.c_long, [P(System), P(Sampler), C.c_uint32, C.c_long, C.c_long, v, v,
                                              P(C.c_long)]),
            "hh_sf_moment": (i, [v, v, C.c_long, d, i, P(d)]),
            "hh_engine_create": (i, [i, P(v)]),
            "hh_engine_destroy": (None, [v]),
            "hh_engine_info": (i, [v, C.c_char_p, C.c_size_t]),
            "hh_cf_run": (i, [v, P(System), P(CFParams), i, v, v, v, P(CFStats)]),
            "hh_cf_items": (i, [v, i, v, v, v, v, v]),
            "hh_cf_to_sf": (i, [v, v, C.c_long, d, v, v, P(d)]),
            "hh_cf_to_sf_noalloc": (i, [v, v, C.c_long, d, v, v]),
            "hh_m0": (i, [v, P(System), P(Sampler), C.c_uint32, C.c_long, i, d, P(d), P(d), P(d)]),
            "hh_m2": (i, [v, P(System), P(Sampler), C.c_uint32, C.c_long, i, d, P(d), P(d), P(d)]),
            "hh_eval": (i, [v, P(System), i, v, C.c_long, v]),
            "hh_cf_array": (i, [v, P(System), P(CFParams), P(CFArrayParams), v, C.c_long, v, v, P(CFStats)]),
            "hh_prmu": (i, [v, P(System), P(PrmuParams), v, C.c_long, v, v, P(C.c_long), P(CFStats)]),
            "hh_conf_parse": (i, [C.c_char_p, v]),
            "hh_session_create": (i, [v, P(SessionOpts), P(v)]),
            "hh_submit_cf": (i, [v, P(System), P(CFParams), v, C.c_long, v, P(C.c_uint64)]),
            "hh_submit_m0": (i, [v, P(System), P(Sampler), C.c_uint32, C.c_long, i, d, P(C.c_uint64)]),
            "hh_poll": (i, [v, C.c_uint64, d, P(RequestInfo)]),
            "hh_session_destroy": (i, [v]),
            "hh_pf_analytic": (d, [P(System), d]),
            "hh_vegas": (i, [v, P(System), P(VegasParams), P(VegasResult)]),
            "hh_cf_moment": (i, [v, C.c_long, d, d, i, P(d)]),
            "hh_fit_baseline": (i, [v, v, C.c_long, C.c_long, v]),
            "hh_loess": (i, [v, v, C.c_long, d, C.c_long, C.c_long, d, C.c_long, v, v]),
            "hh_pipnn_create": (i, [v, v, i, P(v)]),
            "hh_pipnn_eval": (i, [v, v, C.c_long, v, v, P(d)]),
            "hh_pipnn_destroy": (None, [v]),
        }
        for name, (res, args) in sigs.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _check(rc):
    if rc < 0:
        raise HHError(lib().hh_last_error().decode())
    return rc


def mt_stream(seed, n):
    out = np.zeros(n, dtype=np.uint32)
    lib().hh_mt_fill(seed, out.ctypes.data_as(C.c_void_p), n)
    return out


def sample_ics(system, sampler, seed, ntraj, pes=None):
    """Initial conditions [ntraj][QP] (host, bit-exact reference stream).  `pes` is an optional
    plugin pes_lab address (int) used by the rejection step."""
    ic = np.zeros((ntraj, qp_size(system.system)), dtype=np.float64)
    nd = C.c_long(0)
    n = _check(lib().hh_sample_ics(C.byref(system), C.byref(sampler), seed, ntraj, pes,
                                   ic.ctypes.data_as(C.c_void_p), C.byref(nd)))
    return ic, n, nd.value


def sample_ics_from(system, sampler, seed, skip_draws, ntraj, pes=None):
    """sample_ics continuing the stream after its first skip_draws draws; the returned draw
    count includes the skipped ones."""
    ic = np.zeros((ntraj, qp_size(system.system)), dtype=np.float64)
    nd = C.c_long(0)
    n = _check(lib().hh_sample_ics_from(C.byref(system), C.byref(sampler), seed, skip_draws, ntraj, pes,
                                        ic.ctypes.data_as(C.c_void_p), C.byref(nd)))
    return ic, n, nd.value


def sf_moment(nu, sf, n, ntraj=0.0):
    """compute_Mn_from_sf_using_classical_detailed_balance (hh_sf_moment)."""
    nu = np.ascontiguousarray(nu, dtype=np.float64)
    sf = np.ascontiguousarray(sf, dtype=np.float64)
    out = C.c_double(0)
    _check(lib().hh_sf_moment(nu.ctypes.data_as(C.c_void_p), sf.ctypes.data_as(C.c_void_p), nu.size, float(ntraj), n,
                              C.byref(out)))
    return out.value


def pf_analytic(system, T):
    """analytic_full_partition_function_by_V (hh_pf_analytic)."""
    v = lib().hh_pf_analytic(C.byref(system), float(T))
    if v < 0:
        raise HHError(lib().hh_last_error().decode())
    return v


def cf_moment(cf, dt, n, ntraj=0.0):
    """M0 (n = 0) or M2 (n = 2) of a host CF (hh_cf_moment); ntraj > 0 for an unnormalised CF."""
    cf = np.ascontiguousarray(cf, dtype=np.float64)
    out = C.c_double()
    _check(lib().hh_cf_moment(cf.ctypes.data_as(C.c_void_p), len(cf), float(dt), float(ntraj), int(n), C.byref(out)))
    return out.value


def fit_baseline(t, cf, begin=16384):
    """FIT_BASELINE's wing-model fit (hh_fit_baseline): (A, B, C) of C + A / (1 + B^2 t^2)."""
    t = np.ascontiguousarray(t, dtype=np.float64)
    cf = np.ascontiguousarray(cf, dtype=np.float64)
    abc = np.zeros(3)
    _check(lib().hh_fit_baseline(t.ctypes.data_as(C.c_void_p), cf.ctypes.data_as(C.c_void_p), len(t), int(begin),
                                 abc.ctypes.data_as(C.c_void_p)))
    return tuple(abc)


def loess(x, y, grid_max, npoints, ws_min=21, ws_step=0.1, ws_delay=10):
    """SMOOTH's LOESS smoothing (hh_loess) on npoints points from 0 to grid_max: (grid, values)."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    g, out = np.zeros(npoints), np.zeros(npoints)
    _check(lib().hh_loess(x.ctypes.data_as(C.c_void_p), y.ctypes.data_as(C.c_void_p), len(x), float(grid_max),
                          int(npoints), int(ws_min), float(ws_step), int(ws_delay), g.ctypes.data_as(C.c_void_p),
                          out.ctypes.data_as(C.c_void_p)))
    return g, out


def cf_params(L, dt=200.0, rtol=1e-12, atol=None, Rcut=40.0, ppf=1.0, halves=HALVES_FORWARD, engine=ENGINE_BATCHED,
              precision=PREC_FP64, max_steps=50000, block=0, grid=0, zimmermann=False, orbit_rtol=0.0):
    """hh_cf_params; orbit_rtol: mixed precision hands orbiting trajectories over to FP64 at this
    tolerance (0: 1e-12; negative: off)."""
    return CFParams(rtol, rtol if atol is None else atol, dt, Rcut, ppf, L, halves, max_steps, engine, precision,
                    block, grid, 1 if zimmermann else 0, orbit_rtol)


class Engine:
    """One device of the engine (HIP stream, surface tables, workspaces)."""

    def __init__(self, device=0):
        import torch  # device memory for callers; the engine itself only needs HIP
        self.torch = torch
        self.device = device
        h = C.c_void_p()
        _check(lib().hh_engine_create(device, C.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            lib().hh_engine_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self):
        import json
        buf = C.create_string_buffer(512)
        _check(lib().hh_engine_info(self.h, buf, 512))
        return json.loads(buf.value.decode())

    def to_device_soa(self, ic):
        """[ntraj][QP] host array -> device tensor SoA [QP][ntraj] (float64)."""
        t = self.torch.from_numpy(np.ascontiguousarray(ic.T)).to(f"cuda:{self.device}")
        return t.contiguous()

    def cf_run(self, system, params, ics_dev, out=None):
        """ics_dev: list of device tensors SoA [QP][ntraj].  Returns (list of CF tensors, stats)."""
        torch = self.torch
        nj = len(ics_dev)
        if out is None:
            out = [torch.zeros(params.L, dtype=torch.float64, device=f"cuda:{self.device}") for _ in range(nj)]
        icp = (C.c_void_p * nj)(*[t.data_ptr() for t in ics_dev])
        ntr = (C.c_long * nj)(*[t.shape[1] for t in ics_dev])
        cfp = (C.c_void_p * nj)(*[t.data_ptr() for t in out])
        st = CFStats()
        torch.cuda.synchronize(self.device)
        _check(lib().hh_cf_run(self.h, C.byref(system), C.byref(params), nj, icp, ntr, cfp, C.byref(st)))
        return out, st

    def items(self, job, n_items, L, rows=False, zimmermann=False):
        """Per-item results of job `job` of the last cf_run; rows [n_items][L], or [n_items][3][L]
        dipole components for a Zimmermann run."""
        ln = np.zeros(n_items, np.int32)
        stt = np.zeros(n_items, np.int32)
        nr = np.zeros(n_items, np.int32)
        dr = np.zeros(n_items, np.float32)
        rw = (np.zeros((n_items, 3, L)) if zimmermann else np.zeros((n_items, L))) if rows else None
        p = lambda a: a.ctypes.data_as(C.c_void_p) if a is not None else None
        _check(lib().hh_cf_items(self.h, job, p(ln), p(stt), p(nr), p(dr), p(rw)))
        return dict(len=ln, status=stt, nrhs=nr, e_drift=dr, rows=rw)

    def vegas(self, system, integrand, T, Rmin, Rmax, pair_state=PAIR_FREE, niterations=12, npoints=1000000,
              seed=1):
        """VEGAS phase-space integral (hh_vegas): returns the VegasResult."""
        p = VegasParams(integrand, pair_state, float(T), float(Rmin), float(Rmax), int(niterations), int(npoints),
                        int(seed))
        r = VegasResult()
        _check(lib().hh_vegas(self.h, C.byref(system), C.byref(p), C.byref(r)))
        return r

    def cf_to_sf(self, cf_dev, dt):
        """Apodizes cf_dev in place (as the reference) and returns (nu, sf, ms) device tensors."""
        torch = self.torch
        n = cf_dev.numel()
        nu = torch.empty_like(cf_dev)
        sf = torch.empty_like(cf_dev)
        ms = C.c_double(0)
        torch.cuda.synchronize(self.device)
        _check(lib().hh_cf_to_sf(self.h, C.c_void_p(cf_dev.data_ptr()), n, dt, C.c_void_p(nu.data_ptr()),
                                 C.c_void_p(sf.data_ptr()), C.byref(ms)))
        return nu, sf, ms.value

    def eval(self, system, what, y_dev):
        """Device surfaces at phase points y_dev (SoA [QP][n]): EVAL_PES -> [n], EVAL_DIPOLE ->
        [3][n], EVAL_HAMILTONIAN -> [n], EVAL_RHS -> [QP][n], EVAL_DIPOLE_RATE (|d mu/dt|^2) -> [n]."""
        n = y_dev.shape[1]
        rows = {EVAL_PES: 1, EVAL_DIPOLE: 3, EVAL_HAMILTONIAN: 1, EVAL_RHS: y_dev.shape[0], EVAL_DIPOLE_RATE: 1}[what]
        out = self.torch.empty((rows, n), dtype=self.torch.float64, device=y_dev.device)
        self.torch.cuda.synchronize(self.device)
        _check(lib().hh_eval(self.h, C.byref(system), what, C.c_void_p(y_dev.data_ptr()), n,
                             C.c_void_p(out.data_ptr())))
        return out

    def prmu(self, system, params, ic_dev, sf_dev=None):
        """PR/mu spectrum (hh_prmu) of the R0-injected ensemble ic_dev (SoA [QP][ntraj]): adds
        every counted trajectory's weighted power spectrum into sf_dev ([nfreq], zeros when None).
        Returns (sf_dev, counted [ntraj] numpy, ncounted, stats)."""
        torch = self.torch
        if sf_dev is None:
            sf_dev = torch.zeros(params.nfreq, dtype=torch.float64, device=f"cuda:{self.device}")
        n = ic_dev.shape[1]
        counted = np.zeros(n, dtype=np.int32)
        nc = C.c_long(0)
        st = CFStats()
        torch.cuda.synchronize(self.device)
        _check(lib().hh_prmu(self.h, C.byref(system), C.byref(params), C.c_void_p(ic_dev.data_ptr()), n,
                             C.c_void_p(sf_dev.data_ptr()), counted.ctypes.data_as(C.c_void_p), C.byref(nc),
                             C.byref(st)))
        return sf_dev, counted, nc.value, st

    def cf_array(self, system, params, ic_dev, T_base, pair_state, pesmin, temps, ratios=None):
        """CORRELATION_ARRAY (hh_cf_array): the CF of the ensemble ic_dev (SoA [QP][ntraj], sampled
        at T_base) reweighted to each of `temps`.  Returns (CFs [ntemp][L] device tensor, nstar
        [ntemp] numpy, stats)."""
        torch = self.torch
        t = np.ascontiguousarray(temps, dtype=np.float64)
        r = None if ratios is None else np.ascontiguousarray(ratios, dtype=np.float64)
        ap = CFArrayParams(T_base, pair_state, len(t), pesmin, t.ctypes.data_as(C.c_void_p),
                           r.ctypes.data_as(C.c_void_p) if r is not None else None)
        out = torch.zeros((len(t), params.L), dtype=torch.float64, device=f"cuda:{self.device}")
        nstar = np.zeros(len(t))
        st = CFStats()
        torch.cuda.synchronize(self.device)
        _check(lib().hh_cf_array(self.h, C.byref(system), C.byref(params), C.byref(ap), C.c_void_p(ic_dev.data_ptr()),
                                 ic_dev.shape[1], C.c_void_p(out.data_ptr()), nstar.ctypes.data_as(C.c_void_p),
                                 C.byref(st)))
        return out, nstar, st

    def m0(self, system, sampler, seed, npoints, nstreams, ppf=1.0):
        m, q, ms = C.c_double(), C.c_double(), C.c_double()
        self.torch.cuda.synchronize(self.device)
        _check(lib().hh_m0(self.h, C.byref(system), C.byref(sampler), seed, npoints, nstreams, ppf, C.byref(m),
                           C.byref(q), C.byref(ms)))
        return m.value, q.value, ms.value

    def m2(self, system, sampler, seed, npoints, nstreams, ppf=1.0):
        """Rejection-sampler M2 (hh_m2): the hh_m0 streams averaging |d mu/dt|^2, times SecondCoeff."""
        m, q, ms = C.c_double(), C.c_double(), C.c_double()
        self.torch.cuda.synchronize(self.device)
        _check(lib().hh_m2(self.h, C.byref(system), C.byref(sampler), seed, npoints, nstreams, ppf, C.byref(m),
                           C.byref(q), C.byref(ms)))
        return m.value, q.value, ms.value


class Session:
    """Serving session on an Engine: continuous batching of CF and M0 requests (see
    hh_session_create).  Device buffers passed to submit_cf must stay alive until the request
    has finished; this wrapper keeps references to them until then."""

    def __init__(self, engine, slice_ms=10.0, max_device_gib=0.0, block=0):
        self.engine = engine
        opts = SessionOpts(slice_ms, max_device_gib, block)
        h = C.c_void_p()
        _check(lib().hh_session_create(engine.h, C.byref(opts), C.byref(h)))
        self.h = h
        self._keep = {}

    def submit_cf(self, system, params, ic_dev, out=None):
        """ic_dev: device tensor SoA [QP][ntraj]; returns (request id, CF output tensor)."""
        torch = self.engine.torch
        if out is None:
            out = torch.zeros(params.L, dtype=torch.float64, device=ic_dev.device)
        rid = C.c_uint64()
        torch.cuda.synchronize(ic_dev.device)
        _check(lib().hh_submit_cf(self.h, C.byref(system), C.byref(params), C.c_void_p(ic_dev.data_ptr()),
                                  ic_dev.shape[1], C.c_void_p(out.data_ptr()), C.byref(rid)))
        self._keep[rid.value] = (ic_dev, out)
        return rid.value, out

    def submit_m0(self, system, sampler, seed, npoints, nstreams, ppf=1.0):
        rid = C.c_uint64()
        _check(lib().hh_submit_m0(self.h, C.byref(system), C.byref(sampler), seed, npoints, nstreams, ppf,
                                  C.byref(rid)))
        return rid.value

    def poll(self, rid, timeout_ms=-1.0):
        """None while the request runs; its RequestInfo once finished."""
        info = RequestInfo()
        r = _check(lib().hh_poll(self.h, rid, timeout_ms, C.byref(info)))
        if r == 0:
            return None
        self._keep.pop(rid, None)
        return info

    def close(self):
        if self.h:
            _check(lib().hh_session_destroy(self.h))
            self.h = None
            self._keep.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
