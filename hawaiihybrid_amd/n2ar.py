"""N2-Ar PIP-NN surfaces: weights, network descriptors and the device evaluator.

The weights are the reference's trained models (PES-IDS/npz/n2-ar-nonrigid-17-32-1-f12.npz,
n2-ar-dipoleq-effquad-{short,long}.npz), shipped as data under data/n2ar/ and read with
numpy's pickle-free loader.  Each network layer is stored as [in][out] (x W + b), the layout
the reference's Eigen maps produce (PES-IDS/mlp.hpp:300-309).
"""
import ctypes as C
import os

import numpy as np

from . import HHError, _check, lib

DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "n2ar")
PES_FILE = "n2-ar-nonrigid-17-32-1-f12.npz"
DIP_FILES = ("n2-ar-dipoleq-effquad-short.npz", "n2-ar-dipoleq-effquad-long.npz")
# PES value at infinite separation cached by the reference for this model
# (PES-IDS/ai_pes_n2_ar_pip_nn.cpp:131)
PES_EN_INF = 0.0384490895371528

ACT_SILU, ACT_TANH = 0, 1
MLP_BF16, MLP_BF16X3 = 0, 1


class MlpDesc(C.Structure):
    _fields_ = [("n_layers", C.c_int), ("dims", C.c_int * 5), ("activation", C.c_int),
                ("W", C.c_void_p * 4), ("b", C.c_void_p * 4)]


class PipnnDesc(C.Structure):
    _fields_ = [("pes", MlpDesc), ("pes_xmean", C.c_void_p), ("pes_xscale", C.c_void_p),
                ("pes_ymean", C.c_double), ("pes_yscale", C.c_double), ("pes_en_inf", C.c_double),
                ("dip", (MlpDesc * 3) * 2), ("dip_xmean", C.c_void_p * 2), ("dip_xscale", C.c_void_p * 2)]


def _load(name):
    with np.load(os.path.join(DATA, name), allow_pickle=False) as z:
        return {k: np.ascontiguousarray(z[k], dtype=np.float64) if z[k].dtype != np.int64 else z[k].copy()
                for k in z.files}


def load_weights():
    """Dict with 'pes' = (layers, xmean, xscale, ymean, yscale) and 'dip' = [short, long], each
    (subnets, xmean, xscale) with subnets = [layers of subnet 0, 1, 2]; layers = [(W, b), ...]."""
    z = _load(PES_FILE)
    arch = [int(a) for a in z["architecture"]]
    layers = [(z[f"{2 * k}.weight"], z[f"{2 * k}.bias"]) for k in range(len(arch) - 1)]
    pes = (layers, z["xscaler.mean"], z["xscaler.scale"], float(z["yscaler.mean"][0]), float(z["yscaler.scale"][0]))
    dips = []
    for f in DIP_FILES:
        z = _load(f)
        nsub = sum(1 for k in z if k.startswith("architecture"))
        subnets = []
        for s in range(nsub):
            a = [int(v) for v in z[f"architecture.{s}"]]
            subnets.append([(z[f"blocks.{s}.{2 * k}.weight"], z[f"blocks.{s}.{2 * k}.bias"]) for k in range(len(a) - 1)])
        dips.append((subnets, z["xscaler.mean"], z["xscaler.scale"]))
    return {"pes": pes, "dip": dips}


def synthetic_pes(width, n_hidden=3, seed=2024):
    """The benchmark's synthetic PES shape: 17 polynomials -> n_hidden x width (tanh) -> 1, weights and
    biases ~ U(-1/sqrt(fan_in), 1/sqrt(fan_in)) from a fixed seed (PyTorch's nn.Linear default)."""
    rng = np.random.default_rng(seed)
    dims = [17] + [width] * n_hidden + [1]
    layers = []
    for k in range(len(dims) - 1):
        bound = 1.0 / np.sqrt(dims[k])
        layers.append((rng.uniform(-bound, bound, (dims[k], dims[k + 1])), rng.uniform(-bound, bound, dims[k + 1])))
    return layers


def _mlp_desc(layers, act, keep):
    d = MlpDesc()
    d.n_layers = len(layers)
    d.dims[0] = layers[0][0].shape[0]
    for k, (W, b) in enumerate(layers):
        W = np.ascontiguousarray(W, dtype=np.float64)
        b = np.ascontiguousarray(b, dtype=np.float64)
        keep += [W, b]
        d.dims[k + 1] = W.shape[1]
        d.W[k] = W.ctypes.data
        d.b[k] = b.ctypes.data
    d.activation = act
    return d


def make_desc(weights, pes_layers=None, pes_act=ACT_SILU, with_dipole=True):
    """C descriptor (and the arrays it points into) from load_weights(); pes_layers replaces the
    PES network (the synthetic shapes keep the reference's input scaler, ymean = 0, yscale = 1)."""
    keep = []
    d = PipnnDesc()
    layers, xm, xs, ym, ys = weights["pes"]
    en_inf = PES_EN_INF
    if pes_layers is not None:
        layers, ym, ys, en_inf = pes_layers, 0.0, 1.0, 0.0
    d.pes = _mlp_desc(layers, pes_act, keep)
    xm = np.ascontiguousarray(xm)
    xs = np.ascontiguousarray(xs)
    keep += [xm, xs]
    d.pes_xmean, d.pes_xscale = xm.ctypes.data, xs.ctypes.data
    d.pes_ymean, d.pes_yscale, d.pes_en_inf = ym, ys, en_inf
    if with_dipole:
        for m, (subnets, dxm, dxs) in enumerate(weights["dip"]):
            for s, sub in enumerate(subnets):
                d.dip[m][s] = _mlp_desc(sub, ACT_SILU, keep)
            dxm = np.ascontiguousarray(dxm)
            dxs = np.ascontiguousarray(dxs)
            keep += [dxm, dxs]
            d.dip_xmean[m], d.dip_xscale[m] = dxm.ctypes.data, dxs.ctypes.data
    return d, keep


class PipNN:
    """The N2-Ar surfaces on the device (hh_pipnn_*): energies (Hartree) and lab-frame dipoles
    of batches of phase points given as lab coordinates (Phi, Theta, R, phi1, theta1).

    precision MLP_BF16X3 (the default) splits every operand into a BF16 head and tail and takes
    three matrix-core products: within 0.25 cm^-1 of the FP64 networks, the mode for dynamics.
    MLP_BF16 takes one product: 1.6x the throughput, median error 2e-3 cm^-1 but outliers of
    ~100 cm^-1 on the repulsive wall and ~10% of the largest dipole, so it is a throughput mode,
    not a parity-grade surface."""

    def __init__(self, engine, precision=MLP_BF16X3, pes_layers=None, pes_act=ACT_SILU, with_dipole=True, weights=None):
        self.engine = engine
        self.weights = weights if weights is not None else load_weights()
        desc, keep = make_desc(self.weights, pes_layers, pes_act, with_dipole)
        h = C.c_void_p()
        _check(lib().hh_pipnn_create(engine.h, C.byref(desc), precision, C.byref(h)))
        self.h = h
        self.with_dipole = with_dipole

    def close(self):
        if self.h:
            lib().hh_pipnn_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def eval(self, q_soa, energy=True, dipole=True):
        """q_soa: float64 cuda tensor [5][n].  Returns (E[n] or None, mu[3][n] or None, device ms)."""
        import torch
        if q_soa.dtype != torch.float64 or not q_soa.is_cuda or q_soa.dim() != 2 or q_soa.shape[0] != 5:
            raise HHError("q must be a float64 cuda tensor of shape [5][n]")
        q_soa = q_soa.contiguous()
        n = q_soa.shape[1]
        E = torch.empty(n, dtype=torch.float64, device=q_soa.device) if energy else None
        mu = torch.empty((3, n), dtype=torch.float64, device=q_soa.device) if dipole else None
        ms = C.c_double()
        torch.cuda.synchronize(q_soa.device)
        _check(lib().hh_pipnn_eval(self.h, C.c_void_p(q_soa.data_ptr()), n,
                                   C.c_void_p(E.data_ptr()) if E is not None else None,
                                   C.c_void_p(mu.data_ptr()) if mu is not None else None, C.byref(ms)))
        return E, mu, ms.value


def phase_points(n, T, seed=42, Rmin=4.51, Rmax=40.0):
    """n phase points of N2-Ar from the seeded stream by q_generator + p_generator at T (host,
    bit-exact reference draws); returns the state rows [n][10]."""
    from . import IC_PHASE_SPACE, PAIR_ALL, Sampler, n2_ar, sample_ics
    smp = Sampler(IC_PHASE_SPACE, PAIR_ALL, float(T), Rmin, Rmax, 0.0, 0.0)
    ic, _, _ = sample_ics(n2_ar(), smp, seed, n)
    return ic


def lab_q(states):
    """State rows [n][10] -> lab coordinates [5][n] (Phi, Theta, R, phi1, theta1)."""
    return np.ascontiguousarray(states[:, 0::2].T)
