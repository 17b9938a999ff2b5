#!/usr/bin/env python3
"""Write the golden fixtures under tests/golden/ (run here, where /root/reference exists).

  mt_streams.npz       u32 words from the reference's own thirdparty/mtwist.c
                       (oracle/_ref/libref_mtwist.so): mt_seed32(seed); mt_lrand() ...
  hear_ref.npz         V_HeAr, dV_HeAr, dip_HeAr from the reference's PES-IDS/HeAr.h
                       (oracle/_ref/libref_hear.so) on an R grid
  co2ar_surface_ref.npz  the reference CO2-Ar expressions (ai_pes_co2ar.c pes_co2ar,
                       dpes_dR, dpes_dTheta; ai_ids_co2ar.cpp dipx, dipz) read as arithmetic
                       and evaluated in 40-digit arithmetic at (R, theta) points, with the
                       Legendre functions they read (P_l, and P_l^1 with the Condon-Shortley
                       phase) evaluated from their definitions
  cf_sf_hear_1000.npz  the data columns of the reference's own test fixtures
                       tests/CF-F-He-Ar-1000.0.txt and tests/SF-F-He-Ar-1000.0.txt
                       (the CF_TO_SF known-answer pair of tests/convert-cf-to-sf.conf)
  spectra_hear_1000.npz  every 4th point of the expected spectra of the reference's
                       tests/desymmetrizations.conf and tests/compute-alpha.conf
                       (SPCL, SPD1, SPD2, SPD3, SPD4, SPD4a-F-He-Ar-1000.0.txt: ALPHA of the
                       CF_TO_SF of CF-F-He-Ar-1000.0 without and with each desymmetrization)
  baseline_smooth_hear_1000.npz  the expected outputs of the reference's FIT_BASELINE and SMOOTH
                       tests (tests/convert-cf-to-sf.conf: SF-F-He-Ar-1000.0-with-baseline.txt,
                       ...-with-custom-baseline.txt, the CF_TO_SF of CF-F-He-Ar-1000.0 after
                       FIT_BASELINE() and FIT_BASELINE(8000); tests/smooth.conf:
                       SPCL-F-He-Ar-1000.0-sm.txt, ALPHA of SMOOTH(grid_npoints=1200,
                       grid_max=1200.0, window_size_delay=10, window_size_step=0.05) of
                       SF-F-He-Ar-1000.0)
  add_spectra_hear_100.npz  the data columns of the reference's tests/add-spectra.conf inputs
                       (SPCL-B- and SPCL-F-He-Ar-100.0-raw.txt) and expected sums
                       (SPCL-T-He-Ar-100.0-raw.txt, ...-raw-wn-step=0.05.txt)

Build the reference checkers first: make -C oracle ref
"""
import ctypes as C
import os
import sys

import mpmath as mp
import numpy as np
import sympy as sp

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.abspath(os.path.join(HERE, "..", ".."))
sys.path.insert(0, os.path.join(REPO, "hawaiihybrid_amd", "data"))
import co2ar_parse as cp  # noqa: E402

REF = "/root/reference"
mp.mp.dps = 40


def mt_streams():
    lib = C.CDLL(os.path.join(REPO, "oracle", "_ref", "libref_mtwist.so"))
    lib.mt_seed32.argtypes = [C.c_uint32]
    lib.mt_lrand.restype = C.c_uint32
    out = {}
    for seed, n in ((42, 20000), (43, 2000), (20106943, 2000), (0, 2000), (1, 2000)):
        lib.mt_seed32(seed)
        out[f"seed_{seed}"] = np.array([lib.mt_lrand() for _ in range(n)], dtype=np.uint32)
    # a long prefix as a checksum: 10^6 words of seed 42, xor-folded and summed
    lib.mt_seed32(42)
    words = np.array([lib.mt_lrand() for _ in range(1_000_000)], dtype=np.uint64)
    out["seed_42_1e6_sum"] = np.array([int(words.sum())], dtype=np.uint64)
    out["seed_42_1e6_last"] = words[-16:].astype(np.uint32)
    np.savez_compressed(os.path.join(HERE, "mt_streams.npz"), **out)


def hear():
    lib = C.CDLL(os.path.join(REPO, "oracle", "_ref", "libref_hear.so"))
    for f in ("V_HeAr", "dV_HeAr", "dip_HeAr"):
        getattr(lib, f).restype = C.c_double
        getattr(lib, f).argtypes = [C.c_double]
    R = np.linspace(3.0, 60.0, 571)
    np.savez_compressed(os.path.join(HERE, "hear_ref.npz"), R=R,
                        V=np.array([lib.V_HeAr(x) for x in R]),
                        dV=np.array([lib.dV_HeAr(x) for x in R]),
                        dip=np.array([lib.dip_HeAr(x) for x in R]))


def legendre_values(x, lmax=11):
    """P_l(x) and P_l^1(x) = -(1-x^2)^(1/2) dP_l/dx, l = 0..lmax, in mp arithmetic."""
    X = sp.Symbol("X")
    P, P1 = [], []
    for l in range(lmax + 1):
        pl = sp.legendre(l, X)
        P.append(mp.mpf(sp.N(pl.subs(X, sp.Float(str(x), 45)), 45)))
        dpl = sp.diff(pl, X)
        P1.append(-mp.sqrt(1 - mp.mpf(x) ** 2) * mp.mpf(sp.N(dpl.subs(X, sp.Float(str(x), 45)), 45)))
    return P, P1


def co2ar():
    pes_src = open(os.path.join(REF, "PES-IDS", "ai_pes_co2ar.c")).read()
    ids_src = open(os.path.join(REF, "PES-IDS", "ai_ids_co2ar.cpp")).read()
    HTOCM = mp.mpf("2.1947463136320e5")

    def branch_list(fn, tail=None):
        body = cp._function_body(pes_src if fn.startswith(("pes", "dpes")) else ids_src,
                                 (r"double\s+" + fn + r"\s*\(\s*double\s+R\s*,\s*double\s+Theta\s*\)\s*\{")
                                 if fn.startswith(("pes", "dpes")) else
                                 (r"double\s+AI_IDS_co2_ar::" + fn + r"\s*\([^)]*\)\s*\{"))
        br, _ = cp.split_branches(body)
        return [(u, cp.eval_statements(t, tail_return=tail)) for u, t in br]

    fns = {"V": branch_list("pes_co2ar", "t48"), "dR": branch_list("dpes_dR", "t29"),
           "dT": branch_list("dpes_dTheta", "t86"), "dipx": branch_list("dipx"),
           "dipz": branch_list("dipz")}
    bounds = [u for u, _ in fns["V"][:-1]]
    Rs = [3.9, 4.3]
    for a, b in zip(bounds[:-1], bounds[1:]):
        Rs += [a + 0.3 * (b - a), a + 0.8 * (b - a)]
    Rs += [31.0, 38.5, 45.0, 60.0]
    thetas = [0.35, 1.0, np.pi / 2, 2.2, 2.9]
    pts, vals = [], {k: [] for k in fns}
    for R in Rs:
        b = next((i for i, u in enumerate(bounds) if R < u), len(bounds))
        for th in thetas:
            x = mp.cos(mp.mpf(th))
            P, P1 = legendre_values(float(mp.cos(mp.mpf(th))))
            subs = {cp.R: sp.Float(repr(R), 45), cp.COST: sp.Float(str(x), 45),
                    cp.SINT: sp.Float(str(mp.sin(mp.mpf(th))), 45)}
            for l in range(12):
                subs[cp.psym(l, 0)] = sp.Float(str(P[l]), 45)
                subs[cp.psym(l, 1)] = sp.Float(str(P1[l]), 45)
            for k, br in fns.items():
                e = br[b][1]
                v = mp.mpf(sp.N(e.subs(subs), 40))
                if k in ("V", "dR", "dT"):
                    v = v / HTOCM
                vals[k].append(float(v))
            pts.append((R, th))
    pts = np.array(pts)
    np.savez_compressed(os.path.join(HERE, "co2ar_surface_ref.npz"), R=pts[:, 0], theta=pts[:, 1],
                        **{k: np.array(v) for k, v in vals.items()})


def read_two_columns(path):
    rows = [ln.split() for ln in open(path) if ln.strip() and not ln.startswith("#")]
    return np.array([[float(a), float(b)] for a, b in rows])


def cf_sf():
    cf = read_two_columns(os.path.join(REF, "tests", "CF-F-He-Ar-1000.0.txt"))
    sf = read_two_columns(os.path.join(REF, "tests", "SF-F-He-Ar-1000.0.txt"))
    np.savez_compressed(os.path.join(HERE, "cf_sf_hear_1000.npz"), cf_t=cf[:, 0], cf=cf[:, 1],
                        sf_nu=sf[:, 0], sf=sf[:, 1], T=np.array([1000.0]), ntraj=np.array([5e8]))


def spectra():
    out = {}
    for name in ("SPCL", "SPD1", "SPD2", "SPD3", "SPD4", "SPD4a"):
        sp = read_two_columns(os.path.join(REF, "tests", name + "-F-He-Ar-1000.0.txt"))
        out["nu"] = sp[::4, 0]
        out[name] = sp[::4, 1]
    np.savez_compressed(os.path.join(HERE, "spectra_hear_1000.npz"), step=np.array([4]), **out)


def add_spectra():
    out = {}
    for key, name in (("B", "SPCL-B-He-Ar-100.0-raw"), ("F", "SPCL-F-He-Ar-100.0-raw"),
                      ("T", "SPCL-T-He-Ar-100.0-raw"), ("T005", "SPCL-T-He-Ar-100.0-raw-wn-step=0.05")):
        sp = read_two_columns(os.path.join(REF, "tests", name + ".txt"))
        out[key + "_nu"], out[key] = sp[:, 0], sp[:, 1]
    np.savez_compressed(os.path.join(HERE, "add_spectra_hear_100.npz"), **out)


def baseline_smooth():
    b = read_two_columns(os.path.join(REF, "tests", "SF-F-He-Ar-1000.0-with-baseline.txt"))
    c = read_two_columns(os.path.join(REF, "tests", "SF-F-He-Ar-1000.0-with-custom-baseline.txt"))
    sm = read_two_columns(os.path.join(REF, "tests", "SPCL-F-He-Ar-1000.0-sm.txt"))
    np.savez_compressed(os.path.join(HERE, "baseline_smooth_hear_1000.npz"), sf_baseline=b[:, 1],
                        sf_custom_baseline=c[:, 1], sm_nu=sm[:, 0], sm_alpha=sm[:, 1])


if __name__ == "__main__":
    # all fixtures, or the ones named on the command line
    todo = sys.argv[1:] or ["mt_streams", "hear", "co2ar", "cf_sf", "spectra", "add_spectra", "baseline_smooth"]
    for name in todo:
        globals()[name]()
    print("golden fixtures written to", HERE)
