"""CPU checks of the engine library: it loads, exports every entry point of include/*.h, and
its host-side initial-condition sampler reproduces the reference stream bit-exactly.  No
device calls."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import hawaiihybrid_amd as H
import oracle_lib as O

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for fn in os.listdir(os.path.join(REPO, "include")):
        # hawaii_ref.h is libhawaii.so's (tests/test_ref_api.py)
        if not fn.endswith(".h") or fn == "hawaii_ref.h":
            continue
        text = open(os.path.join(REPO, "include", fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w\s\*]*?\b(hh_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return sorted(names)


def test_library_exports_every_declared_entry_point():
    lib = H.lib()
    names = declared_functions()
    assert len(names) >= 10
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_mt_stream_matches_reference(golden):
    g = golden("mt_streams.npz")
    for seed in (42, 43, 20106943, 0, 1):
        ref = g[f"seed_{seed}"]
        assert np.array_equal(H.mt_stream(seed, ref.size), ref)


@pytest.mark.parametrize("which", ["hear", "co2ar"])
def test_r0_injection_ics_bit_exact(which):
    if which == "hear":
        sh, so, T = H.he_ar(), O.hear(), 300.0
        pes = O.PESMIN_HEAR
    else:
        sh, so, T = H.co2_ar(), O.co2ar(), 200.0
        pes = O.PESMIN_CO2AR
    ic, n, nd = H.sample_ics(sh, H.Sampler(H.IC_R0_INJECT, 1, T, 4.0, 40.0, pes, 40.0), 42, 4000)
    ref, n2, nd2 = O.sample_ics(so, O.Sampler(O.IC_R0_INJECT, 1, T, 4.0, 40.0, pes, 40.0), 42, 4000)
    assert np.array_equal(ic, ref) and n == n2 and nd == nd2


def test_rejection_ics_bit_exact_with_reference_plugin():
    """Rejection sampling with the reference's own compiled He-Ar plugin (PES-IDS/HeAr.h built
    into oracle/_ref) as pes_lab: identical phase points and candidate counts."""
    ref_hear = O.ref_lib("libref_hear.so")
    if ref_hear is None:
        pytest.skip("oracle/_ref not built (needs /root/reference)")
    pes = C.cast(ref_hear.pes_lab, C.c_void_p).value
    smp = H.Sampler(H.IC_REJECTION, H.PAIR_FREE, 300.0, 4.0, 40.0, H.PESMIN_HEAR, 0.0)
    ic, n, nd = H.sample_ics(H.he_ar(), smp, 42, 3000, pes=pes)
    ref, n2, nd2 = O.sample_ics(O.hear(), O.Sampler(0, 1, 300.0, 4.0, 40.0, O.PESMIN_HEAR, 0.0), 42, 3000)
    assert np.array_equal(ic, ref) and n == n2 and nd == nd2


@pytest.mark.parametrize("pair", [H.PAIR_FREE, H.PAIR_BOUND])
def test_rejection_ics_bit_exact_builtin_surfaces(pair):
    smp = H.Sampler(H.IC_REJECTION, pair, 300.0, 4.0, 40.0, H.PESMIN_CO2AR, 0.0)
    ic, n, nd = H.sample_ics(H.co2_ar(), smp, 7, 1500)
    ref, n2, nd2 = O.sample_ics(O.co2ar(), O.Sampler(0, pair, 300.0, 4.0, 40.0, O.PESMIN_CO2AR, 0.0), 7, 1500)
    assert np.array_equal(ic, ref) and n == n2


def test_sampler_rejects_invalid_arguments():
    bad = H.System(H.SYS_HEAR, -1.0, 0.0)
    with pytest.raises(H.HHError, match="reduced mass"):
        H.sample_ics(bad, H.Sampler(H.IC_R0_INJECT, 1, 300.0, 4.0, 40.0, -1e-4, 40.0), 1, 4)
    with pytest.raises(H.HHError, match="Rmin"):
        H.sample_ics(H.he_ar(), H.Sampler(H.IC_REJECTION, 1, 300.0, 40.0, 4.0, -1e-4, 0.0), 1, 4)
