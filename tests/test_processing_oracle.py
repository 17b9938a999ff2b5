"""The PROCESSING spectra against the reference's own expected outputs (tests/desymmetrizations.conf,
tests/compute-alpha.conf: SPCL, SPD1..SPD4a-F-He-Ar-1000.0.txt, every 4th point in
tests/golden/spectra_hear_1000.npz), restated here in numpy/scipy from the formulas
(compute_alpha src/hawaii.c:5904-5924, desymmetrize_* src/hawaii.c:5696-5881) on the reference's
SF and CF files, with the oracle's CF -> SF for D4/D4a.  This pins the restatement the driver
follows; the driver itself is checked against the same vectors on the GPU
(tests/test_gpu_driver.py).  CPU only."""
import numpy as np
import pytest
from scipy.interpolate import CubicSpline

import oracle_lib as O

PLANCK, BOLTZMANN, C_CM, HBAR, ATU = 6.62607015e-34, 1.380649e-23, 2.99792458e10, 1.054571817e-34, 2.4188843265857e-17
MOMENT_SF = 8.0 * np.pi ** 3 * 2.686780111e25 ** 2 / 3.0 / HBAR


def hnukt(nu, T):
    return PLANCK * C_CM * nu / (BOLTZMANN * T)


def alpha(nu, sf, T):
    return MOMENT_SF * nu * (1.0 - np.exp(-hnukt(nu, T))) * sf


def close_to_reference(got, want):
    """the reference files print 11 significant digits: 1e-9 of the peak everywhere, 1e-6
    relative where the value is above 1e-6 of the peak"""
    peak = np.abs(want).max()
    assert np.max(np.abs(got - want)) < 1e-9 * peak
    big = np.abs(want) > 1e-6 * peak
    assert np.max(np.abs(got[big] / want[big] - 1.0)) < 1e-6


@pytest.fixture(scope="module")
def ref(golden):
    return golden("cf_sf_hear_1000.npz"), golden("spectra_hear_1000.npz")


def test_alpha_and_d1_d2_d3(ref):
    g, s = ref
    T = float(g["T"][0])
    nu, sf = g["sf_nu"][::4], g["sf"][::4]
    assert np.allclose(nu, s["nu"], atol=1e-9)
    h = hnukt(nu, T)
    close_to_reference(alpha(nu, sf, T), s["SPCL"])
    close_to_reference(alpha(nu, sf * 2.0 / (1.0 + np.exp(-h)), T), s["SPD1"])
    with np.errstate(invalid="ignore", divide="ignore"):
        d2 = np.where(nu < 1e-10, 0.0, sf * h / (1.0 - np.exp(-h)))
    close_to_reference(alpha(nu, d2, T), s["SPD2"])
    close_to_reference(alpha(nu, sf * np.exp(h / 2.0), T), s["SPD3"])


@pytest.mark.parametrize("which", ["SPD4", "SPD4a"])
def test_d4_d4a_egelstaff(ref, which):
    g, s = ref
    T = float(g["T"][0])
    t, cf = g["cf_t"], g["cf"]
    cc = HBAR / BOLTZMANN / T / ATU / 2.0
    spl = CubicSpline(t, cf, bc_type="natural")
    norm = spl(0.0) / spl(cc) if which == "SPD4a" else 1.0
    te = np.sqrt(t * t + cc * cc)
    keep = te <= t[-1]
    n2 = 1 << int(np.ceil(np.log2(keep.sum())))
    tt, yy = np.zeros(n2), np.zeros(n2)
    tt[:keep.sum()] = t[keep]
    yy[:keep.sum()] = norm * spl(te[keep])
    nu, sf = O.cf_to_sf(tt, yy)
    sfd = sf * np.exp(hnukt(nu, T) / 2.0)
    close_to_reference(alpha(nu, sfd, T)[::4], s[which])
