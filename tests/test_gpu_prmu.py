"""PR/mu spectral function on the device (hh_prmu) against the oracle: the same R0-injected
initial conditions, the oracle's series (orc_prmu_series) and spectra (orc_prmu_power, a real
FFT per component) weighted by SF_COEFF * pr_mu (src/hawaii.c:3395, 3509, 3820).

Tolerances: which trajectories count (ended beyond R0 within L samples) exactly; the summed
spectrum within 1e-8 relative L2 in FP64 (the integrators differ in rounding only; the device
FFT pairs two components in one complex transform), 1e-4 with FP32 surfaces."""
import numpy as np
import pytest

import hawaiihybrid_amd as H
import oracle_lib as O

pytestmark = pytest.mark.gpu

HARTREE, ALU, ATU = 4.3597447222071e-18, 5.29177210903e-11, 2.4188843265857e-17


@pytest.fixture(scope="module")
def eng():
    e = H.Engine(0)
    yield e
    e.close()


def oracle_sf(so, ic, mu, dt, L, R0, nfreq):
    coeff = HARTREE * ALU ** 6 * ATU * R0 * R0 * dt * dt
    sf = np.zeros(nfreq)
    counted = np.zeros(len(ic), dtype=np.int32)
    for t, y0 in enumerate(ic):
        status, n, vec = O.prmu_series(so, y0, dt, L + 2, R0)
        if status != 0 or n > L + 1:
            continue
        counted[t] = 1
        x = np.zeros((3, L))
        x[:, :n - 1] = vec[1:n].T
        sf += coeff * (-y0[5] / mu) * O.prmu_power(x, dt, nfreq)
    return sf, counted


@pytest.mark.parametrize("which,prec", [("hear", H.PREC_FP64), ("co2ar", H.PREC_FP64), ("co2ar", H.PREC_MIXED)])
def test_prmu_matches_oracle(eng, which, prec):
    if which == "hear":
        sh, so, mu, pesmin, n = H.he_ar(), O.hear(), O.MU_HEAR, H.PESMIN_HEAR, 256
    else:
        sh, so, mu, pesmin, n = H.co2_ar(), O.co2ar(), O.MU_CO2AR, H.PESMIN_CO2AR, 48
    dt, L, R0, nfreq = 200.0, 512, 12.0, 200
    smp = H.Sampler(H.IC_R0_INJECT, H.PAIR_FREE, 300.0, 4.0, 40.0, pesmin, R0)
    ic, _, _ = H.sample_ics(sh, smp, 31, n)
    print(f"prmu {which} prec {prec}: {n} trajectories, L {L}", flush=True)
    p = H.PrmuParams(1e-12, 1e-12, dt, R0, L, nfreq, prec, 0, 0)
    sf, counted, nc, st = eng.prmu(sh, p, eng.to_device_soa(ic))
    want, wc = oracle_sf(so, ic, mu, dt, L, R0, nfreq)
    assert np.array_equal(counted, wc)
    assert nc == wc.sum() and 0 < nc <= n
    got = sf.cpu().numpy()
    err = np.linalg.norm(got - want) / np.linalg.norm(want)
    print(f"  counted {nc}/{n}; rel L2 {err:.2e}", flush=True)
    assert err < (1e-8 if prec == H.PREC_FP64 else 1e-4)


def test_prmu_accumulates_and_truncation_flag(eng):
    """sf_dev accumulates over calls; allow_truncating counts every finished trajectory."""
    sh = H.he_ar()
    smp = H.Sampler(H.IC_R0_INJECT, H.PAIR_FREE, 300.0, 4.0, 40.0, H.PESMIN_HEAR, 12.0)
    ic, _, _ = H.sample_ics(sh, smp, 8, 64)
    icd = eng.to_device_soa(ic)
    p = H.PrmuParams(1e-12, 1e-12, 200.0, 12.0, 256, 100, H.PREC_FP64, 0, 0)
    a, ca, na, _ = eng.prmu(sh, p, icd)
    a = a.cpu().numpy()
    b, _, _, _ = eng.prmu(sh, p, icd, eng.torch.from_numpy(a.copy()).cuda())
    np.testing.assert_allclose(b.cpu().numpy(), 2 * a, rtol=1e-15)
    p.allow_truncating = 1
    _, ct, nt, _ = eng.prmu(sh, p, icd)
    assert nt == 64 and np.all(ct == 1) and na < 64


def test_prmu_rejects_bad_input(eng):
    sh = H.he_ar()
    icd = eng.to_device_soa(np.zeros((4, 6)) + 12.0)
    for bad in (H.PrmuParams(1e-12, 1e-12, 200.0, 12.0, 500, 100, 0, 0, 0),    # not a power of two
                H.PrmuParams(1e-12, 1e-12, 200.0, 12.0, 256, 200, 0, 0, 0),    # nfreq > L/2 + 1
                H.PrmuParams(1e-12, 1e-12, 200.0, -1.0, 256, 100, 0, 0, 0)):   # R0 <= 0
        with pytest.raises(H.HHError):
            eng.prmu(sh, bad, icd)


@pytest.mark.parametrize("L", [4, 8192, 65536, 1 << 20])
def test_prmu_fft_lengths(eng, L):
    """The four-step transform's factorisations at the edges (L = 4: 2 x 2; 65536: 256 x 256;
    2^20: 1024 x 1024 with four column tiles) against the oracle's direct radix-2 FFT of the same
    series, over the whole spectrum (nfreq = L/2 + 1); one trajectory at 2^20."""
    sh, so = H.he_ar(), O.hear()
    dt, R0 = 200.0, 12.0
    smp = H.Sampler(H.IC_R0_INJECT, H.PAIR_FREE, 300.0, 4.0, 40.0, H.PESMIN_HEAR, R0)
    ic, _, _ = H.sample_ics(sh, smp, 13, 1 if L > 65536 else 8)
    nfreq = L // 2 + 1
    p = H.PrmuParams(1e-12, 1e-12, dt, R0, L, nfreq, H.PREC_FP64, 0, 1)   # keep truncated ones (L = 4)
    sf, counted, nc, _ = eng.prmu(sh, p, eng.to_device_soa(ic))
    coeff = HARTREE * ALU ** 6 * ATU * R0 * R0 * dt * dt
    want = np.zeros(nfreq)
    for t, y0 in enumerate(ic):
        status, n, vec = O.prmu_series(so, y0, dt, L + 2, R0)
        assert status == 0 and counted[t] == 1
        x = np.zeros((3, L))
        m = min(n - 1, L)
        x[:, :m] = vec[1:m + 1].T
        want += coeff * (-y0[5] / O.MU_HEAR) * O.prmu_power(x, dt, nfreq)
    got = sf.cpu().numpy()
    assert np.linalg.norm(got - want) / np.linalg.norm(want) < 1e-9
