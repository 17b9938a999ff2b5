"""Serving session on the GPU: continuous batching gives every request the same result as the
same request alone (CF bit for bit, by construction: per-item arithmetic does not depend on
the lane or slice an item runs in, and each job reduces in a fixed order), M0 requests equal
hh_m0, and admission follows the device-memory budget."""
import numpy as np
import pytest

import hawaiihybrid_amd as H

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eng():
    e = H.Engine(0)
    yield e
    e.close()


def co2_ics(eng, seed, n, T=200.0):
    smp = H.Sampler(H.IC_R0_INJECT, 1, T, 4.0, 40.0, H.PESMIN_CO2AR, 40.0)
    return eng.to_device_soa(H.sample_ics(H.co2_ar(), smp, seed, n)[0])


def hear_ics(eng, seed, n, T=300.0):
    smp = H.Sampler(H.IC_REJECTION, 1, T, 4.0, 40.0, H.PESMIN_HEAR, 0.0)
    return eng.to_device_soa(H.sample_ics(H.he_ar(), smp, seed, n)[0])


def test_session_cf_equals_cf_run(eng):
    ic = co2_ics(eng, 11, 600)
    p = H.cf_params(512, rtol=1e-10, precision=H.PREC_MIXED)
    ref, _ = eng.cf_run(H.co2_ar(), p, [ic])
    s = H.Session(eng, slice_ms=5.0)
    rid, out = s.submit_cf(H.co2_ar(), p, ic)
    info = s.poll(rid, 120000.0)
    s.close()
    assert info is not None and info.state == H.REQ_DONE, info.error if info else "timed out"
    assert np.array_equal(out.cpu().numpy(), ref[0].cpu().numpy())
    assert info.cf.n_items == 600 and info.latency_ms > 0


def test_session_mix_equals_alone(eng):
    """Four CF requests of two systems, two precisions and three temperatures plus one M0
    request, submitted together: each CF result has the bits of the same request run alone
    through hh_cf_run, and the M0 result equals hh_m0 on the same streams."""
    reqs = [
        (H.co2_ar(), H.cf_params(384, rtol=1e-10, precision=H.PREC_MIXED), co2_ics(eng, 1, 700, 150.0)),
        (H.co2_ar(), H.cf_params(384, rtol=1e-12, precision=H.PREC_FP64), co2_ics(eng, 2, 300, 300.0)),
        (H.he_ar(), H.cf_params(256, rtol=1e-12, halves=H.HALVES_BOTH), hear_ics(eng, 3, 500)),
        (H.co2_ar(), H.cf_params(384, rtol=1e-10, precision=H.PREC_MIXED), co2_ics(eng, 4, 200, 450.0)),
    ]
    alone = [eng.cf_run(sy, p, [ic])[0][0].cpu().numpy() for sy, p, ic in reqs]
    smp = H.Sampler(H.IC_REJECTION, 1, 300.0, 4.0, 40.0, H.PESMIN_HEAR, 0.0)
    m_ref, q_ref, _ = eng.m0(H.he_ar(), smp, 9, 64 * 400, 64)

    s = H.Session(eng, slice_ms=3.0)
    ids = [s.submit_cf(sy, p, ic) for sy, p, ic in reqs]
    mid = s.submit_m0(H.he_ar(), smp, 9, 64 * 400, 64)
    infos = [s.poll(rid, 300000.0) for rid, _ in ids]
    minfo = s.poll(mid, 300000.0)
    s.close()
    for (rid, out), info, ref in zip(ids, infos, alone):
        assert info is not None and info.state == H.REQ_DONE, info.error if info else "timed out"
        assert np.array_equal(out.cpu().numpy(), ref)
    assert minfo.state == H.REQ_DONE and minfo.kind == H.REQ_M0
    assert minfo.m0 == m_ref and minfo.m0_err == q_ref


def test_session_admission_by_device_memory(eng):
    """With a budget that holds one request's buffers at a time, the second request is
    admitted only after the first has finished; a request larger than the budget fails."""
    p = H.cf_params(4096, rtol=1e-10, precision=H.PREC_MIXED)
    one = 300 * (4096 * 8 + 16)
    s = H.Session(eng, slice_ms=5.0, max_device_gib=1.5 * one / 2**30)
    a = s.submit_cf(H.co2_ar(), p, co2_ics(eng, 21, 300))
    b = s.submit_cf(H.co2_ar(), p, co2_ics(eng, 22, 300))
    big = s.submit_cf(H.co2_ar(), p, co2_ics(eng, 23, 1000))
    ia, ib, ibig = s.poll(a[0], 300000.0), s.poll(b[0], 300000.0), s.poll(big[0], 300000.0)
    s.close()
    assert ia.state == H.REQ_DONE and ib.state == H.REQ_DONE
    assert ib.admit_ms >= ia.done_ms - 1e-6
    assert ibig.state == H.REQ_FAILED and b"memory" in ibig.error
