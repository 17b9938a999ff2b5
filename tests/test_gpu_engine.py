"""GPU parity of the engine against the CPU oracle (run with -m gpu on an MI355X).

Tolerances:
  * FP64 engines, non-chaotic sizes (L <= 1024): per-trajectory CF rows within 1e-6 relative
    L2 of the oracle for >= 99% of trajectories and ensemble CF within 1e-6 relative L2.
    Both integrate the same equations with the same DOP853 controller; they differ in the
    last ulp of sin/cos/exp (own device routines vs glibc) and FMA contraction, which an
    orbiting (chaotic) trajectory amplifies, so a bound per trajectory is stated for the bulk.
  * mixed engine (FP32 surfaces, FP64 state and control): ensemble CF within 1e-4 relative L2.
  * integer and index work (lengths, statuses, the sampled initial conditions) exactly.
"""
import numpy as np
import pytest

import hawaiihybrid_amd as H
import oracle_lib as O

pytestmark = pytest.mark.gpu

ALU3 = 5.29177210903e-11 ** 3


@pytest.fixture(scope="module")
def eng():
    e = H.Engine(0)
    yield e
    e.close()


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def run_and_compare(eng, sys_h, sys_o, smp_h, smp_o, ntraj, L, halves, engine, prec, rtol=1e-12, seed=42):
    ic, _, _ = H.sample_ics(sys_h, smp_h, seed, ntraj)
    ref = O.cf_rows(sys_o, O.CFParams(1e-12, 1e-12, 200.0, L, 40.0, halves, 50000), ic, 16)
    cf_ref = O.ensemble_from_rows(ref["crln"])
    p = H.cf_params(L, rtol=rtol, halves=halves, engine=engine, precision=prec)
    out, st = eng.cf_run(sys_h, p, [eng.to_device_soa(ic)])
    n_items = ntraj * (2 if halves == H.HALVES_BOTH else 1)
    it = eng.items(0, n_items, L, rows=True)
    if halves == H.HALVES_BOTH:
        rows = 0.5 * (it["rows"][0::2] + it["rows"][1::2]) * ALU3
    else:
        rows = it["rows"] * ALU3
    per = np.array([rel_l2(rows[i], ref["crln"][i]) for i in range(ntraj)])
    return dict(cf=out[0].cpu().numpy(), cf_ref=cf_ref, per=per, items=it, ref=ref, stats=st)


def test_engine_is_gfx950(eng):
    assert eng.info()["arch"].startswith("gfx950")


@pytest.mark.parametrize("engine", [H.ENGINE_PLAIN, H.ENGINE_BATCHED])
def test_hear_r0_fp64_matches_oracle(eng, engine):
    """He-Ar at its full configuration: 4096 trajectories at 300 K from R0 = 40 bohr, FP64 at
    rtol = atol = 1e-12, 2048 samples at dt = 200 a.u."""
    r = run_and_compare(eng, H.he_ar(), O.hear(), H.Sampler(H.IC_R0_INJECT, 1, 300.0, 4.0, 40.0, H.PESMIN_HEAR, 40.0),
                        None, 4096, 2048, H.HALVES_FORWARD, engine, H.PREC_FP64)
    assert rel_l2(r["cf"], r["cf_ref"]) < 1e-10
    assert np.quantile(r["per"], 0.99) < 1e-10
    assert np.array_equal(r["items"]["len"], r["ref"]["len_f"])
    assert np.all(r["items"]["status"] == 0)


@pytest.mark.parametrize("engine", [H.ENGINE_PLAIN, H.ENGINE_BATCHED])
def test_hear_rejection_both_halves_matches_oracle(eng, engine):
    """CORRELATION_SINGLE semantics: rejection-sampled free pairs, forward + backward halves,
    crln = 1/2 (c_f + c_b) ALU^3."""
    r = run_and_compare(eng, H.he_ar(), O.hear(), H.Sampler(H.IC_REJECTION, 1, 300.0, 4.0, 40.0, H.PESMIN_HEAR, 0.0),
                        None, 256, 512, H.HALVES_BOTH, engine, H.PREC_FP64)
    assert rel_l2(r["cf"], r["cf_ref"]) < 1e-10
    it, ref = r["items"], r["ref"]
    assert np.array_equal(it["len"][0::2], ref["len_f"]) and np.array_equal(it["len"][1::2], ref["len_b"])


@pytest.mark.parametrize("engine", [H.ENGINE_PLAIN, H.ENGINE_BATCHED])
def test_co2ar_fp64_matches_oracle(eng, engine):
    r = run_and_compare(eng, H.co2_ar(), O.co2ar(),
                        H.Sampler(H.IC_R0_INJECT, 1, 200.0, 4.0, 40.0, H.PESMIN_CO2AR, 40.0), None, 1024, 1024,
                        H.HALVES_FORWARD, engine, H.PREC_FP64)
    assert rel_l2(r["cf"], r["cf_ref"]) < 1e-6
    assert np.quantile(r["per"], 0.99) < 1e-6
    assert np.mean(r["items"]["len"] == r["ref"]["len_f"]) > 0.99
    # same work as the oracle's integrator, to within the chaotic few
    assert abs(r["stats"].nrhs / r["ref"]["nrhs"].sum() - 1.0) < 0.02


@pytest.mark.parametrize("rtol", [1e-10])
def test_co2ar_mixed_within_1e4_of_fp64_oracle(eng, rtol):
    r = run_and_compare(eng, H.co2_ar(), O.co2ar(),
                        H.Sampler(H.IC_R0_INJECT, 1, 200.0, 4.0, 40.0, H.PESMIN_CO2AR, 40.0), None, 1024, 1024,
                        H.HALVES_FORWARD, H.ENGINE_BATCHED, H.PREC_MIXED, rtol=rtol)
    assert rel_l2(r["cf"], r["cf_ref"]) < 1e-4
    assert r["stats"].n_failed == 0


def test_co2ar_mixed_full_ensemble_within_1e4(eng):
    """The benchmark's workload at full size, on its first timed step: 65536 CO2-Ar trajectories
    (trajectories 327680..393215 of the seed-42 stream, T = 200 K, R0 = Rcut = 40 bohr incoming),
    L = 4096, dt = 200.  The mixed engine (FP32 surfaces, FP64 state, rtol 1e-10) hands every
    trajectory that reaches an outer radial turning point (an orbiting complex, chaotic) over to FP64
    at rtol 1e-12; the whole-ensemble CF is then within 1e-4 relative L2 of the FP64 oracle, no
    trajectory filtered out.  Kept in mixed precision, the orbiting trajectories alone move the CF
    by ~1e-2."""
    n, L, step = 65536, 4096, 5
    smp = H.Sampler(H.IC_R0_INJECT, 1, 200.0, 4.0, 40.0, H.PESMIN_CO2AR, 40.0)
    ic = np.ascontiguousarray(H.sample_ics(H.co2_ar(), smp, 42, (step + 1) * n)[0][step * n:])
    import ctypes as C
    cf_ref = np.zeros(L)
    O.lib().orc_cf_ensemble(C.byref(O.co2ar()), C.byref(O.CFParams(1e-12, 1e-12, 200.0, L, 40.0, 1, 50000)),
                            O.ptr(ic), n, 16, 1.0, O.ptr(cf_ref), None)
    x = eng.to_device_soa(ic)
    out, st = eng.cf_run(H.co2_ar(), H.cf_params(L, rtol=1e-10, precision=H.PREC_MIXED), [x])
    err = rel_l2(out[0].cpu().numpy(), cf_ref)
    print(f"  mixed + FP64 orbits: CF rel L2 {err:.3e}, {st.n_orbit} trajectories handed over, "
          f"{st.integrate_ms:.1f} ms", flush=True)
    assert err < 1e-4
    assert 100 < st.n_orbit < 2000
    assert st.n_failed == 0
    off, st_off = eng.cf_run(H.co2_ar(), H.cf_params(L, rtol=1e-10, precision=H.PREC_MIXED, orbit_rtol=-1.0), [x])
    err_off = rel_l2(off[0].cpu().numpy(), cf_ref)
    print(f"  mixed, orbits kept in FP32 surfaces: CF rel L2 {err_off:.3e}", flush=True)
    assert st_off.n_orbit == 0
    assert err_off > 1e-3


def test_co2ar_mixed_equals_the_fp64_engine_where_chaos_sets_the_floor(eng):
    """The first 65536 trajectories of the seed-42 stream hold an orbiting complex (trajectory 35198:
    15 radial turning points, bound for all 4096 samples) whose row depends on the last bit of the
    arithmetic: the FP64 engine at the oracle's tolerance is 2.2e-4 from the CPU oracle on this
    ensemble (the device's own sin/cos, FMA contraction and unit-vector surface chain against glibc and
    the reference's rotation matrices), so no FP64 implementation meets 1e-4 against another here.
    The mixed engine hands that trajectory (and every other orbiting one) to the FP64 pool, and its
    whole-ensemble CF is within 2e-5 of the FP64 engine's."""
    n, L = 65536, 4096
    smp = H.Sampler(H.IC_R0_INJECT, 1, 200.0, 4.0, 40.0, H.PESMIN_CO2AR, 40.0)
    x = eng.to_device_soa(H.sample_ics(H.co2_ar(), smp, 42, n)[0])
    mix, st = eng.cf_run(H.co2_ar(), H.cf_params(L, rtol=1e-10, precision=H.PREC_MIXED), [x])
    f64, _ = eng.cf_run(H.co2_ar(), H.cf_params(L, rtol=1e-12, precision=H.PREC_FP64), [x])
    err = rel_l2(mix[0].cpu().numpy(), f64[0].cpu().numpy())
    print(f"  mixed vs FP64 engine: CF rel L2 {err:.3e}, {st.n_orbit} handed over", flush=True)
    assert err < 2e-5


def test_orbit_handover_rows_are_the_fp64_rows(eng):
    """An item handed over to FP64 is integrated again from its initial condition by the FP64 pool
    at the hand-over tolerance: its row, length and status are then the bits of the batched FP64
    engine at that tolerance, and its right-hand sides count both runs.  L = 4096 samples (8.2e5 au):
    a thermal pair injected at 40 bohr needs ~2e5 au to reach the well, so orbiting only shows on
    the configuration's full trajectory length."""
    n, L = 8192, 4096
    smp = H.Sampler(H.IC_R0_INJECT, 1, 200.0, 4.0, 40.0, H.PESMIN_CO2AR, 40.0)
    x = eng.to_device_soa(H.sample_ics(H.co2_ar(), smp, 7, n)[0])
    _, st = eng.cf_run(H.co2_ar(), H.cf_params(L, rtol=1e-10, precision=H.PREC_MIXED), [x])
    mix = eng.items(0, n, L, rows=True)
    _, _ = eng.cf_run(H.co2_ar(), H.cf_params(L, rtol=1e-12, precision=H.PREC_FP64), [x])
    f64 = eng.items(0, n, L, rows=True)
    same = np.all(mix["rows"] == f64["rows"], axis=1) & (mix["len"] == f64["len"]) & (mix["status"] == f64["status"])
    print(f"  {st.n_orbit} handed over, {int(same.sum())} rows identical to FP64", flush=True)
    assert st.n_orbit > 0
    assert same.sum() >= st.n_orbit
    assert not np.any(mix["status"] == 6)   # no item is left with the hand-over status
    # the handed-over items spent their mixed-precision right-hand sides as well
    assert np.sum(mix["nrhs"][same] > f64["nrhs"][same]) >= st.n_orbit


def test_job_result_identical_alone_and_in_a_mix(eng):
    """Each job's CF depends only on its own items and a fixed reduction order: running it
    alone or streamed with other jobs through one persistent launch gives identical bits."""
    smp = H.Sampler(H.IC_R0_INJECT, 1, 200.0, 4.0, 40.0, H.PESMIN_CO2AR, 40.0)
    ics = [eng.to_device_soa(H.sample_ics(H.co2_ar(), smp, s, 300)[0]) for s in (1, 2, 3)]
    p = H.cf_params(512, engine=H.ENGINE_BATCHED, precision=H.PREC_MIXED, rtol=1e-10)
    alone, _ = eng.cf_run(H.co2_ar(), p, [ics[1]])
    mixed, _ = eng.cf_run(H.co2_ar(), p, ics)
    assert np.array_equal(alone[0].cpu().numpy(), mixed[1].cpu().numpy())


def test_batched_engine_reproduces_the_plain_engine_bit_for_bit(eng):
    """FP64: the resident lane pool (256-lane workgroups, one wavefront per SIMD, wavefront-shared
    sampling, items handed out by a device queue) and the plain engine (one item per thread,
    lane-serial sampling) run the same arithmetic per item, so 3 jobs of 8192 trajectories give
    the same bits in both, and again on a second run of the pool.  Other workgroup sizes are
    refused: the batched engine has one launch shape."""
    smp = H.Sampler(H.IC_R0_INJECT, 1, 200.0, 4.0, 40.0, H.PESMIN_CO2AR, 40.0)
    ic = H.sample_ics(H.co2_ar(), smp, 9, 3 * 8192)[0]
    xs = [eng.to_device_soa(ic[j * 8192:(j + 1) * 8192]) for j in range(3)]
    outs = []
    for block in (0, 256):
        p = H.cf_params(512, engine=H.ENGINE_BATCHED, precision=H.PREC_FP64, rtol=1e-10, block=block)
        cf, st = eng.cf_run(H.co2_ar(), p, xs)
        print(f"  block {st.block} grid {st.grid}", flush=True)
        assert st.block == 256
        outs.append([c.cpu().numpy() for c in cf])
    cf, _ = eng.cf_run(H.co2_ar(), H.cf_params(512, engine=H.ENGINE_PLAIN, precision=H.PREC_FP64, rtol=1e-10), xs)
    outs.append([c.cpu().numpy() for c in cf])
    for o in outs[1:]:
        for a, b in zip(outs[0], o):
            assert np.array_equal(a, b)
    for block in (64, 128):
        with pytest.raises(H.HHError, match="256"):
            eng.cf_run(H.co2_ar(), H.cf_params(512, precision=H.PREC_FP64, rtol=1e-10, block=block), xs)


def test_pool_launch_timing_is_the_run_own_pool(eng):
    """hh_cf_stats.pool_launches / pool_kernel_ms count and time the launches of the run's own
    pool only: for a mixed run with hand-overs they are fewer than the launches of both pools,
    and each launch, timed from its slice's start to its own end, lies within its slice, so
    their sum is at most the slices' wall time."""
    smp = H.Sampler(H.IC_R0_INJECT, 1, 200.0, 4.0, 40.0, H.PESMIN_CO2AR, 40.0)
    ic = H.sample_ics(H.co2_ar(), smp, 42, 2 * 8192)[0]
    xs = [eng.to_device_soa(ic[j * 8192:(j + 1) * 8192]) for j in range(2)]
    _, st = eng.cf_run(H.co2_ar(), H.cf_params(1024, rtol=1e-10, precision=H.PREC_MIXED), xs)
    print(f"  launches {st.launches}, own pool {st.pool_launches} in {st.pool_kernel_ms:.2f} ms, "
          f"slices {st.integrate_ms:.2f} ms, handed over {st.n_orbit}", flush=True)
    assert st.n_orbit > 0
    assert 0 < st.pool_launches < st.launches
    assert 0.0 < st.pool_kernel_ms <= st.integrate_ms * (1.0 + 1e-6) + 1e-3
    _, s64 = eng.cf_run(H.co2_ar(), H.cf_params(1024, rtol=1e-12, precision=H.PREC_FP64), xs)
    assert s64.pool_launches == s64.launches and 0.0 < s64.pool_kernel_ms <= s64.integrate_ms * (1.0 + 1e-6) + 1e-3


def test_mixed_precision_runs_reproducibly(eng):
    """Three runs of the mixed engine on 3 jobs of 8192 trajectories give the same bits and the
    same right-hand-side counts per item."""
    smp = H.Sampler(H.IC_R0_INJECT, 1, 200.0, 4.0, 40.0, H.PESMIN_CO2AR, 40.0)
    ic = H.sample_ics(H.co2_ar(), smp, 42, 3 * 8192)[0]
    xs = [eng.to_device_soa(ic[j * 8192:(j + 1) * 8192]) for j in range(3)]
    runs = []
    for _ in range(3):
        cf, st = eng.cf_run(H.co2_ar(), H.cf_params(512, precision=H.PREC_MIXED, rtol=1e-10), xs)
        assert st.block == 256
        runs.append(([c.cpu().numpy() for c in cf], [eng.items(j, 8192, 512)["nrhs"] for j in range(3)]))
    for cf, nrhs in runs[1:]:
        for a, b in zip(runs[0][0], cf):
            assert np.array_equal(a, b)
        for a, b in zip(runs[0][1], nrhs):
            assert np.array_equal(a, b)


@pytest.mark.parametrize("engine", [H.ENGINE_PLAIN, H.ENGINE_BATCHED])
def test_step_limit_drops_trajectory_like_cvode(eng, engine):
    """max_steps accepted steps without reaching the next sample ends a half with the step-limit
    status (CVode's CV_TOO_MUCH_WORK, src/trajectory.c:15, 62).  When that half is the
    trajectory's last, correlation_eval returns -1 and the caller drops the trajectory
    (src/hawaii.c:3001), which still counts in ntraj.  At 1000 K with max_steps = 6, one of
    these 256 He-Ar trajectories ends that way in the oracle."""
    smp = H.Sampler(H.IC_REJECTION, 1, 1000.0, 4.0, 40.0, H.PESMIN_HEAR, 0.0)
    ic, _, _ = H.sample_ics(H.he_ar(), smp, 5, 256)
    ref = O.cf_rows(O.hear(), O.CFParams(1e-12, 1e-12, 200.0, 256, 40.0, O.HALVES_BOTH, 6), ic, 8)
    assert np.sum(ref["status"] == 4) == 1
    p = H.cf_params(256, halves=H.HALVES_BOTH, engine=engine, max_steps=6)
    out, st = eng.cf_run(H.he_ar(), p, [eng.to_device_soa(ic)])
    it = eng.items(0, 512, 256)
    assert np.array_equal(it["status"][1::2], ref["status"])
    assert st.n_stepfail >= 1
    cf_ref = O.ensemble_from_rows(ref["crln"])
    assert rel_l2(out[0].cpu().numpy(), cf_ref) < 1e-10


def test_cf_to_sf_reference_kat(eng, golden):
    g = golden("cf_sf_hear_1000.npz")
    cf = eng.torch.from_numpy(g["cf"].copy()).cuda()
    nu, sf, _ = eng.cf_to_sf(cf, 200.0)
    nu, sf = nu.cpu().numpy(), sf.cpu().numpy()
    peak = np.abs(g["sf"]).max()
    assert np.max(np.abs(nu - g["sf_nu"])) < 1e-6
    assert np.max(np.abs(sf - g["sf"])) < 1e-9 * peak
    big = np.abs(g["sf"]) > 1e-8 * peak
    assert np.max(np.abs(sf[big] / g["sf"][big] - 1.0)) < 1e-4


@pytest.mark.parametrize("n", [2048, 4096])
def test_cf_to_sf_matches_oracle(eng, n):
    rng = np.random.default_rng(n)
    cf = np.exp(-np.arange(n) / 300.0) * (1 + 0.1 * rng.standard_normal(n))
    t = 200.0 * np.arange(n)
    onu, osf = O.cf_to_sf(t, cf)
    d = eng.torch.from_numpy(cf.copy()).cuda()
    nu, sf, _ = eng.cf_to_sf(d, 200.0)
    assert np.max(np.abs(sf.cpu().numpy() - osf)) < 1e-12 * np.abs(osf).max()
    assert np.max(np.abs(nu.cpu().numpy() - onu)) < 1e-9


def test_m0_matches_oracle_streams(eng):
    """M0 over 64 independent streams seeded seed+s equals the oracle's calculate_M0 per
    stream combined as mpi_calculate_M0 combines ranks."""
    smp = H.Sampler(H.IC_REJECTION, 1, 300.0, 4.0, 40.0, H.PESMIN_HEAR, 0.0)
    m, q, _ = eng.m0(H.he_ar(), smp, 7, 64 * 200, 64)
    mo = [O.m0(O.hear(), O.Sampler(0, 1, 300.0, 4.0, 40.0, O.PESMIN_HEAR, 0.0), 7 + s, 200)[0] for s in range(64)]
    assert abs(m - np.mean(mo)) < 1e-12 * abs(np.mean(mo))
    assert q > 0


@pytest.mark.parametrize("which", ["hear", "co2ar"])
def test_m2_matches_oracle_streams(eng, which):
    """M2 (calculate_M2) over 64 streams seeded seed+s equals the oracle's per-stream M2
    combined as mpi_calculate_M2 combines ranks.  The per-point value differences the dipole
    with h = 1e-3, which scales the surfaces' last-bit differences by ~1/h: 1e-8 relative."""
    if which == "hear":
        sh, so, T, ps, pesmin = H.he_ar(), O.hear(), 300.0, H.PAIR_FREE, H.PESMIN_HEAR
    else:
        sh, so, T, ps, pesmin = H.co2_ar(), O.co2ar(), 100.0, H.PAIR_BOUND, H.PESMIN_CO2AR
    smp = H.Sampler(H.IC_REJECTION, ps, T, 4.0, 40.0, pesmin, 0.0)
    m, q, _ = eng.m2(sh, smp, 7, 64 * 100, 64, 1.5)
    res = [O.m2(so, O.Sampler(0, ps, T, 4.0, 40.0, pesmin, 0.0), 7 + s, 100, 1.5) for s in range(64)]
    mo = np.mean([r[0] for r in res])
    # q: the per-stream sums of squared deviations averaged, as the ranks' are
    qo = 1.5 * H.SECOND_COEFF * np.sqrt(np.mean([(r[1] / (1.5 * H.SECOND_COEFF)) ** 2 * 100 * 99 for r in res])
                                        / 100 / 99)
    assert abs(m - mo) < 1e-8 * abs(mo)
    assert abs(q - qo) < 1e-7 * qo


@pytest.mark.parametrize("which", ["hear", "co2ar"])
def test_dipole_rate_matches_oracle(eng, which):
    """hh_eval(HH_EVAL_DIPOLE_RATE), |d mu/dt|^2 at sampled phase points, against the oracle."""
    if which == "hear":
        sh, so = H.he_ar(), O.hear()
        smp = H.Sampler(H.IC_REJECTION, 3, 300.0, 3.5, 40.0, H.PESMIN_HEAR, 0.0)
    else:
        sh, so = H.co2_ar(), O.co2ar()
        smp = H.Sampler(H.IC_REJECTION, 3, 300.0, 4.0, 40.0, H.PESMIN_CO2AR, 0.0)
    y, _, _ = H.sample_ics(sh, smp, 23, 400)
    f = eng.eval(sh, H.EVAL_DIPOLE_RATE, eng.to_device_soa(y)).cpu().numpy()[0]
    fo = O.dipole_rate_sq(so, y)
    assert np.all(fo > 0)
    np.testing.assert_allclose(f, fo, rtol=1e-8)


def test_m0_rejects_unreachable_sampler(eng):
    bad = H.Sampler(H.IC_REJECTION, H.PAIR_BOUND, 300.0, 39.0, 40.0, H.PESMIN_HEAR, 0.0)
    with pytest.raises(H.HHError):
        eng.m0(H.he_ar(), bad, 1, 128, 64)


@pytest.mark.parametrize("which", ["hear", "co2ar"])
def test_batched_surfaces_match_oracle(eng, which):
    """hh_eval (device pes_lab / dipole_lab / Hamiltonian / rhs in FP64) at sampled phase points
    against the oracle's restated surfaces.  The device evaluates its own sin/cos and the
    CO2-Ar tables re-expanded about each interval's centre: agreement to rounding."""
    if which == "hear":
        sh, so = H.he_ar(), O.hear()
        smp = H.Sampler(H.IC_REJECTION, 3, 300.0, 3.5, 40.0, H.PESMIN_HEAR, 0.0)
    else:
        sh, so = H.co2_ar(), O.co2ar()
        smp = H.Sampler(H.IC_REJECTION, 3, 300.0, 4.0, 40.0, H.PESMIN_CO2AR, 0.0)
    y, _, _ = H.sample_ics(sh, smp, 17, 400)
    yd = eng.to_device_soa(y)
    V = eng.eval(sh, H.EVAL_PES, yd).cpu().numpy()[0]
    D = eng.eval(sh, H.EVAL_DIPOLE, yd).cpu().numpy()
    E = eng.eval(sh, H.EVAL_HAMILTONIAN, yd).cpu().numpy()[0]
    F = eng.eval(sh, H.EVAL_RHS, yd).cpu().numpy()
    L = O.lib()
    for i in range(y.shape[0]):
        q = np.ascontiguousarray(y[i, 0::2])
        v = L.orc_pes(so, O.ptr(q))
        assert abs(V[i] - v) <= 1e-11 * abs(v) + 1e-16
        mu = np.zeros(3)
        L.orc_dipole(so, O.ptr(q), O.ptr(mu))
        assert np.allclose(D[:, i], mu, rtol=1e-11, atol=1e-16)
        assert abs(E[i] - L.orc_hamiltonian(so, O.ptr(np.ascontiguousarray(y[i])))) <= 1e-11 * abs(E[i]) + 1e-15
        f = np.zeros(y.shape[1])
        L.orc_rhs(so, O.ptr(np.ascontiguousarray(y[i])), O.ptr(f))
        assert np.allclose(F[:, i], f, rtol=1e-9, atol=1e-15)
