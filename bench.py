#!/usr/bin/env python3
"""CO2-Ar trajectory-ensemble benchmark on one MI355X.

Workload: a step is one CO2-Ar dipole correlation function over `--ntraj` trajectories
(T = 200 K, R0 = Rcut = 40 bohr, incoming radial momentum, L = 4096 samples at dt = 200 a.u.;
IC mode r0_inject, forward half), from consecutive chunks of the seed-42 Mersenne Twister
stream: integrate every trajectory, record mu(0).mu(t_k), reduce to the ensemble CF.  The
initial conditions are sampled on the host (bit-exact reference stream) and resident in HBM
before any timed region.  W warm-up steps, then K timed steps.

Engines, timed on the same K batches:
  plain    FP64 at the reference tolerance (rtol = atol = 1e-12), one trajectory per thread,
           one launch per step (each launch waits for its slowest trajectory)
  batched  resident lane pool (time-sliced persistent kernel, longest-first queue, continuous
           batching: all K steps are admitted together and stream through the pool), FP32
           surfaces with FP64 state and error control at rtol = atol = --rtol-mixed; a trajectory
           that reaches an outer radial turning point (an orbiting complex, chaotic) is handed
           over to the FP64 pool at rtol = atol = 1e-12 and integrated again from its initial
           condition, concurrently with the rest (its own stream)
`value` is the batched engine's trajectories/s over the K steps; `speedup_vs_plain` divides it
by the plain engine's.  Also reported: the plain engine given all K steps in one launch, the
batched engine in FP64, the batched mixed engine on the K steps repeated --deep times in one batch
(`batched_mixed_traj_per_s_deep`: the K-step batch ends with a tail of a few long trajectories,
which a deeper batch runs beside later steps' bulk), accuracy against the CPU oracle for the first timed step (FP64 DOP853
at 1e-12), the roofline of the pool kernel, a serving run with Poisson arrivals (p50/p99
request latency), the N2-Ar PIP-NN surfaces on the matrix cores (2^20 phase points at
100-500 K; BF16 and split-BF16 errors against the FP64 oracle), and the CPU oracle timed on
the host cores.  One JSON line on stdout.

Other modes: `--single INDEX` times one trajectory of the workload alone on each engine (the
latency that bounds a batch's tail); `--traffic-from FETCH_DIR WRITE_DIR` turns two rocprofv3
counter runs into the pool kernel's bytes per launch (profiles/).
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.join(HERE, "tests"))

ALU3 = 5.29177210903e-11 ** 3
MI355X_FP64_VECTOR_TFLOPS = 78.6
MI355X_FP32_VECTOR_TFLOPS = 157.3
MI355X_HBM_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def traffic_from_counters(fetch_dir, write_dir, out_path, kernel="pool_cf_kernel<1, float, 256, 1>"):
    """Bytes per launch of the mixed CO2-Ar pool kernel from two rocprofv3 counter runs of this
    script with --calibrate-copy (one with --pmc FETCH_SIZE, one with --pmc WRITE_SIZE; CSV
    output).  Each counter is scaled by what it reads for the 1 GiB device-to-device copy that
    --calibrate-copy runs first (the largest __amd_rocclr_copyBuffer dispatch), then averaged
    over the kernel's dispatches.  Writes the JSON that the roofline's `traffic` reads."""
    import csv
    import glob

    def per_dispatch(d, counter):
        vals = {}
        for f in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
            for r in csv.DictReader(open(f)):
                if r.get("Counter_Name") != counter:
                    continue
                key = (r["Kernel_Name"], r["Dispatch_Id"])
                vals[key] = vals.get(key, 0.0) + float(r["Counter_Value"] or 0)
        return vals

    args = "--steps 20 --warmup 5 --no-serve --no-n2ar --no-ch4co2 --no-oracle --no-hbm --deep 0 --calibrate-copy"
    res = {"kernel": kernel,
           "workload": "CO2-Ar, 65536 trajectories per step, T=200 K, L=4096, batched mixed engine rtol=1e-10; "
                       "every dispatch of the kernel in the run (5 warm-up and 20 timed steps)",
           "commands": [f"rocprofv3 --pmc FETCH_SIZE --output-format csv -d {fetch_dir} -o fetch -- python3 bench.py {args}",
                        f"rocprofv3 --pmc WRITE_SIZE --output-format csv -d {write_dir} -o write -- python3 bench.py {args}",
                        f"python3 bench.py --traffic-from {fetch_dir} {write_dir}"],
           "counters": {}}
    for name, d in (("fetch", fetch_dir), ("write", write_dir)):
        counter = "FETCH_SIZE" if name == "fetch" else "WRITE_SIZE"
        v = per_dispatch(d, counter)
        copy = max((x for (k, _), x in v.items() if "copyBuffer" in k), default=0.0)
        scale = (1 << 30) / 1024.0 / copy if copy > 0 else None
        ks = [x for (k, _), x in v.items() if kernel in k]
        res["counters"][name] = {"counter": counter, "copy_kib": copy, "scale": scale, "dispatches": len(ks),
                                 "kib_per_dispatch": float(np.mean(ks)) if ks else None}
    f, w = res["counters"]["fetch"], res["counters"]["write"]
    if f["kib_per_dispatch"] is not None and w["kib_per_dispatch"] is not None and f["scale"] and w["scale"]:
        res["hbm_bytes_per_launch"] = 1024.0 * (f["kib_per_dispatch"] * f["scale"] + w["kib_per_dispatch"] * w["scale"])
    json.dump(res, open(out_path, "w"), indent=1)
    return res


def hbm_section(H, eng, stm, mixed_len, plain_len, plain_reduce_ms, L, prmu_ntraj, prmu_L, sf_n):
    """HBM rooflines of the data-movement kernels beside the integrator, bytes counted from
    what each kernel loads and stores:
      cf_reduce   the CF reduction (cf_reduce_partial + _final per job): 8 B per recorded
                  sample (rows are read up to the item's length), 8 B of length and status per
                  item, and the per-chunk partial sums and compensations (256 items per chunk,
                  16 B per column and chunk) written and read back, 8 B per CF value written.
                  `alone`: the first timed step's job reduced after the plain engine's
                  integration (nothing else on the device); `in_batch`: the K jobs of the timed
                  mixed batch, each reduced while the pools integrate the others.
      prmu        the PR/mu spectrum pass (pack + Connes window, four-step FFT of two complex
                  series per trajectory, power, accumulation) of He-Ar trajectories at FFT length
                  prmu_L: pack reads 24 B per recorded sample and writes 32 B per point, each FFT
                  step reads and writes 64 B per point, power reads 48 B and writes 8 B per kept
                  frequency, accumulation reads 8 B per frequency and trajectory; time = the
                  engine's events around the pass (integration excluded).
      cf_to_sf    idct_cf_to_sf of an sf_n-point CF (Connes window, IDCT via an M = 2n-point
                  complex FFT): the pack kernel reads 8 B and writes 16 B per point of M, each of
                  the log2(M) global radix-2 passes (M > 8192) reads and writes 32 B per point of
                  M, the unpack kernel reads 16 B and writes 24 B per CF point; `bytes_algorithmic`
                  is the transform's own input and output (8 B read, 16 B written per CF point)."""
    import torch
    out = {}

    def reduce_bytes(lens):
        nchunks = -(-len(lens[0]) // 256)
        return sum(8.0 * float(np.minimum(ln, L).sum()) + 8.0 * len(ln) + 2 * nchunks * L * 16.0 + 8.0 * L
                   for ln in lens)

    def rate(b, ms):
        return {"bytes": b, "ms": ms, "GBps": b / (ms * 1e-3) / 1e9, "frac_hbm": b / (ms * 1e-3) / 1e9 / MI355X_HBM_GBS}

    out["cf_reduce"] = {"alone": dict(rate(reduce_bytes([plain_len]), plain_reduce_ms), jobs=1),
                        "in_batch": dict(rate(reduce_bytes(mixed_len), stm.reduce_ms), jobs=len(mixed_len))}
    # PR/mu spectrum
    smp = H.Sampler(H.IC_R0_INJECT, H.PAIR_FREE, 300.0, 4.0, 40.0, H.PESMIN_HEAR, 12.0)
    ic, _, _ = H.sample_ics(H.he_ar(), smp, 7, prmu_ntraj)
    nfreq = prmu_L // 2 + 1
    pp = H.PrmuParams(1e-12, 1e-12, 200.0, 12.0, prmu_L, nfreq, H.PREC_FP64, 0, 1)
    x = eng.to_device_soa(ic)
    eng.prmu(H.he_ar(), pp, x)                        # warm-up (buffers)
    _, counted, nc, st = eng.prmu(H.he_ar(), pp, x)
    rec = float(np.minimum(eng.items(0, prmu_ntraj, prmu_L)["len"], prmu_L).sum())
    nt = prmu_ntraj
    pb = 24.0 * rec + nt * prmu_L * (32.0 + 2 * 64.0) + nt * nfreq * (48.0 + 8.0 + 8.0)
    out["prmu_spectrum"] = dict(rate(pb, st.reduce_ms), trajectories=nt, fft_length=prmu_L, nfreq=nfreq)
    del x
    # CF -> SF
    cf = torch.rand(sf_n, dtype=torch.float64, device="cuda:0")
    eng.cf_to_sf(cf.clone(), 200.0)
    _, _, ms = eng.cf_to_sf(cf, 200.0)
    M = 2 * sf_n
    passes = int(math.log2(M)) if M > 8192 else 1
    mb = 24.0 * M + passes * 32.0 * M + 40.0 * sf_n
    out["cf_to_sf"] = dict(rate(mb, ms), points=sf_n, fft_passes=passes, bytes_algorithmic=24.0 * sf_n,
                           GBps_algorithmic=24.0 * sf_n / (ms * 1e-3) / 1e9)
    del cf
    torch.cuda.synchronize()
    return out


def single_trajectory(H, eng, index, ntraj, L, rtol_mixed):
    """One trajectory of the workload alone on the device, on each engine (a warm-up run, then a
    timed one): wall time, right-hand sides and time per right-hand side.  The longest
    trajectories of a batch run this way at its end, so this is the latency that bounds the
    batch's tail."""
    import torch
    smp = H.Sampler(H.IC_R0_INJECT, H.PAIR_FREE, 200.0, 4.0, 40.0, H.PESMIN_CO2AR, 40.0)
    ic = H.sample_ics(H.co2_ar(), smp, 42, ntraj)[0]
    if not 0 <= index < ntraj:
        raise SystemExit(f"--single: index {index} outside [0, {ntraj})")
    x = eng.to_device_soa(ic[index:index + 1])
    out = {"trajectory": index, "workload": f"CO2-Ar, T=200 K, R0=Rcut=40 bohr incoming, L={L}, dt=200 a.u."}
    for name, p in (("plain_fp64", H.cf_params(L, rtol=1e-12, engine=H.ENGINE_PLAIN, precision=H.PREC_FP64)),
                    ("batched_fp64", H.cf_params(L, rtol=1e-12, engine=H.ENGINE_BATCHED, precision=H.PREC_FP64)),
                    ("batched_mixed", H.cf_params(L, rtol=rtol_mixed, engine=H.ENGINE_BATCHED,
                                                  precision=H.PREC_MIXED))):
        eng.cf_run(H.co2_ar(), p, [x])
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        _, st = eng.cf_run(H.co2_ar(), p, [x])
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) * 1e3
        out[name] = {"ms": ms, "nrhs": st.nrhs, "us_per_rhs": 1e3 * ms / max(st.nrhs, 1),
                     "handed_over_to_fp64": st.n_orbit}
        log(name, out[name])
    print(json.dumps(out), flush=True)


def serving_mix(H, eng, rate_hz, seconds, seed, log_every=25):
    """Poisson arrivals of CF requests (He-Ar or CO2-Ar, T ~ U[50, 500] K, ntraj ~ 2^U[10, 16])
    and M0 requests (10^U[6, 8] accepted points of the rejection sampler), 4:1; inputs staged in
    HBM before the clock starts.  M0 slices run on their own stream beside the pools' slices.
    Latency percentiles per request class."""
    import torch
    rng = np.random.default_rng(seed)
    n_req = max(1, int(rng.poisson(rate_hz * seconds)))
    t_arr = np.sort(rng.uniform(0.0, seconds, n_req))
    reqs = []
    for i in range(n_req):
        if rng.uniform() < 0.8:
            co2 = rng.uniform() < 0.5
            T = float(rng.uniform(50.0, 500.0))
            ntraj = int(2 ** rng.uniform(10, 16))
            if co2:
                sysd, pes, L = H.co2_ar(), H.PESMIN_CO2AR, 4096
                p = H.cf_params(L, rtol=1e-10, precision=H.PREC_MIXED)
            else:
                sysd, pes, L = H.he_ar(), H.PESMIN_HEAR, 2048
                p = H.cf_params(L, rtol=1e-12, precision=H.PREC_FP64)
            smp = H.Sampler(H.IC_R0_INJECT, H.PAIR_FREE, T, 4.0, 40.0, pes, 40.0)
            ic = eng.to_device_soa(H.sample_ics(sysd, smp, 1000 + i, ntraj)[0])
            reqs.append(("cf", sysd, p, ic, ntraj))
        else:
            co2 = rng.uniform() < 0.5
            sysd = H.co2_ar() if co2 else H.he_ar()
            pes = H.PESMIN_CO2AR if co2 else H.PESMIN_HEAR
            T = float(rng.uniform(50.0, 500.0))
            npts = int(10 ** rng.uniform(6, 8))
            smp = H.Sampler(H.IC_REJECTION, H.PAIR_FREE, T, 4.0, 40.0, pes, 0.0)
            reqs.append(("m0", sysd, smp, npts, 65536))
    torch.cuda.synchronize()
    s = H.Session(eng, slice_ms=10.0)
    ids = []
    t0 = time.perf_counter()
    for i, (ta, r) in enumerate(zip(t_arr, reqs)):
        now = time.perf_counter() - t0
        if ta > now:
            time.sleep(ta - now)
        if r[0] == "cf":
            rid, _ = s.submit_cf(r[1], r[2], r[3])
        else:
            rid = s.submit_m0(r[1], r[2], 7 + i, r[3], r[4])
        cls = ("cf-co2ar" if r[1].system == H.SYS_CO2AR else "cf-hear") if r[0] == "cf" else "m0"
        ids.append((rid, r[0], r[4] if r[0] == "cf" else r[3], cls))
    infos = [(kind, n, s.poll(rid, 600000.0), cls) for rid, kind, n, cls in ids]
    wall = time.perf_counter() - t0
    s.close()
    done = [(k, n, inf, c) for k, n, inf, c in infos if inf.state == H.REQ_DONE]
    lat = np.array([inf.latency_ms for _, _, inf, _ in done])
    failed = len(infos) - len(done)
    traj = sum(n for k, n, _, _ in done if k == "cf")
    pts = sum(n for k, n, _, _ in done if k == "m0")
    out = {"requests": len(infos), "cf_requests": int(sum(1 for i in infos if i[0] == "cf")),
           "m0_requests": int(sum(1 for i in infos if i[0] == "m0")), "failed": failed,
           "arrival_rate_hz": rate_hz, "arrival_window_s": seconds, "wall_s": wall,
           "p50_ms": float(np.percentile(lat, 50)) if lat.size else None,
           "p99_ms": float(np.percentile(lat, 99)) if lat.size else None,
           "aggregate_traj_per_s": traj / wall, "m0_points_per_s": pts / wall,
           "m0_points_range": "10^6 - 10^8", "per_class": {}}
    for cls in ("cf-hear", "cf-co2ar", "m0"):
        lc = np.array([inf.latency_ms for _, _, inf, c in done if c == cls])
        out["per_class"][cls] = {"requests": int(sum(1 for i in infos if i[3] == cls)),
                                 "p50_ms": float(np.percentile(lc, 50)) if lc.size else None,
                                 "p99_ms": float(np.percentile(lc, 99)) if lc.size else None}
    return out


def n2ar_section(H, eng, npoints, sample_check):
    """N2-Ar PIP-NN surfaces (the reference's trained PES and dipole models, and the synthetic
    17 -> 3 x 128 tanh shape) at `npoints` phase points from q_generator + p_generator at 64
    temperatures spread over 100-500 K: device time per evaluation, points/s, matrix-core
    TFLOP/s of the network products, and the error against the FP64 oracle on a subsample."""
    import ctypes as C
    import torch
    import oracle_lib as O
    from hawaiihybrid_amd import n2ar
    HT = 2.1947463136320e5
    nchunk = 64
    per = (npoints + nchunk - 1) // nchunk
    pts = np.concatenate([n2ar.phase_points(per, 100.0 + 400.0 * (c + 0.5) / nchunk, seed=1000 + c)
                          for c in range(nchunk)])[:npoints]
    q = n2ar.lab_q(pts)
    qd = torch.from_numpy(q).to("cuda:0")
    W = n2ar.load_weights()
    m, keep = O.n2ar_model(W, n2ar.PES_EN_INF)
    idx = np.linspace(0, npoints - 1, sample_check).astype(np.int64)
    refE = np.array([O.lib().orc_n2ar_pes(C.byref(m), O.ptr(np.ascontiguousarray(q[:, i]))) for i in idx]) * HT
    refmu = np.zeros((3, len(idx)))
    mu = np.zeros(3)
    for j, i in enumerate(idx):
        O.lib().orc_n2ar_dipole(C.byref(m), O.ptr(np.ascontiguousarray(q[:, i])), O.ptr(mu))
        refmu[:, j] = mu
    musc = np.max(np.linalg.norm(refmu, axis=0))
    # algorithmic matrix-core FLOPs per point: PES 17-32-1; dipole 2 models x (119-16-2, 119-16-2, 70-8-1)
    f_pes = 2 * (17 * 32 + 32)
    f_dip = 2 * 2 * (2 * (119 * 16 + 16 * 2) + 70 * 8 + 8)
    f_syn = 2 * (17 * 128 + 2 * 128 * 128 + 128)
    out = {"points": npoints, "T_K": "64 temperatures evenly over 100-500", "check_points": len(idx)}
    for prec, name in ((n2ar.MLP_BF16, "bf16"), (n2ar.MLP_BF16X3, "bf16x3")):
        model = n2ar.PipNN(eng, precision=prec, weights=W)
        model.eval(qd)                      # warm-up (code objects)
        times = []
        for _ in range(5):
            E, MU, ms = model.eval(qd)
            times.append(ms)
        ms = float(np.median(times))
        E = E.cpu().numpy()[idx] * HT
        MU = MU.cpu().numpy()[:, idx]
        dE = np.abs(E - refE)
        out[name] = {"ms": ms, "points_per_s": npoints / (ms * 1e-3),
                     "matrix_tflops": npoints * (f_pes + f_dip) / (ms * 1e-3) / 1e12,
                     "dE_cm1_median": float(np.median(dE)), "dE_cm1_p99": float(np.quantile(dE, 0.99)),
                     "dE_cm1_p999": float(np.quantile(dE, 0.999)), "dE_cm1_max": float(dE.max()),
                     "dE_rel_max_above_100cm1": float(np.max((dE / np.abs(refE))[np.abs(refE) > 100.0])),
                     "dmu_over_max_mu": float(np.max(np.abs(MU - refmu)) / musc),
                     "dmu_over_max_mu_p99": float(np.quantile(np.max(np.abs(MU - refmu), axis=0), 0.99) / musc)}
        model.close()
        syn = n2ar.synthetic_pes(128)
        sm = n2ar.PipNN(eng, precision=prec, pes_layers=syn, pes_act=n2ar.ACT_TANH, with_dipole=False, weights=W)
        sm.eval(qd, dipole=False)
        times = [sm.eval(qd, dipole=False)[2] for _ in range(5)]
        ms = float(np.median(times))
        out[name]["synthetic_17_128x3_1_tanh"] = {"ms": ms, "points_per_s": npoints / (ms * 1e-3),
                                                   "matrix_tflops": npoints * f_syn / (ms * 1e-3) / 1e12}
        sm.close()
    return out


def ch4co2_section(H, eng, ntraj, L, check_n):
    """CH4-CO2 (spherical top + linear rotor, 16-dim phase space; stand-in PES): `ntraj`
    R0-injected trajectories at 300 K (R0 = Rcut = 40 bohr), L samples at dt = 200 a.u., in jobs of
    65536.  The mixed engine (stand-in PES and dipole in FP32, state and error control in FP64,
    rtol 1e-10, orbiting pairs handed over to FP64 at 1e-12) against the all-FP64 engine at
    1e-12 on the whole ensemble (CF relative L2, per-trajectory energy drift |E(end) - E0|), and
    both against the CPU oracle on the first `check_n` trajectories."""
    import oracle_lib as O
    import torch
    smp = H.Sampler(H.IC_R0_INJECT, H.PAIR_FREE, 300.0, 4.75, 40.0, H.PESMIN_CH4CO2, 40.0)
    ic, _, _ = H.sample_ics(H.ch4_co2(), smp, 42, ntraj)
    job = 65536
    xs = [eng.to_device_soa(np.ascontiguousarray(ic[j:j + job])) for j in range(0, ntraj, job)]
    out = {"trajectories": ntraj, "L": L, "T_K": 300.0, "jobs": len(xs)}
    res = {}
    for name, prec, rtol in (("mixed", H.PREC_MIXED, 1e-10), ("fp64", H.PREC_FP64, 1e-12)):
        p = H.cf_params(L, rtol=rtol, precision=prec)
        eng.cf_run(H.ch4_co2(), p, xs[:1])          # warm-up (code objects, pool)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        cfs, st = eng.cf_run(H.ch4_co2(), p, xs)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        drift = np.concatenate([eng.items(j, min(job, ntraj - j * job), L)["e_drift"] for j in range(len(xs))])
        cf = sum(c.cpu().numpy() * min(job, ntraj - j * job) for j, c in enumerate(cfs)) / ntraj
        res[name] = (cf, drift)
        out[name] = {"traj_per_s": ntraj / dt, "s": dt, "nrhs": st.nrhs, "rtol": rtol,
                     "handed_over_to_fp64": st.n_orbit,
                     "energy_drift_hartree_median": float(np.median(drift)),
                     "energy_drift_hartree_p99": float(np.quantile(drift, 0.99)),
                     "energy_drift_hartree_max": float(drift.max())}
        log(f"CH4-CO2 {name}: {ntraj / dt:.0f} traj/s, {st.nrhs} right-hand sides, {st.n_orbit} handed over")
    out["cf_rel_l2_mixed_vs_fp64_engine"] = rel(res["mixed"][0], res["fp64"][0])
    if check_n:
        n = check_n
        ref = O.cf_rows(O.ch4co2(), O.CFParams(1e-12, 1e-12, 200.0, L, 40.0, 1, 50000), ic[:n],
                        min(16, os.cpu_count() or 1))
        cf_ref = O.ensemble_from_rows(ref["crln"])
        sub = {}
        for name, prec, rtol in (("mixed", H.PREC_MIXED, 1e-10), ("fp64", H.PREC_FP64, 1e-12)):
            cfs, _ = eng.cf_run(H.ch4_co2(), H.cf_params(L, rtol=rtol, precision=prec),
                                [eng.to_device_soa(np.ascontiguousarray(ic[:n]))])
            sub[name] = rel(cfs[0].cpu().numpy(), cf_ref)
        out["oracle_subsample"] = {"trajectories": n, "cf_rel_l2_mixed": sub["mixed"], "cf_rel_l2_fp64": sub["fp64"]}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=8)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--ntraj", type=int, default=65536)
    ap.add_argument("--L", type=int, default=4096)
    ap.add_argument("--rtol-mixed", type=float, default=1e-10)
    ap.add_argument("--cpu-sample", type=int, default=131072)
    ap.add_argument("--serve-seconds", type=float, default=8.0)
    ap.add_argument("--serve-rate", type=float, default=24.0, help="serving mix: Poisson arrivals per second")
    ap.add_argument("--deep", type=int, default=4,
                    help="also time the batched mixed engine on the timed steps repeated this many times in one batch (0: skip)")
    ap.add_argument("--no-oracle", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-n2ar", action="store_true")
    ap.add_argument("--no-ch4co2", action="store_true")
    ap.add_argument("--no-hbm", action="store_true")
    ap.add_argument("--traffic-from", nargs=2, metavar=("FETCH_DIR", "WRITE_DIR"),
                    help="no run: turn two rocprofv3 counter runs (FETCH_SIZE, WRITE_SIZE) of this script "
                         "into profiles/pool_cf_kernel_co2ar_mixed_traffic.json")
    ap.add_argument("--ch4co2-ntraj", type=int, default=262144)
    ap.add_argument("--ch4co2-L", type=int, default=1024)
    ap.add_argument("--n2ar-points", type=int, default=1 << 20)
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run a 1 GiB device-to-device copy first (reference kernel for HBM byte counters)")
    ap.add_argument("--single", type=int, default=None, metavar="INDEX",
                    help="latency of one trajectory alone on the device: trajectory INDEX of the first "
                         "--ntraj of the workload's stream, on the plain FP64, batched FP64 and batched "
                         "mixed engines (one JSON line; what bounds a batch's tail)")
    args = ap.parse_args()
    if args.traffic_from:
        r = traffic_from_counters(args.traffic_from[0], args.traffic_from[1],
                                  os.path.join(HERE, "profiles", "pool_cf_kernel_co2ar_mixed_traffic.json"))
        print(json.dumps(r))
        return
    if args.gpus != 1:
        print("bench.py runs on exactly one GPU: this engine is measured on one MI355X; running the "
              "ensemble over several GPUs is a design note (trajectories shard by index), not a "
              "measured configuration", file=sys.stderr)
        sys.exit(2)

    import torch
    import hawaiihybrid_amd as H

    eng = H.Engine(0)
    log("device", eng.info())
    if args.single is not None:
        single_trajectory(H, eng, args.single, args.ntraj, args.L, args.rtol_mixed)
        return
    if args.calibrate_copy:
        a = torch.ones(1 << 27, dtype=torch.float64, device="cuda:0")   # 1 GiB
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b
    sysd = H.co2_ar()
    smp = H.Sampler(H.IC_R0_INJECT, H.PAIR_FREE, 200.0, 4.0, 40.0, H.PESMIN_CO2AR, 40.0)
    W, K, n, L = args.warmup, args.steps, args.ntraj, args.L
    t0 = time.time()
    ic_all, _, _ = H.sample_ics(sysd, smp, 42, (W + K) * n)
    log(f"sampled {(W + K) * n} initial conditions in {time.time() - t0:.2f} s (host, reference stream)")
    ics = [eng.to_device_soa(ic_all[b * n:(b + 1) * n]) for b in range(W + K)]
    warm, timed = ics[:W], ics[W:]
    torch.cuda.synchronize()

    p_plain = H.cf_params(L, rtol=1e-12, engine=H.ENGINE_PLAIN, precision=H.PREC_FP64)
    p_mixed = H.cf_params(L, rtol=args.rtol_mixed, engine=H.ENGINE_BATCHED, precision=H.PREC_MIXED)
    p_b64 = H.cf_params(L, rtol=1e-12, engine=H.ENGINE_BATCHED, precision=H.PREC_FP64)

    # plain FP64: one launch per step
    for x in warm:
        eng.cf_run(sysd, p_plain, [x])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    plain_k_ms = []
    for x in timed:
        out, st = eng.cf_run(sysd, p_plain, [x])
        plain_k_ms.append(st.integrate_ms)
    torch.cuda.synchronize()
    t_plain = time.perf_counter() - t0
    log(f"plain fp64, one launch per step: {t_plain * 1e3:.1f} ms for {K} steps -> {K * n / t_plain:.0f} traj/s")

    # plain FP64 with every timed step in one launch (for comparison)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    eng.cf_run(sysd, p_plain, timed)
    torch.cuda.synchronize()
    t_plain_one = time.perf_counter() - t0
    log(f"plain fp64, all steps in one launch: {t_plain_one * 1e3:.1f} ms")

    # batched FP64 (same tolerance as plain)
    if warm:
        eng.cf_run(sysd, p_b64, warm)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    outs64, st64 = eng.cf_run(sysd, p_b64, timed)
    torch.cuda.synchronize()
    t_b64 = time.perf_counter() - t0
    b64_cf0 = outs64[0].cpu().numpy()
    b64_cf1 = outs64[1].cpu().numpy() if K > 1 else None
    del outs64
    log(f"batched fp64: {t_b64 * 1e3:.1f} ms")

    # batched mixed: the headline
    if warm:
        eng.cf_run(sysd, p_mixed, warm)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    outs, stm = eng.cf_run(sysd, p_mixed, timed)
    torch.cuda.synchronize()
    t_mixed = time.perf_counter() - t0
    mixed_cf0 = outs[0].cpu().numpy()
    mixed_cf1 = outs[1].cpu().numpy() if K > 1 else None
    mixed_items0 = eng.items(0, n, L, rows=True)
    mixed_len = [mixed_items0["len"]] + [eng.items(j, n, L)["len"] for j in range(1, K)]
    # per-trajectory rows of the plain engine for the first timed step (outside the timings)
    out, st_plain0 = eng.cf_run(sysd, p_plain, [timed[0]])
    plain_cf0 = out[0].cpu().numpy()
    plain_items0 = eng.items(0, n, L, rows=True)
    log(f"batched mixed: {t_mixed * 1e3:.1f} ms for {K} steps -> {K * n / t_mixed:.0f} traj/s; kernel "
        f"{stm.integrate_ms:.1f} ms in {stm.launches} slices, reduce {stm.reduce_ms:.2f} ms, nrhs {stm.nrhs}, "
        f"{stm.n_orbit} trajectories handed over to FP64 ({stm.nrhs_orbit} right-hand sides)")

    value = K * n / t_mixed
    plain_rate = K * n / t_plain

    # the same engine with a batch `--deep` times deeper (the timed steps repeated): what it
    # sustains when each step's tail of long trajectories runs beside later steps' bulk
    deep_rate = None
    if args.deep > 1:
        deep = timed * args.deep
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        _, std = eng.cf_run(sysd, p_mixed, deep)
        torch.cuda.synchronize()
        t_deep = time.perf_counter() - t0
        deep_rate = len(deep) * n / t_deep
        log(f"batched mixed, {len(deep)} steps in one batch: {t_deep * 1e3:.1f} ms -> {deep_rate:.0f} traj/s")

    # roofline of the dominant kernel, the mixed-precision pool kernel (vector ALU: no matrix
    # work).  Algorithmic work counts right-hand sides and dipole evaluations only (rhs, dip of
    # flops.json): a right-hand side of the mixed kernel is rhs_f64 FP64 operations (Hamilton's
    # equations) and rhs - rhs_f64 FP32 ones (the surface gradient), a dipole dip_f64 FP64 (the
    # mu(0).mu(t) product) and the rest FP32.  Each precision is priced at its own vector peak:
    # peak = the rate of this operation mix with both at their peaks.  The kernel's launches and
    # their durations are the engine's own HIP events on the pool's stream (pool_launches,
    # pool_kernel_ms); the FP64 orbit runs are the other kernel of the same slices, reported
    # beside it over the slices' wall time.
    fl = json.load(open(os.path.join(HERE, "hawaiihybrid_amd", "data", "flops.json")))["co2ar"]
    nr_m, ns_m = stm.nrhs - stm.nrhs_orbit, stm.nsamples - stm.nsamples_orbit
    f32 = nr_m * (fl["rhs"] - fl["rhs_f64"]) + ns_m * (fl["dip"] - fl["dip_f64"])
    f64 = nr_m * fl["rhs_f64"] + ns_m * fl["dip_f64"]
    f64_orbit = stm.nrhs_orbit * fl["rhs"] + stm.nsamples_orbit * fl["dip"]
    nlaunch = max(stm.pool_launches, 1)
    launch_ms = stm.pool_kernel_ms / nlaunch
    secs = stm.pool_kernel_ms * 1e-3
    t_at_peak = f32 / (MI355X_FP32_VECTOR_TFLOPS * 1e12) + f64 / (MI355X_FP64_VECTOR_TFLOPS * 1e12)
    achieved = (f32 + f64) / secs / 1e12
    peak = (f32 + f64) / t_at_peak / 1e12
    traffic, traffic_src = None, None
    tpath = os.path.join(HERE, "profiles", "pool_cf_kernel_co2ar_mixed_traffic.json")
    if os.path.exists(tpath):
        traffic = json.load(open(tpath)).get("hbm_bytes_per_launch")
        traffic_src = ("profiles/pool_cf_kernel_co2ar_mixed_traffic.json (rocprofv3 --pmc FETCH_SIZE, WRITE_SIZE "
                       "of this workload, scaled by a 1 GiB copy; bench.py --traffic-from)")
    roofline = {"bound": "valu", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                "frac": achieved / peak, "traffic": traffic, "traffic_source": traffic_src,
                "kernel": "pool_cf_kernel<CO2-Ar, float, 256 lanes, 1 wavefront/SIMD> (the mixed-precision pool)",
                "pipe": "vector ALU, FP32 and FP64 (no matrix work)",
                "achieved_fp32_tflops": f32 / secs / 1e12, "peak_fp32_tflops": MI355X_FP32_VECTOR_TFLOPS,
                "achieved_fp64_tflops": f64 / secs / 1e12, "peak_fp64_tflops": MI355X_FP64_VECTOR_TFLOPS,
                "launches": stm.pool_launches, "avg_launch_ms": launch_ms,
                "algorithmic_flop_per_launch": (f32 + f64) / nlaunch,
                "slices": {"launches_both_pools": stm.launches, "wall_ms": stm.integrate_ms,
                           "fp64_orbit_flop": f64_orbit,
                           "achieved_tflops_both_pools": (f32 + f64 + f64_orbit) / (stm.integrate_ms * 1e-3) / 1e12},
                "flop_model": "hawaiihybrid_amd/data/flops.json: right-hand sides x rhs + samples x dip, "
                              "split into FP32 and FP64 parts"}

    result = {
        "metric": "CO2-Ar traj/s batched vs plain GPU at 1e-4 CF rel err; p50/p99 request latency",
        "value": value, "unit": "trajectories/s", "n_gpus": 1, "steps": K, "warmup": W,
        "ms_per_step": t_mixed / K * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64 state and error control, f32 surfaces",
        "data": "synthetic: seeded initial conditions from the reference Mersenne Twister stream",
        "config": {"workload": f"CO2-Ar dipole correlation function, {n} trajectories per step, T=200 K, "
                               f"R0=Rcut=40 bohr incoming, L={L}, dt=200 a.u.",
                   "ntraj_per_step": n, "L": L, "rtol_batched_mixed": args.rtol_mixed, "rtol_plain": 1e-12},
        "plain_fp64_traj_per_s": plain_rate, "speedup_vs_plain": value / plain_rate,
        "batched_mixed_traj_per_s_deep": deep_rate, "deep_steps": K * args.deep if deep_rate else 0,
        "plain_definition": "FP64 rtol=atol=1e-12, one thread per trajectory, one launch per step",
        "plain_fp64_one_launch_traj_per_s": K * n / t_plain_one,
        "speedup_vs_plain_one_launch": value / (K * n / t_plain_one),
        "batched_fp64_traj_per_s": K * n / t_b64,
        "batched_mixed_orbit_handover": {"trajectories": stm.n_orbit, "of": K * n,
                                          "fp64_rhs_fraction": stm.nrhs_orbit / max(stm.nrhs, 1),
                                          "rtol_fp64": 1e-12},
        "plain_fp64_kernel_ms_per_step": float(np.mean(plain_k_ms)),
        "roofline": roofline,
    }

    if not args.no_hbm:
        result["hbm"] = hbm_section(H, eng, stm, mixed_len, plain_items0["len"], st_plain0.reduce_ms, L, 1024,
                                    1 << 16, 1 << 22)
        log("hbm", result["hbm"])

    if not args.no_serve:
        # a fixed offered load below the mix's capacity, so latencies compare across engine
        # versions: at 60 requests/s the M0 requests alone (20% of them, 2e7 points on average
        # for 10^U[6, 8]) ask for more points per second than the rejection sampler delivers
        # (1.8e8/s), and the backlog sets p99
        rate = args.serve_rate
        log(f"serving mix: Poisson arrivals at {rate:.1f} requests/s for {args.serve_seconds:.0f} s")
        sv = serving_mix(H, eng, rate, args.serve_seconds, 5)
        log("serving", sv)
        result["serving"] = sv

    if not args.no_n2ar:
        log(f"N2-Ar PIP-NN surfaces: {args.n2ar_points} phase points")
        result["n2ar_pipnn"] = n2ar_section(H, eng, args.n2ar_points, 16384 if not args.no_oracle else 64)
        log("n2ar", result["n2ar_pipnn"])

    if not args.no_ch4co2:
        log(f"CH4-CO2: {args.ch4co2_ntraj} trajectories, L={args.ch4co2_L}")
        result["ch4co2"] = ch4co2_section(H, eng, args.ch4co2_ntraj, args.ch4co2_L, 0 if args.no_oracle else 512)
        log("ch4co2", result["ch4co2"])

    if not args.no_oracle:
        import ctypes as C
        import oracle_lib as O
        cores = min(16, os.cpu_count() or 1)
        ic0 = ic_all[W * n:(W + 1) * n]
        t0 = time.time()
        ref = O.cf_rows(O.co2ar(), O.CFParams(1e-12, 1e-12, 200.0, L, 40.0, 1, 50000), ic0, cores)
        t_or = time.time() - t0
        cf_ref = O.ensemble_from_rows(ref["crln"])
        prow = plain_items0["rows"] * ALU3
        mrow = mixed_items0["rows"] * ALU3
        dn = np.maximum(np.linalg.norm(ref["crln"], axis=1), 1e-300)
        dp = np.linalg.norm(prow - ref["crln"], axis=1) / dn
        dm = np.linalg.norm(mrow - ref["crln"], axis=1) / dn
        # whole ensemble, every trajectory counted (the metric's 1e-4 bar); for information, the
        # trajectories on which two FP64 integrations (GPU plain, CPU oracle) already disagree
        # beyond 1e-6 (chaotic: their rows depend on the last bit of the arithmetic)
        # The floor that chaos sets for any two FP64 integrations of this ensemble: the oracle
        # again, from initial conditions one unit in the last place away (every component scaled
        # by 1 + 2^-52).  A pair bound in an orbiting complex for the whole record has a row of
        # 2-4 % of the ensemble CF's norm, and where two integrations of it part depends on the
        # last bits of the arithmetic.
        t0 = time.time()
        ref_1ulp = O.cf_rows(O.co2ar(), O.CFParams(1e-12, 1e-12, 200.0, L, 40.0, 1, 50000),
                             np.ascontiguousarray(ic0 * (1.0 + 2.0 ** -52)), cores)
        t_floor = time.time() - t0
        floor = rel(O.ensemble_from_rows(ref_1ulp["crln"]), cf_ref)
        result["accuracy_vs_fp64_oracle"] = {
            "step": "first timed step, whole ensemble", "oracle": "CPU DOP853 FP64 rtol=atol=1e-12 (oracle/)",
            "cf_rel_l2_batched_mixed": rel(mixed_cf0, cf_ref), "cf_rel_l2_batched_fp64": rel(b64_cf0, cf_ref),
            "cf_rel_l2_plain_fp64": rel(plain_cf0, cf_ref), "bar": 1e-4,
            "cf_rel_l2_oracle_vs_oracle_1ulp": floor,
            "cf_rel_l2_batched_mixed_vs_batched_fp64": rel(mixed_cf0, b64_cf0),
            "chaotic_trajectories_fp64": int((dp >= 1e-6).sum()),
            "per_traj_rel_l2_batched_mixed_median": float(np.median(dm)),
            "per_traj_rel_l2_batched_mixed_p99": float(np.quantile(dm, 0.99)),
            "max_energy_drift_hartree": float(stm.max_e_drift)}
        log("accuracy", result["accuracy_vs_fp64_oracle"],
            f"(oracle {t_or:.1f} s, 1-ulp oracle {t_floor:.1f} s on {cores} threads)")
        m = min(args.cpu_sample, K * n)
        cf = np.zeros(L)
        t0 = time.time()
        O.lib().orc_cf_ensemble(C.byref(O.co2ar()), C.byref(O.CFParams(1e-12, 1e-12, 200.0, L, 40.0, 1, 50000)),
                                O.ptr(np.ascontiguousarray(ic_all[W * n:W * n + m])), m, cores, 1.0, O.ptr(cf), None)
        t_cpu = time.time() - t0
        if K > 1 and m == 2 * n:
            # the same comparison over the first two timed steps (131072 trajectories of the
            # oracle's CPU-baseline run), each engine's CF the mean of its two per-step CFs
            result["accuracy_vs_fp64_oracle"]["first_two_steps"] = {
                "trajectories": m,
                "cf_rel_l2_batched_mixed": rel(0.5 * (mixed_cf0 + mixed_cf1), cf),
                "cf_rel_l2_batched_fp64": rel(0.5 * (b64_cf0 + b64_cf1), cf),
                "cf_rel_l2_batched_mixed_vs_batched_fp64": rel(0.5 * (mixed_cf0 + mixed_cf1),
                                                               0.5 * (b64_cf0 + b64_cf1))}
            log("accuracy over the first two steps", result["accuracy_vs_fp64_oracle"]["first_two_steps"])
        model = "unknown"
        try:
            with open("/proc/cpuinfo") as f:
                model = next(line.split(":", 1)[1].strip() for line in f if line.startswith("model name"))
        except (OSError, StopIteration):
            pass
        result["cpu_baseline"] = {"value": m / t_cpu, "unit": "trajectories/s", "cores": cores, "kind": "port",
                                  "cpu": model, "host_cpus_visible": os.cpu_count(),
                                  "sample": f"first {m} trajectories of the timed steps, CPU oracle "
                                            f"(DOP853 FP64, rtol=atol=1e-12), {cores} threads, {t_cpu:.1f} s"}
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
