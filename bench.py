#!/usr/bin/env python3
"""CO2-Ar trajectory-ensemble benchmark on one MI355X.

A step is one batch of `--ntraj` CO2-Ar trajectories (T = 200 K, R0 = Rcut = 40 bohr,
incoming momentum, L = 4096 samples at dt = 200 a.u.) taken from consecutive chunks of the
seed-42 Mersenne Twister stream: sample, integrate, record the dipole correlation function,
reduce to the ensemble CF.  Initial conditions are resident in HBM before timing.

Timed engines (same batches):
  plain    FP64, one trajectory per thread, one launch per batch
  batched  persistent kernel + device work queue, FP32 surfaces with FP64 state and error
           control (rtol = atol = --rtol-mixed), all timed batches streamed through one launch
The headline `value` is the batched engine's trajectories/s; `speedup_vs_plain` compares it
with the plain FP64 engine.  Accuracy of both against the CPU oracle (FP64) is reported for
the first timed batch.  One JSON line on stdout.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.join(HERE, "tests"))


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=4)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--ntraj", type=int, default=65536)
    ap.add_argument("--L", type=int, default=4096)
    ap.add_argument("--rtol-mixed", type=float, default=1e-10)
    ap.add_argument("--cpu-sample", type=int, default=16384)
    ap.add_argument("--no-oracle", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on exactly one GPU: the trajectory ensemble of this path is measured on "
              "one MI355X (multi-GPU sharding is a design note, not a measured configuration)", file=sys.stderr)
        sys.exit(2)

    import torch
    import hawaiihybrid_amd as H

    eng = H.Engine(0)
    info = eng.info()
    log("device", info)
    sysd = H.co2_ar()
    smp = H.Sampler(H.IC_R0_INJECT, H.PAIR_FREE, 200.0, 4.0, 40.0, H.PESMIN_CO2AR, 40.0)
    nb = args.warmup + args.steps
    n = args.ntraj
    t0 = time.time()
    ic_all, _, _ = H.sample_ics(sysd, smp, 42, nb * n)
    log(f"sampled {nb * n} initial conditions in {time.time() - t0:.2f} s (host, bit-exact stream)")
    ics = [eng.to_device_soa(ic_all[b * n:(b + 1) * n]) for b in range(nb)]
    warm, timed = ics[:args.warmup], ics[args.warmup:]
    torch.cuda.synchronize()

    p_plain = H.cf_params(args.L, rtol=1e-12, engine=H.ENGINE_PLAIN, precision=H.PREC_FP64)
    p_mixed = H.cf_params(args.L, rtol=args.rtol_mixed, engine=H.ENGINE_BATCHED, precision=H.PREC_MIXED)

    # plain FP64: one launch per batch
    for x in warm:
        eng.cf_run(sysd, p_plain, [x])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    plain_cf0 = None
    for i, x in enumerate(timed):
        out, st = eng.cf_run(sysd, p_plain, [x])
        if i == 0:
            plain_cf0 = out[0].cpu().numpy()
            plain_items0 = eng.items(0, n, args.L, rows=True)
    torch.cuda.synchronize()
    t_plain = time.perf_counter() - t0
    log(f"plain fp64: {t_plain * 1e3:.1f} ms for {args.steps} steps -> {args.steps * n / t_plain:.0f} traj/s")

    # batched mixed: all timed batches through one persistent launch
    if warm:
        eng.cf_run(sysd, p_mixed, warm)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    outs, stm = eng.cf_run(sysd, p_mixed, timed)
    torch.cuda.synchronize()
    t_mixed = time.perf_counter() - t0
    mixed_cf0 = outs[0].cpu().numpy()
    mixed_items0 = eng.items(0, n, args.L, rows=True)
    log(f"batched mixed: {t_mixed * 1e3:.1f} ms for {args.steps} steps -> {args.steps * n / t_mixed:.0f} traj/s; "
        f"kernel {stm.integrate_ms:.1f} ms, reduce {stm.reduce_ms:.2f} ms, nrhs {stm.nrhs}, grid {stm.grid}x{stm.block}")

    value = args.steps * n / t_mixed
    plain_rate = args.steps * n / t_plain

    # roofline of the persistent kernel (FP64 vector pipe: state, error control, EOM)
    fl = json.load(open(os.path.join(HERE, "hawaiihybrid_amd", "data", "flops.json")))["co2ar"]
    nsteps = max(stm.nrhs - 2 * stm.n_items, 0) / 13.0
    flops = stm.nrhs * fl["rhs"] + stm.nsamples * (fl["dip"] + fl["sample"]) + nsteps * fl["step"]
    achieved = flops / (stm.integrate_ms * 1e-3) / 1e12
    peak = 78.6
    traffic = None
    tpath = os.path.join(HERE, "profiles", "traffic_persistent_cf_co2ar.json")
    if os.path.exists(tpath):
        traffic = json.load(open(tpath)).get("bytes_per_launch")
    roofline = {"bound": "mfma", "pipe": "fp64 vector ALU (no matrix work in this kernel)", "achieved": achieved,
                "peak": peak, "unit": "TFLOP/s", "frac": achieved / peak, "traffic": traffic,
                "kernel": "persistent_cf_kernel<CO2-Ar, float>", "launch_ms": stm.integrate_ms,
                "algorithmic_flop": flops}

    result = {
        "metric": "CO2-Ar traj/s batched vs plain GPU at 1e-4 CF rel err; p50/p99 request latency",
        "value": value, "unit": "trajectories/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": t_mixed / args.steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64 state + f32 surfaces", "data": "synthetic (seeded initial conditions)",
        "config": {"workload": f"CO2-Ar dipole correlation function, {n} trajectories per step, T=200 K, "
                               f"R0=Rcut=40 bohr (incoming), L={args.L}, dt=200 a.u.",
                   "ntraj_per_step": n, "L": args.L, "rtol_mixed": args.rtol_mixed, "rtol_plain": 1e-12},
        "plain_fp64_traj_per_s": plain_rate, "speedup_vs_plain": value / plain_rate,
        "roofline": roofline,
    }

    if not args.no_oracle:
        import oracle_lib as O
        cores = min(16, os.cpu_count() or 1)
        ic0 = ic_all[args.warmup * n:(args.warmup + 1) * n]
        t0 = time.time()
        ref = O.cf_rows(O.co2ar(), O.CFParams(1e-12, 1e-12, 200.0, args.L, 40.0, 1, 50000), ic0, cores)
        t_or = time.time() - t0
        cf_ref = O.ensemble_from_rows(ref["crln"])
        alu3 = 5.29177210903e-11 ** 3

        def rel(a, b):
            return float(np.linalg.norm(a - b) / np.linalg.norm(b))
        # trajectories where two FP64 integrations (GPU plain vs CPU oracle) disagree are chaotic
        d = np.linalg.norm(plain_items0["rows"] * alu3 - ref["crln"], axis=1) / np.maximum(
            np.linalg.norm(ref["crln"], axis=1), 1e-300)
        regular = d < 1e-6
        mrows = mixed_items0["rows"] * alu3
        result["accuracy_vs_fp64_oracle"] = {
            "batch": "first timed batch", "cf_rel_l2_plain_fp64": rel(plain_cf0, cf_ref),
            "cf_rel_l2_batched_mixed": rel(mixed_cf0, cf_ref),
            "chaotic_trajectories": int((~regular).sum()),
            "cf_rel_l2_batched_mixed_regular": rel(mrows[regular].sum(0), ref["crln"][regular].sum(0)),
            "cf_rel_l2_plain_fp64_regular": rel((plain_items0["rows"] * alu3)[regular].sum(0),
                                                ref["crln"][regular].sum(0)),
        }
        log("accuracy", result["accuracy_vs_fp64_oracle"], f"(oracle {t_or:.1f} s on {cores} threads)")
        # CPU baseline: the oracle on a bounded prefix of the batch
        m = min(args.cpu_sample, n)
        cf = np.zeros(args.L)
        t0 = time.time()
        O.lib().orc_cf_ensemble(C.byref(O.co2ar()), C.byref(O.CFParams(1e-12, 1e-12, 200.0, args.L, 40.0, 1, 50000)),
                                O.ptr(np.ascontiguousarray(ic0[:m])), m, cores, 1.0, O.ptr(cf), None)
        t_cpu = time.time() - t0
        result["cpu_baseline"] = {"value": m / t_cpu, "unit": "trajectories/s", "cores": cores, "kind": "port",
                                  "sample": f"first {m} trajectories of the first timed batch, FP64 oracle "
                                            f"(same DOP853 at rtol=atol=1e-12), {cores} threads, {t_cpu:.1f} s"}
    print(json.dumps(result))


if __name__ == "__main__":
    main()
