"""Graph-captured slices of the batched engine.

With HH_GRAPHS=1 the slices of a lane pool run as captured graphs (repack when the queue has
drained, the pool kernel, copies of the counters into pinned memory), replayed many times per
run with the host's reductions and bookkeeping between replays.  Every replay must give the same bits as the
directly launched slices (an engine created with HH_GRAPHS=0), in FP64 and in mixed
precision.  The second and later runs on one engine reuse its lane pool, whose device queue head
the first run left past the end (pool_admit clamps it back before the new jobs).
"""
import numpy as np
import pytest

import hawaiihybrid_amd as H

pytestmark = pytest.mark.gpu


def engine(monkeypatch, graphs):
    """An engine with graph-captured (the default) or directly launched slices (read at
    creation)."""
    monkeypatch.setenv("HH_GRAPHS", "1" if graphs else "0")
    e = H.Engine(0)
    monkeypatch.delenv("HH_GRAPHS")
    return e


def run(eng, ic_dev, prec, L=512):
    p = H.cf_params(L, rtol=1e-10, engine=H.ENGINE_BATCHED, precision=prec)
    outs, st = eng.cf_run(H.co2_ar(), p, ic_dev)
    return [o.cpu().numpy() for o in outs], st


def test_graph_replays_equal_direct_launches(monkeypatch):
    smp = H.Sampler(H.IC_R0_INJECT, H.PAIR_FREE, 200.0, 4.0, 40.0, H.PESMIN_CO2AR, 40.0)
    ic, _, _ = H.sample_ics(H.co2_ar(), smp, 42, 3 * 8192)
    print("graphs: 3 jobs of 8192 CO2-Ar trajectories, L=512", flush=True)
    g = engine(monkeypatch, True)
    d = engine(monkeypatch, False)
    xs = [g.to_device_soa(ic[j * 8192:(j + 1) * 8192]) for j in range(3)]
    for prec in (H.PREC_FP64, H.PREC_MIXED):
        ref, st = run(d, xs, prec)
        for k in range(3):   # three runs on the graph engine's one pool
            out, st = run(g, xs, prec)
            print(f"  precision {prec} run {k}: {st.launches} slices", flush=True)
            for a, b in zip(ref, out):
                assert np.array_equal(a, b)
    print("  graph replays and direct launches agree bit for bit", flush=True)
    g.close()
    d.close()
