"""The lane pool's work-queue head protocol, as a host-side model (no GPU).

A lane pool persists across hh_cf_run calls and session admissions.  Its device head `next` is
advanced by one atomic fetch-add per lane that asks for an item (pool_cf_kernel's refill); a lane
that draws an index >= total stays idle, so the head ends a slice up to one fetch per idle lane
past the end.  Jobs are appended at `total` (pool_admit, hawaiihybrid_amd/csrc/hh_pool.cpp).  The
protocol is correct when every item index in [0, total) is handed out exactly once:

* pool_admit clamps the device head back to the current `total` on the stream before extending it,
  so the overshoot of earlier slices never skips the new job's first items;
* a graph-captured slice also clamps the head at its start, with the `total` the graph was captured
  for (graphs are cached per admission state);
* the host's copy of the head is clamped to `total` when it is read.

The model replays that protocol over random slices, in both launch modes, across two runs of the
pool and across admissions between slices, and checks the head against `total` at every
admission.  Leaving out the admission clamp (the protocol with only the graph's clamp) loses items,
which the last test shows.
"""
import random

import pytest


class PoolModel:
    def __init__(self, lanes, graphs, admission_clamp=True):
        self.lanes = lanes
        self.graphs = graphs
        self.admission_clamp = admission_clamp
        self.next = 0          # device head
        self.total = 0         # host: items ever admitted
        self.handed = []       # item indices handed out, in order
        self.heads_at_admission = []

    def admit(self, n):
        if self.admission_clamp:
            self.next = min(self.next, self.total)    # run_clamp_next(next, total) on the stream
        self.heads_at_admission.append((self.next, self.total))
        self.total += n

    def slice(self, rng):
        if self.graphs:
            self.next = min(self.next, self.total)    # the graph's own clamp, captured at this total
        # idle lanes ask for items until the slice ends: each ask is one fetch-add
        for _ in range(rng.randrange(0, 3 * self.lanes)):
            i = self.next
            self.next += 1
            if i < self.total:
                self.handed.append(i)

    def drained(self):
        return min(self.next, self.total) >= self.total


def run_calls(model, rng, calls, mid_admissions):
    """`calls` runs of the pool, each admitting its jobs then slicing until the queue drains; with
    mid_admissions, further jobs arrive between slices (the serving session's admissions)."""
    for jobs in calls:
        for n in jobs:
            model.admit(n)
        extra = list(mid_admissions)
        slices = 0
        while not model.drained() or extra:
            model.slice(rng)
            slices += 1
            if extra and rng.random() < 0.5:
                model.admit(extra.pop())
            assert slices < 100000


@pytest.mark.parametrize("graphs", [False, True])
@pytest.mark.parametrize("seed", range(5))
def test_every_item_handed_out_once_across_runs_and_admissions(graphs, seed):
    rng = random.Random(seed)
    m = PoolModel(lanes=64, graphs=graphs)
    calls = [[rng.randrange(1, 500) for _ in range(rng.randrange(1, 4))] for _ in range(3)]
    run_calls(m, rng, calls, mid_admissions=[rng.randrange(1, 300) for _ in range(3)])
    assert sorted(m.handed) == list(range(m.total))
    # at every admission the device head is at most the end of the items admitted so far
    assert all(head <= total for head, total in m.heads_at_admission)


def test_without_the_admission_clamp_a_second_run_loses_items():
    """The protocol with only the graph's clamp (captured for the post-admission total): the first
    run's overshoot carries into the second, whose first items are never handed out, so its job
    never completes."""
    rng = random.Random(1)
    m = PoolModel(lanes=64, graphs=True, admission_clamp=False)
    m.admit(100)
    while not m.drained():
        m.slice(rng)
    overshoot = m.next - m.total
    assert overshoot > 0
    m.admit(100)
    for _ in range(50):
        m.slice(rng)
    lost = set(range(m.total)) - set(m.handed)
    assert lost == set(range(100, 100 + overshoot))
