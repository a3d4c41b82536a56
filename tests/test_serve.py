"""CPU checks of the serving loop: the job distribution, the systems it generates, the scheduling and
latency accounting of the worker pool, and the admission, stepping and retirement of the continuous-batching
server (with stand-ins for the engine, so no GPU is needed)."""
import math
import threading
import time

import numpy as np
import pytest

from autopas_amd import serve


def test_job_distribution():
    jobs = serve.make_jobs(4000, rate=100.0, seed=7)
    n = np.array([j.n for j in jobs])
    assert n.min() >= 256 and n.max() <= 65536
    # log-uniform: the median sits near the geometric mean of the range (4096)
    assert 3000 < np.median(n) < 5600
    rho = np.array([j.density for j in jobs])
    assert rho.min() >= 0.4 and rho.max() <= 0.85
    steps = np.array([j.steps for j in jobs])
    assert steps.min() >= 10 and steps.max() <= 1000
    assert [j.seed for j in jobs[:3]] == [5000, 5001, 5002]
    arr = np.array([j.arrival for j in jobs])
    assert np.all(np.diff(arr) > 0)
    gaps = np.diff(np.concatenate([[0.0], arr]))
    assert abs(gaps.mean() - 0.01) < 0.001  # Poisson process at 100 jobs/s
    # reproducible, and arrivals all at zero without a rate
    again = serve.make_jobs(4000, rate=100.0, seed=7)
    assert [(j.n, j.steps, j.arrival) for j in jobs] == [(j.n, j.steps, j.arrival) for j in again]
    assert all(j.arrival == 0.0 for j in serve.make_jobs(10))


@pytest.mark.parametrize("n,density", [(256, 0.85), (1000, 0.4), (4097, 0.6)])
def test_job_system(n, density):
    job = serve.Job(index=0, n=n, density=density, steps=10, temperature=1.0, seed=5000)
    L, (x, y, z, vx, vy, vz) = serve.job_system(job)
    assert len(x) == n
    assert L == pytest.approx((n / density) ** (1 / 3))
    for a in (x, y, z):
        assert a.min() > -0.06 and a.max() < L + 0.06
    assert L > 2 * (serve.RC + serve.SKIN)  # every serving job fits the fixed-point kernels
    assert abs(vx.sum()) < 1e-9


class _FakeEngine:
    def close(self):
        pass


def test_server_schedules_and_accounts_latency():
    jobs = serve.make_jobs(24, rate=400.0, seed=3)
    busy = []
    lock = threading.Lock()
    active = [0]

    def run(eng, job, kernel):
        with lock:
            active[0] += 1
            busy.append(active[0])
        time.sleep(0.002 + job.steps * 1e-5)
        with lock:
            active[0] -= 1
        return job.n * 10, 1, {}

    with serve.Server(workers=3, make_engine=_FakeEngine) as srv:
        wall = srv.run(jobs, run=run)
    assert max(busy) <= 3 and max(busy) >= 2
    for j in jobs:
        assert j.start >= j.arrival - 1e-3
        assert j.finish > j.start
        assert j.latency >= j.finish - j.start - 1e-9
        assert j.pairs == j.n * 10
    s = serve.latency_summary(jobs, wall)
    assert s["p50_ms"] <= s["p99_ms"]
    assert s["pairs"] == sum(j.n for j in jobs) * 10
    assert wall >= max(j.arrival for j in jobs)


def test_server_reports_failures():
    jobs = serve.make_jobs(4)

    def run(eng, job, kernel):
        if job.index == 2:
            raise ValueError("boom")
        return 0, 0, {}

    with serve.Server(workers=2, make_engine=_FakeEngine) as srv:
        with pytest.raises(RuntimeError, match="job 2 failed"):
            srv.run(jobs, run=run)


class _FakeEnsemble:
    """Stands in for an ensemble context: tracks the box list through edits (each box tagged by the first x
    of its particles, which identifies the job), the steps every box has run, and the particles in flight."""

    def __init__(self, *args, **kwargs):
        self.boxes = []  # [tag, count, steps run]
        self.peak = 0
        self.most = 0
        self.kwargs = kwargs
        self.clock = 0  # steps run since the box list was last empty
        self.admit_clocks = []  # the clock at each edit that added boxes to a running batch
        self.layouts = []  # (particles held, cluster size, tile groups) at each set_layout

    def memory_info(self):
        return 1 << 40, 1 << 41

    def set_layout(self, cluster_size, tile_groups=0, build_groups=0):
        self.layouts.append((sum(b[1] for b in self.boxes), cluster_size, tile_groups))

    def edit_boxes(self, keep, box_min=None, box_max=None, counts=None, x=None, **parts):
        assert len(set(keep)) == len(keep) and all(0 <= k < len(self.boxes) for k in keep)
        kept = [self.boxes[k] for k in keep]
        if not kept:
            self.clock = 0
        elif counts is not None:
            self.admit_clocks.append(self.clock)
        added = []
        if counts is not None:
            ends = np.cumsum(counts)
            for n, e in zip(counts, ends):
                added.append([float(x[e - n]), int(n), 0])
        self.boxes = kept + added
        self.peak = max(self.peak, sum(b[1] for b in self.boxes))
        self.most = max(self.most, len(self.boxes))

    def step(self, steps, dt, kernel):
        assert steps >= 1
        time.sleep(1e-4 * steps)  # a step takes time, so batches overlap the arrivals
        self.clock += steps
        for b in self.boxes:
            b[2] += steps

        class R:
            pair_interactions = 0
            rebuilds = 0
        return R()

    def download_range(self, first, count):
        offsets = np.cumsum([0] + [b[1] for b in self.boxes])
        k = int(np.searchsorted(offsets, first))
        assert offsets[k] == first and self.boxes[k][1] == count
        return dict(tag=self.boxes[k][0], steps=self.boxes[k][2])

    def close(self):
        pass


@pytest.mark.parametrize("max_jobs,aligned,lanes,tiles", [(None, True, 1, 1), (3, True, 1, 1), (None, False, 1, 1),
                                                          (None, True, 3, 1), (None, True, 2, "size")])
def test_batched_server_admits_steps_and_retires(monkeypatch, max_jobs, aligned, lanes, tiles):
    # jobs arrive over ~20 ms; the budget holds a few at once; every job must run exactly its own steps,
    # start no earlier than it arrived, and be read back from its own place in its batch; aligned, jobs join
    # a running batch only at multiples of the rebuild frequency of its steps; with several lanes the jobs
    # spread over the lanes' contexts, each within its share of the budget
    monkeypatch.setattr(serve.engine, "Engine", _FakeEnsemble)
    jobs = serve.make_jobs(16, rate=800.0, seed=9, n_range=(256, 4000), steps_range=(3, 40))
    budget = 6000
    with serve.BatchedServer(particle_budget=budget, keep_results=True, epoch=10, max_jobs=max_jobs,
                             shared_rebuilds=True, aligned=aligned, rebuild_frequency=10, lanes=lanes,
                             tile_groups=tiles) as srv:
        engines = [lane.engine for lane in srv.lanes]
        assert len(engines) == lanes and srv.engine is engines[0]
        assert all(e.kwargs["shared_rebuilds"] is True and e.kwargs["rebuild_frequency"] == 10 for e in engines)
        srv.run(jobs)
        assert srv.edits >= 3 and srv.epochs >= 4
    share = budget // lanes
    for e in engines:
        assert e.peak <= max(share, max(j.n for j in jobs))
        if max_jobs is not None:
            assert e.most <= max_jobs
    # "size": a lane's tile shape follows the particles it holds (job_layout's rule), set only on a change
    layouts = [lay for e in engines for lay in e.layouts]
    if tiles == "size":
        assert all(g == serve.job_layout(n)[2] and m == 8 for n, m, g in layouts), layouts
    else:
        assert not layouts
    clocks = [c for e in engines for c in e.admit_clocks]
    if aligned:
        assert all(c % 10 == 0 for c in clocks), clocks
    else:
        assert any(c % 10 != 0 for c in clocks), clocks
    assert {j.worker for j in jobs} == set(range(lanes))
    for j in jobs:
        tag = float(serve.job_system(j)[1][0][0])
        assert j.result["tag"] == tag, j.index
        assert j.result["steps"] == j.steps, j.index
        assert j.start >= j.arrival - 1e-9 and j.finish >= j.start
