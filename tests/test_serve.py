"""CPU checks of the serving loop: the job distribution, the systems it generates, and the scheduling and
latency accounting of the worker pool (with a stand-in for the engine, so no GPU is needed)."""
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
