"""GPU checks of job serving: contexts reused across boxes (ljg_set_box) and concurrent workers give the
same trajectories as one fresh context per job."""
import numpy as np
import pytest

import autopas_amd as ap
from autopas_amd import serve
from conftest import announce

pytestmark = pytest.mark.gpu


def fresh_run(job, kernel):
    L, (x, y, z, vx, vy, vz) = serve.job_system(job)
    with ap.Engine((0.0, 0.0, 0.0), (L, L, L), serve.RC, serve.SKIN, cluster_size=4, rebuild_frequency=10,
                   dynamic_rebuild=True) as e:
        e.upload(x, y, z, vx, vy, vz)
        rs = e.step(job.steps, serve.DT, kernel)
        return rs.pair_interactions, e.download()


@pytest.mark.parametrize("kernel", [ap.KERNEL_CLUSTER_MIXED, ap.KERNEL_CLUSTER_JROWS_MIXED])
def test_served_jobs_match_fresh_contexts(kernel):
    announce(f"serving: 6 jobs on 2 workers vs fresh contexts, kernel={ap.KERNEL_NAMES[kernel]}")
    jobs = serve.make_jobs(6, rate=None, seed=11, n_range=(256, 3000), steps_range=(10, 40))
    with serve.Server(workers=2, kernel=kernel, keep_results=True) as srv:
        srv.run(jobs)
    for job in jobs:
        pairs, ref = fresh_run(job, kernel)
        assert job.pairs == pairs, job.index
        for k in ("x", "y", "z", "vx", "vy", "vz", "fx", "fy", "fz"):
            assert np.array_equal(job.result[k], ref[k]), (job.index, k)


def test_set_box_rejects_small_box():
    announce("ljg_set_box rejects a box shorter than rc + skin")
    with ap.Engine((0.0, 0.0, 0.0), (10.0, 10.0, 10.0), 2.5, 0.3) as e:
        with pytest.raises(ap.EngineError) as ei:
            e.set_box((0.0, 0.0, 0.0), (2.0, 10.0, 10.0))
        assert ei.value.status == 1
