"""Job serving on one GPU: independent MD jobs (small periodic LJ boxes) arrive over time and are run by
a pool of workers, each owning one engine context (its own HIP stream) that is reused from job to job.

The unit of work is md-flexible's whole simulation loop for one box (``Simulation.cpp:230-350``): the
particles are uploaded, ``steps`` velocity-Verlet steps run on the device with lists rebuilt every
``rebuild_frequency`` steps or on a skin/2 displacement (``LogicHandler.h:987-1015``), and the final
state is downloaded.  Latency is measured from a job's arrival to the end of its download.

Job parameters follow the serving workload: N log-uniform in [256, 65536], density uniform in
[0.4, 0.85], steps uniform in {10, ..., 1000}, temperature uniform in [0.7, 1.5], a jittered simple-cubic
lattice (jitter +-0.05 sigma) with ceil(N^(1/3)) sites per side (the first N sites) filling the cube
L = (N / rho)^(1/3), and a per-job seed 5000 + j.  Arrivals are a Poisson process.  All draws come from
numpy's Philox4x32 generator.

Worker threads spend their time inside the C ABI, which releases the GIL, so the contexts' streams run
concurrently on the device.
"""
import collections
import math
import queue
import threading
import time
from dataclasses import dataclass, field

import numpy as np

from . import engine

RC = 2.5
SKIN = 0.3
DT = 0.002


@dataclass
class Job:
    index: int
    n: int
    density: float
    steps: int
    temperature: float
    seed: int
    arrival: float = 0.0  # seconds after the start of the run
    # filled in by the server
    start: float = math.nan
    finish: float = math.nan
    pairs: int = 0
    rebuilds: int = 0
    worker: int = -1
    result: dict = field(default_factory=dict)

    @property
    def box_length(self):
        return (self.n / self.density) ** (1.0 / 3.0)

    @property
    def latency(self):
        return self.finish - self.arrival


def make_jobs(count, rate=None, seed=5, n_range=(256, 65536), density_range=(0.4, 0.85), steps_range=(10, 1000),
              temperature_range=(0.7, 1.5), first_seed=5000):
    """Draw ``count`` jobs; with ``rate`` (jobs/s) their arrival times form a Poisson process, else all
    arrive at t = 0 (closed loop)."""
    rng = np.random.Generator(np.random.Philox(seed))
    lo, hi = math.log(n_range[0]), math.log(n_range[1])
    jobs = []
    t = 0.0
    for j in range(count):
        n = int(round(math.exp(rng.uniform(lo, hi))))
        n = min(max(n, n_range[0]), n_range[1])
        job = Job(index=j, n=n, density=float(rng.uniform(*density_range)),
                  steps=int(rng.integers(steps_range[0], steps_range[1] + 1)),
                  temperature=float(rng.uniform(*temperature_range)), seed=first_seed + j)
        if rate:
            t += float(rng.exponential(1.0 / rate))
            job.arrival = t
        jobs.append(job)
    return jobs


def job_system(job):
    """Jittered simple-cubic lattice of the job (positions, velocities) in the cube [0, L)^3."""
    k = int(math.ceil(job.n ** (1.0 / 3.0) - 1e-9))
    L = job.box_length
    a = L / k
    x, y, z, vx, vy, vz, _ = engine.generate_lattice(job.seed, job.n, (k, k, k), a, (0.0, 0.0, 0.0), 0.05,
                                                     job.temperature)
    return L, (x, y, z, vx, vy, vz)


KERNEL_BY_SIZE = None  # run_job's kernel argument: pick the layout from the job's size (job_layout)


def job_layout(n):
    """Layout for a job of n particles: (kernel, cluster size, i-clusters per wavefront).  Serving-sized boxes
    (up to 65,536 particles) are bound by launch and rebuild latency, where the cluster tiles (whose build
    has no expansion pass) with few i-clusters per wavefront were fastest on MI355X at every size from 512
    to 64,000 particles: one i-cluster per wavefront up to 2048 clusters, four above (measured at density
    0.75, us per step for 1 / 2 / 4 / 8 i-clusters per wavefront: 256 clusters 19 / 22 / 22 / 27,
    2048 clusters 25 / 27 / 27 / 31, 4096 clusters 30 / 31 / 27 / 32, 8192 clusters 39 / 37 / 34 / 36)."""
    clusters = (n + 7) // 8
    return engine.KERNEL_CLUSTER_MIXED, 8, (1 if clusters <= 2048 else 4)


def run_job(eng, job, kernel):
    """Run one job on an engine context (box, upload, steps, download); returns (pairs, rebuilds, state).
    kernel KERNEL_BY_SIZE picks the layout from the job's size (job_layout)."""
    L, (x, y, z, vx, vy, vz) = job_system(job)
    eng.set_box((0.0, 0.0, 0.0), (L, L, L))
    if kernel is KERNEL_BY_SIZE:
        kernel, m, g = job_layout(job.n)
        eng.set_layout(m, g)
    eng.upload(x, y, z, vx, vy, vz)
    rs = eng.step(job.steps, DT, kernel)
    return rs.pair_interactions, rs.rebuilds, eng.download()


class Server:
    """A pool of ``workers`` threads, each with one engine context, serving jobs first come first served."""

    def __init__(self, workers=4, kernel=engine.KERNEL_CLUSTER_MIXED, cluster_size=4, tile_groups=0,
                 keep_results=False, make_engine=None):
        self.workers = workers
        self.kernel = kernel
        self.keep_results = keep_results
        if make_engine is None:
            def make_engine():
                return engine.Engine((0.0, 0.0, 0.0), (8.0, 8.0, 8.0), RC, SKIN, cluster_size=cluster_size,
                                     rebuild_frequency=10, dynamic_rebuild=True, tile_groups=tile_groups)
        self._make_engine = make_engine
        self._engines = [make_engine() for _ in range(workers)]
        self.errors = []

    def close(self):
        for e in self._engines:
            e.close()
        self._engines = []

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def run(self, jobs, run=run_job):
        """Serve ``jobs`` (arrival times relative to now); returns the wall time from start to the last
        completion.  Each job gets start/finish/latency, pairs and rebuilds filled in."""
        q = queue.Queue()
        self.errors = []
        t0 = time.perf_counter()

        def feeder():
            for job in sorted(jobs, key=lambda j: j.arrival):
                delay = job.arrival - (time.perf_counter() - t0)
                if delay > 0:
                    time.sleep(delay)
                q.put(job)
            for _ in range(self.workers):
                q.put(None)

        def worker(w):
            eng = self._engines[w]
            while True:
                job = q.get()
                if job is None:
                    return
                job.worker = w
                job.start = time.perf_counter() - t0
                try:
                    pairs, rebuilds, state = run(eng, job, self.kernel)
                except Exception as exc:  # recorded and re-raised by run()
                    self.errors.append((job.index, exc))
                    pairs, rebuilds, state = 0, 0, {}
                job.finish = time.perf_counter() - t0
                job.pairs, job.rebuilds = pairs, rebuilds
                if self.keep_results:
                    job.result = state

        threads = [threading.Thread(target=worker, args=(w,), daemon=True) for w in range(self.workers)]
        for th in threads:
            th.start()
        feeder()
        for th in threads:
            th.join()
        if self.errors:
            idx, exc = self.errors[0]
            raise RuntimeError(f"job {idx} failed: {exc}") from exc
        return max(j.finish for j in jobs) if jobs else 0.0


class _Lane:
    """One batching context of a BatchedServer: its engine (an ensemble context with its own stream), the
    jobs it runs as boxes, the jobs dispatched to it and not yet admitted, and its phase."""

    def __init__(self, eng):
        self.engine = eng
        self.active = []  # [job, steps left], in box order
        self.inbox = collections.deque()  # dispatched jobs, in arrival order
        self.phase = 0  # batch steps since the last phase boundary
        self.groups = 1  # i-clusters per wavefront of its force launches

    def work_left(self):
        """Particle-steps still to run for the jobs running and dispatched here."""
        return sum(j.n * left for j, left in list(self.active)) + sum(j.n * j.steps for j in list(self.inbox))

    def particles(self):
        """Particles of the jobs running and dispatched here (a batch's step time grows with them)."""
        return sum(j.n for j, left in list(self.active) if left > 0) + sum(j.n for j in list(self.inbox))


class BatchedServer:
    """Continuous batching: every running job is a box of an ensemble context, so each step advances all
    running jobs of the context in the same launches.  Between epochs of at most ``epoch`` steps the
    finished jobs are downloaded and retired and the waiting jobs are admitted, first come first served,
    while the particles in flight stay within ``particle_budget`` (the HBM budget) and the batch holds at
    most ``max_jobs`` jobs (fewer jobs in a batch finish sooner each); the running jobs' state stays resident
    on the device (``Engine.edit_boxes``).  An epoch ends early when a running job reaches its step count, so
    every job runs exactly its own number of steps.

    ``lanes`` ensemble contexts batch at once, each on its own thread and HIP stream, so the batches' launches
    overlap on the device; an arriving job is dispatched to the lane with the least work left (``dispatch``
    "work": particle-steps of its running and waiting jobs) or the fewest particles ("particles"), and each
    lane holds a share of the particle budget.  One lane is a
    single batch.  A batch of serving jobs holds a few thousand to a few tens of thousands of clusters, where
    the force launch is bound by the latency of one wavefront's list: one i-cluster per wavefront
    (``tile_groups`` 1) spreads it over the most lanes (measured on MI355X: a 1-box batch runs at the
    speed of the box's own context with it, 10-60% slower with 4).

    Each job keeps the rebuild schedule it has alone (its own frequency and displacement trigger, its
    layout and lists kept through every change of the batch), so its trajectory is bit-identical to the job
    run alone.  With ``shared_rebuilds`` a batch has one schedule (a trigger in any job rebuilds all), which
    takes fewer, larger rebuilds; a job's trajectory then equals its run alone up to the rounding of the
    FP32 pair sums, which its rebuilds at other steps take in another order.

    With ``aligned`` (the default) waiting jobs are admitted only at the batch's phase boundaries, every
    ``rebuild_frequency`` steps of the batch: a job's scheduled rebuilds fall every ``rebuild_frequency`` of
    its own steps, so jobs admitted at phase boundaries are due together and the batch takes one
    partial rebuild per period instead of one every step or two (only a job whose displacement trigger fires
    leaves the common phase).  A job that arrives mid-period waits at most ``rebuild_frequency`` batch
    steps (``epoch`` steps at most when the lane runs several periods at once, as it does while no job
    waits for it); finished jobs are retired at any step."""

    def __init__(self, particle_budget=None, kernel=engine.KERNEL_CLUSTER_MIXED, cluster_size=8, tile_groups=1,
                 epoch=50, keep_results=False, bytes_per_particle=1024, max_jobs=None, shared_rebuilds=False,
                 aligned=True, rebuild_frequency=10, lanes=1, dispatch="particles"):
        if lanes < 1:
            raise ValueError("a batched server needs at least one lane")
        if dispatch not in ("work", "particles"):
            raise ValueError("dispatch is 'work' or 'particles'")
        self.dispatch = _Lane.work_left if dispatch == "work" else _Lane.particles
        self.kernel = kernel
        # "size": each lane's tile shape follows the particles it holds, as job_layout picks it for one job
        # (results then equal each job alone up to the rounding of the FP32 pair sums)
        self.size_tiles = tile_groups == "size"
        if self.size_tiles:
            tile_groups = 1
        self.cluster_size = cluster_size
        self.epoch = epoch
        self.aligned = aligned
        self.frequency = rebuild_frequency
        self.max_jobs = max_jobs  # jobs in one batch at most (None: only the particle budget limits)
        self.keep_results = keep_results
        self.shared_rebuilds = shared_rebuilds
        self.lanes = []
        try:
            for _ in range(lanes):
                self.lanes.append(_Lane(engine.Engine(
                    (0.0, 0.0, 0.0), (8.0, 8.0, 8.0), RC, SKIN, cluster_size=cluster_size,
                    rebuild_frequency=rebuild_frequency, dynamic_rebuild=True, tile_groups=tile_groups,
                    shared_rebuilds=shared_rebuilds)))
        except Exception:
            self.close()
            raise
        if particle_budget is None:
            free, _ = self.engine.memory_info()
            particle_budget = int(0.8 * free / bytes_per_particle)
        self.particle_budget = particle_budget
        self.lane_budget = max(1, particle_budget // lanes)
        self.pairs = 0
        self.epochs = 0
        self.edits = 0
        self.rebuild_passes = 0  # build passes of the batches (each over the boxes due at that step)
        self._lock = threading.Lock()  # the counters above, summed over the lanes' threads

    @property
    def engine(self):
        """The first lane's context (the only one with one lane)."""
        return self.lanes[0].engine

    def close(self):
        for lane in self.lanes:
            lane.engine.close()
        self.lanes = []

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _edit(self, lane, keep, admit):
        lo = hi = cnt = None
        parts = {}
        if admit:
            systems = [job_system(j) for j in admit]
            lo = np.zeros((len(admit), 3))
            hi = np.array([[L] * 3 for L, _ in systems])
            cnt = np.array([j.n for j in admit], dtype=np.int64)
            for k, key in enumerate(("x", "y", "z", "vx", "vy", "vz")):
                parts[key] = np.concatenate([st[k] for _, st in systems])
        lane.engine.edit_boxes(keep, lo, hi, cnt, **parts)
        with self._lock:
            self.edits += 1

    def _serve_lane(self, w, t0, cond, state):
        """The loop of lane w: retire, admit from the lane's inbox, run an epoch; returns when the feeder is
        done and the lane has nothing left."""
        lane = self.lanes[w]
        while True:
            with cond:
                while not lane.active and not lane.inbox:
                    if state["fed"] or state["error"]:
                        return
                    cond.wait()
                if state["error"]:
                    return
            now = time.perf_counter() - t0
            active = lane.active
            done = [k for k, (_, left) in enumerate(active) if left == 0]
            if done:
                offsets = np.concatenate([[0], np.cumsum([j.n for j, _ in active])])
                for k in done:
                    job = active[k][0]
                    if self.keep_results:
                        job.result = lane.engine.download_range(int(offsets[k]), job.n)
                    job.finish = time.perf_counter() - t0
            keep = [k for k, (_, left) in enumerate(active) if left > 0]
            inflight = sum(active[k][0].n for k in keep)
            if not keep:
                lane.phase = 0  # an empty batch starts a new phase
            admit = []
            with cond:
                # the admission time is read under the lock the feeder appends under: a job that arrived
                # after `now` above is admitted at a later time than its arrival
                now = time.perf_counter() - t0
                while (lane.phase == 0 or not self.aligned) and lane.inbox and (
                        self.max_jobs is None or len(keep) + len(admit) < self.max_jobs) and (
                        inflight + lane.inbox[0].n <= self.lane_budget or not (keep or admit)):
                    job = lane.inbox.popleft()
                    job.start = now
                    job.worker = w
                    admit.append(job)
                    inflight += job.n
            if admit or done:
                if keep or admit:
                    self._edit(lane, keep, admit)
                lane.active = active = [active[k] for k in keep] + [[j, j.steps] for j in admit]
            if not active:
                with cond:  # jobs wait for the phase boundary only in a running batch
                    if lane.inbox:
                        continue
                    if not state["fed"]:
                        cond.wait(0.001)
                continue
            steps = min(self.epoch, min(left for _, left in active))
            if self.aligned:
                # to the next phase boundary, or with no job waiting to a later one within the epoch
                periods = 1 if lane.inbox else max(1, self.epoch // self.frequency)
                steps = min(steps, self.frequency - lane.phase + (periods - 1) * self.frequency)
            if self.size_tiles:
                g = job_layout(sum(j.n for j, _ in active))[2]
                if g != lane.groups:
                    lane.engine.set_layout(self.cluster_size, g)
                    lane.groups = g
            rs = lane.engine.step(steps, DT, self.kernel)
            lane.phase = (lane.phase + steps) % self.frequency
            with self._lock:
                self.pairs += rs.pair_interactions
                self.rebuild_passes += rs.rebuilds
                self.epochs += 1
            for a in active:
                a[1] -= steps

    def run(self, jobs):
        """Serve ``jobs`` (arrival times relative to now); returns the wall time to the last completion."""
        t0 = time.perf_counter()
        cond = threading.Condition()
        state = dict(fed=False, error=None)
        for lane in self.lanes:
            lane.active = []
            lane.inbox.clear()
            lane.phase = 0

        def lane_main(w):
            try:
                self._serve_lane(w, t0, cond, state)
            except BaseException as exc:  # re-raised by run()
                with cond:
                    state["error"] = state["error"] or exc
                    cond.notify_all()

        threads = [threading.Thread(target=lane_main, args=(w,), daemon=True) for w in range(len(self.lanes))]
        for th in threads:
            th.start()
        try:
            for job in sorted(jobs, key=lambda j: j.arrival):
                delay = job.arrival - (time.perf_counter() - t0)
                if delay > 0:
                    time.sleep(delay)
                with cond:
                    if state["error"]:
                        break
                    lane = min(self.lanes, key=self.dispatch)
                    lane.inbox.append(job)
                    cond.notify_all()
        finally:
            with cond:
                state["fed"] = True
                cond.notify_all()
            for th in threads:
                th.join()
        if state["error"] is not None:
            raise state["error"]
        return max(j.finish for j in jobs) if jobs else 0.0


def serve_open_loop_batched(count, load=0.8, seed=5, calibration_jobs=None, rate=None, **server_kw):
    """The serving measurement with continuous batching: calibrate the batched server's throughput on a
    closed-loop batch, then serve ``count`` jobs arriving as a Poisson process at ``load`` times it, or at
    ``rate`` jobs/s when given (the same arrival stream another server was measured on)."""
    with BatchedServer(**server_kw) as srv:
        calib = make_jobs(calibration_jobs or max(8, count // 4), None, seed=seed + 1)
        wall = srv.run(calib)
        capacity = len(calib) / wall
        arrival_rate = rate if rate is not None else load * capacity
        jobs = make_jobs(count, arrival_rate, seed=seed)
        srv.pairs = srv.epochs = srv.edits = srv.rebuild_passes = 0
        wall = srv.run(jobs)
        out = latency_summary(jobs, wall)
        out.update(pairs=int(srv.pairs), pairs_per_s=srv.pairs / wall if wall else 0.0,
                   load=arrival_rate / capacity, arrival_rate_jobs_per_s=arrival_rate, capacity_jobs_per_s=capacity,
                   kernel=engine.KERNEL_NAMES[srv.kernel], epochs=srv.epochs, batch_edits=srv.edits,
                   rebuild_passes=srv.rebuild_passes, aligned=srv.aligned, lanes=len(srv.lanes),
                   particle_budget=srv.particle_budget, epoch_steps=srv.epoch,
                   schedule="shared" if srv.shared_rebuilds else "per job")
    return out, jobs


def latency_summary(jobs, wall):
    lat = np.array([j.latency for j in jobs])
    pairs = sum(j.pairs for j in jobs)
    return dict(jobs=len(jobs), p50_ms=float(np.percentile(lat, 50) * 1e3), p99_ms=float(np.percentile(lat, 99) * 1e3),
                mean_ms=float(lat.mean() * 1e3), wall_s=wall, pairs=int(pairs), pairs_per_s=pairs / wall if wall else 0.0,
                jobs_per_s=len(jobs) / wall if wall else 0.0)


def serve_open_loop(count, workers=4, kernel=engine.KERNEL_CLUSTER_MIXED, load=0.8, seed=5, cluster_size=4,
                    calibration_jobs=None, **job_kw):
    """The serving measurement: calibrate the pool's throughput on a closed-loop batch, then serve
    ``count`` jobs arriving as a Poisson process at ``load`` times that throughput."""
    with Server(workers, kernel, cluster_size) as srv:
        calib = make_jobs(calibration_jobs or max(8, count // 4), None, seed=seed + 1, **job_kw)
        wall = srv.run(calib)
        capacity = len(calib) / wall
        jobs = make_jobs(count, load * capacity, seed=seed, **job_kw)
        wall = srv.run(jobs)
    out = latency_summary(jobs, wall)
    out.update(workers=workers, load=load, capacity_jobs_per_s=capacity,
               kernel="cluster tiles M8, tile shape by job size" if kernel is KERNEL_BY_SIZE
               else engine.KERNEL_NAMES[kernel])
    return out, jobs
