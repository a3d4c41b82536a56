"""Batched ensemble: many independent periodic LJ boxes advanced together in one engine context.

Box b of the ensemble workload has N_b = round(exp(U(ln 64, ln 8192))) particles, density
rho_b ~ U(0.3, 0.9), cube side L_b = (N_b / rho_b)^(1/3), a jittered simple-cubic lattice with
k = ceil(N_b^(1/3)) sites per side (spacing L_b / k, the first N_b sites, jitter U(-0.05, 0.05) sigma),
velocities Maxwell-Boltzmann at T_b ~ U(0.7, 1.5) with zero net momentum, and its own seed 4000 + b.
The draws come from numpy's Philox4x32 generator; the lattices from the engine's seeded generator.

md-flexible would run each box as a simulation of its own (one AutoPas instance each,
``AutoPasImpl.h:55-113``).  Here the boxes' particles are concatenated in box order and every launch
of the step covers all boxes (``Engine.set_boxes``); each box keeps the tower grid, cluster layout and
arithmetic it has alone.
"""
import math
from dataclasses import dataclass

import numpy as np

from . import engine

RC = 2.5
SKIN = 0.3
DT = 0.002


@dataclass
class Box:
    index: int
    n: int
    density: float
    temperature: float
    seed: int

    @property
    def length(self):
        return (self.n / self.density) ** (1.0 / 3.0)


def make_boxes(count, seed=4, n_range=(64, 8192), density_range=(0.3, 0.9), temperature_range=(0.7, 1.5),
               first_seed=4000):
    """Draw ``count`` boxes of the ensemble workload."""
    rng = np.random.Generator(np.random.Philox(seed))
    lo, hi = math.log(n_range[0]), math.log(n_range[1])
    boxes = []
    for b in range(count):
        n = int(round(math.exp(rng.uniform(lo, hi))))
        n = min(max(n, n_range[0]), n_range[1])
        boxes.append(Box(index=b, n=n, density=float(rng.uniform(*density_range)),
                         temperature=float(rng.uniform(*temperature_range)), seed=first_seed + b))
    return boxes


def box_system(box):
    """Positions and velocities of one box in [0, L)^3."""
    k = int(math.ceil(box.n ** (1.0 / 3.0) - 1e-9))
    L = box.length
    x, y, z, vx, vy, vz, _ = engine.generate_lattice(box.seed, box.n, (k, k, k), L / k, (0.0, 0.0, 0.0), 0.05,
                                                     box.temperature)
    return dict(x=x, y=y, z=z, vx=vx, vy=vy, vz=vz)


def ensemble_arrays(boxes):
    """Box bounds (nboxes x 3 each), counts and the concatenated particle arrays of ``boxes``."""
    lo = np.zeros((len(boxes), 3))
    hi = np.array([[b.length] * 3 for b in boxes])
    counts = np.array([b.n for b in boxes], dtype=np.int64)
    parts = [box_system(b) for b in boxes]
    state = {k: np.concatenate([p[k] for p in parts]) for k in ("x", "y", "z", "vx", "vy", "vz")}
    return lo, hi, counts, state


def box_slices(counts):
    """Slice of each box in the concatenated arrays."""
    ends = np.cumsum(counts)
    return [slice(int(e - n), int(e)) for n, e in zip(counts, ends)]


def ensemble_engine(boxes, cluster_size=4, rebuild_frequency=10, dynamic_rebuild=True, tile_groups=0, arrays=None,
                    whole_cluster_rows=False):
    """An engine context holding ``boxes`` as one ensemble, particles uploaded."""
    lo, hi, counts, st = arrays if arrays is not None else ensemble_arrays(boxes)
    e = engine.Engine(lo[0], hi[0], RC, SKIN, cluster_size=cluster_size, rebuild_frequency=rebuild_frequency,
                      dynamic_rebuild=dynamic_rebuild, tile_groups=tile_groups, whole_cluster_rows=whole_cluster_rows)
    try:
        e.set_boxes(lo, hi, counts)
        e.upload(st["x"], st["y"], st["z"], st["vx"], st["vy"], st["vz"])
    except Exception:
        e.close()
        raise
    return e
