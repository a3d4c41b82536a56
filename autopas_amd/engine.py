"""ctypes binding of the engine's C ABI (include/ljgpu.h).

This is the thin host layer the tests and ``bench.py`` drive; it owns no compute.  Loading fails loudly
when ``libljgpu.so`` has not been built: there is no CPU fallback.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libljgpu.so")

OK = 0
ERR_ARG = 1
ERR_HIP = 2
ERR_STATE = 3
ERR_CAPACITY = 4
ERR_UNSUPPORTED = 5
ERR_NODEVICE = 6

KERNEL_LC_FP64 = 0
KERNEL_CLUSTER_FP64 = 1
KERNEL_CLUSTER_MIXED = 2
KERNEL_VERLET_MIXED = 3
KERNEL_CLUSTER_JROWS_MIXED = 4
KERNEL_CLUSTER_FP64_N3 = 5
KERNEL_CLUSTER_MIXED_N3 = 6
KERNEL_AUTO = 100
MAX_TUNE_CANDIDATES = 16
KERNEL_NAMES = {
    KERNEL_LC_FP64: "lc_fp64",
    KERNEL_CLUSTER_FP64: "cluster_fp64",
    KERNEL_CLUSTER_MIXED: "cluster_fp32pair_fp64acc",
    KERNEL_VERLET_MIXED: "verlet_fp32pair_fp64acc",
    KERNEL_CLUSTER_JROWS_MIXED: "cluster_jrows_fp32pair_fp64acc",
    KERNEL_CLUSTER_FP64_N3: "cluster_fp64_newton3",
    KERNEL_CLUSTER_MIXED_N3: "cluster_fp32pair_fp64acc_newton3",
}

_d = ctypes.c_double
_i32 = ctypes.c_int32
_i64 = ctypes.c_int64
_p = ctypes.c_void_p


class Params(ctypes.Structure):
    _fields_ = [
        ("box_min", _d * 3),
        ("box_max", _d * 3),
        ("cutoff", _d),
        ("skin", _d),
        ("cluster_size", _i32),
        ("rebuild_frequency", _i32),
        ("dynamic_rebuild", _i32),
        ("apply_shift", _i32),
        ("device", _i32),
        ("tile_groups", _i32),
        ("build_groups", _i32),
        ("tuning_samples", _i32),
        ("tuning_interval", _i32),
        ("list_mode", _i32),
        ("shared_rebuilds", _i32),
        ("graphs", _i32),
        ("whole_cluster_rows", _i32),
        ("persistent", _i32),
        ("row_headroom", _i32),
    ]


class Globals(ctypes.Structure):
    _fields_ = [
        ("potential_energy", _d),
        ("virial", _d),
        ("pair_count", _i64),
        ("candidate_evals", _i64),
    ]


class RunStats(ctypes.Structure):
    _fields_ = [
        ("steps", _i64),
        ("rebuilds", _i64),
        ("pair_interactions", _i64),
        ("force_ms", _d),
        ("force_launches", _i64),
        ("rebuild_ms", _d),
        ("wall_ms", _d),
    ]


# name -> (restype, argtypes)
SIGNATURES = {
    "ljg_create": (_i32, [ctypes.POINTER(Params), ctypes.POINTER(_p)]),
    "ljg_destroy": (_i32, [_p]),
    "ljg_last_error": (ctypes.c_char_p, [_p]),
    "ljg_device_name": (_i32, [_p, ctypes.c_char_p, _i32]),
    "ljg_memory_info": (_i32, [_p, _p, _p]),
    "ljg_set_box": (_i32, [_p, _p, _p]),
    "ljg_set_open_box": (_i32, [_p, _p, _p]),
    "ljg_set_owned": (_i32, [_p, _i64]),
    "ljg_set_boxes": (_i32, [_p, _i32, _p, _p, _p]),
    "ljg_get_box_globals": (_i32, [_p, _i32, _p, _p, _p]),
    "ljg_edit_boxes": (_i32, [_p, _i32, _p, _i32, _p, _p, _p] + [_p] * 8),
    "ljg_download_range": (_i32, [_p, _i64, _i64] + [_p] * 9),
    "ljg_get_tuning": (_i32, [_p, _p]),
    "ljg_set_types": (_i32, [_p, _i32, _p, _p, _p]),
    "ljg_upload": (_i32, [_p, _i64, _p, _p, _p, _p, _p, _p, _p, _p]),
    "ljg_set_forces": (_i32, [_p, _p, _p, _p]),
    "ljg_update_positions": (_i32, [_p] + [_p] * 7),
    "ljg_rebuild": (_i32, [_p, _i32]),
    "ljg_set_shift": (_i32, [_p, _i32]),
    "ljg_set_layout": (_i32, [_p, _i32, _i32, _i32]),
    "ljg_bin": (_i32, [_p]),
    "ljg_build_pairs": (_i32, [_p]),
    "ljg_force": (_i32, [_p, _i32, _i32, ctypes.POINTER(Globals)]),
    "ljg_step": (_i32, [_p, _i32, _d, _i32, ctypes.POINTER(RunStats)]),
    "ljg_set_timing": (_i32, [_p, _i32]),
    "ljg_download": (_i32, [_p] + [_p] * 9),
    "ljg_get_cells": (_i32, [_p, _p, _p]),
    "ljg_cluster_sizes": (_i32, [_p, _p, _p, _p, _p, _p, _p]),
    "ljg_get_clusters": (_i32, [_p, _p, _p, _p, _p, _p, _p]),
    "ljg_get_member_rows": (_i32, [_p, _p, _p]),
    "ljg_get_verlet_lists": (_i32, [_p, _p, _p]),
    "ljg_generate_lattice": (_i32, [ctypes.c_uint64, _i64, _i64, _i64, _i64, _d, _p, _d, _d, _i32] + [_p] * 7),
    "ljg_generate_slab": (_i32, [ctypes.c_uint64, _p, _d, _d, _i64, _i64, _d, _i32, _d, _i64] + [_p] * 7),
    "ljg_write_checkpoint": (_i32, [ctypes.c_char_p, ctypes.c_char_p, _i64, _i32, _i64] + [_p] * 11
                             + [ctypes.c_char_p, _i32]),
    "ljg_read_checkpoint": (_i32, [ctypes.c_char_p, _p] + [_p] * 11),
    "ljg_force_lc_noalloc": (_i32, [ctypes.POINTER(Params), _i32, _p, _p, _i64] + [_p] * 10),
}

_lib = None


def load_library(path=LIB_PATH):
    """Load libljgpu.so (once).  Raises if it is missing: the engine has no fallback path."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(
            f"{path} is missing: build the HIP engine first (python -c 'import __graft_entry__ as g; g.build()')"
        )
    lib = ctypes.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(_p)


class EngineError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"status {status}: {msg}")
        self.status = status


def generate_lattice(seed, n, dims, spacing, box_min=(0.0, 0.0, 0.0), jitter=0.0, temperature=0.0, ntypes=1):
    """Seeded jittered simple-cubic lattice (host-side generator in the engine library)."""
    lib = load_library()
    x, y, z, vx, vy, vz = (np.empty(n) for _ in range(6))
    t = np.empty(n, dtype=np.int32)
    bm = np.asarray(box_min, dtype=np.float64)
    st = lib.ljg_generate_lattice(
        seed, n, dims[0], dims[1], dims[2], spacing, _ptr(bm), jitter, temperature, ntypes,
        _ptr(x), _ptr(y), _ptr(z), _ptr(vx), _ptr(vy), _ptr(vz), _ptr(t),
    )
    if st != OK:
        raise EngineError(st, "ljg_generate_lattice")
    return x, y, z, vx, vy, vz, t


def generate_slab(seed, box, slab, n_liquid, n_gas, dmin=0.85, ntypes=1, temperature=0.0, max_attempts=None):
    """Seeded two-phase system: a liquid slab (z in [slab[0], slab[1])) in a gas, by random sequential
    placement with minimum distance dmin (host-side generator in the engine library)."""
    lib = load_library()
    n = n_liquid + n_gas
    x, y, z, vx, vy, vz = (np.empty(n) for _ in range(6))
    t = np.empty(n, dtype=np.int32)
    b = np.asarray(box, dtype=np.float64)
    attempts = max_attempts if max_attempts is not None else 200 * n
    st = lib.ljg_generate_slab(seed, _ptr(b), slab[0], slab[1], n_liquid, n_gas, dmin, ntypes, temperature,
                               attempts, _ptr(x), _ptr(y), _ptr(z), _ptr(vx), _ptr(vy), _ptr(vz), _ptr(t))
    if st != OK:
        raise EngineError(st, "ljg_generate_slab")
    return x, y, z, vx, vy, vz, t


STATE_KEYS = ("x", "y", "z", "vx", "vy", "vz", "fx", "fy", "fz")


def write_checkpoint(directory, name, iteration, state, type=None, id=None, digits=None):
    """Write state (dict of the nine FP64 arrays) as an md-flexible checkpoint; returns the .pvtu path."""
    lib = load_library()
    arrs = [np.ascontiguousarray(state[k], dtype=np.float64) for k in STATE_KEYS]
    n = len(arrs[0])
    t = np.zeros(n, dtype=np.int32) if type is None else np.ascontiguousarray(type, dtype=np.int32)
    i = np.arange(n, dtype=np.int64) if id is None else np.ascontiguousarray(id, dtype=np.int64)
    digits = digits if digits is not None else max(len(str(iteration)), 1)
    buf = ctypes.create_string_buffer(4096)
    st = lib.ljg_write_checkpoint(str(directory).encode(), str(name).encode(), iteration, digits, n,
                                  *[_ptr(a) for a in arrs], _ptr(t), _ptr(i), buf, len(buf))
    if st != OK:
        raise EngineError(st, "ljg_write_checkpoint")
    return buf.value.decode()


def read_checkpoint(path):
    """Read all pieces of an md-flexible checkpoint: (state dict, type, id)."""
    lib = load_library()
    n = np.zeros(1, dtype=np.int64)
    st = lib.ljg_read_checkpoint(str(path).encode(), _ptr(n), *([None] * 11))
    if st != OK:
        raise EngineError(st, "ljg_read_checkpoint")
    cnt = int(n[0])
    arrs = [np.empty(cnt) for _ in STATE_KEYS]
    t = np.empty(cnt, dtype=np.int32)
    i = np.empty(cnt, dtype=np.int64)
    st = lib.ljg_read_checkpoint(str(path).encode(), _ptr(n), *[_ptr(a) for a in arrs], _ptr(t), _ptr(i))
    if st != OK:
        raise EngineError(st, "ljg_read_checkpoint")
    return dict(zip(STATE_KEYS, arrs)), t, i


class Tuning(ctypes.Structure):
    _fields_ = [
        ("candidates", _i32), ("best", _i32), ("phases", _i32), ("tuning", _i32),
        ("kernel", _i32 * MAX_TUNE_CANDIDATES), ("cluster_size", _i32 * MAX_TUNE_CANDIDATES),
        ("tile_groups", _i32 * MAX_TUNE_CANDIDATES), ("build_groups", _i32 * MAX_TUNE_CANDIDATES),
        ("effective_ms", _d * MAX_TUNE_CANDIDATES),
    ]


class Engine:
    """One periodic box on one GPU, or an ensemble of independent periodic boxes (set_boxes)."""

    def __init__(self, box_min, box_max, cutoff, skin, cluster_size=4, rebuild_frequency=10,
                 dynamic_rebuild=True, apply_shift=True, device=0, tile_groups=0, build_groups=0, tuning_samples=0,
                 tuning_interval=0, list_mode=0, shared_rebuilds=False, graphs=False, whole_cluster_rows=False,
                 persistent=False, row_headroom=0):
        self.lib = load_library()
        p = Params()
        p.box_min[:] = list(box_min)
        p.box_max[:] = list(box_max)
        p.cutoff = cutoff
        p.skin = skin
        p.cluster_size = cluster_size
        p.rebuild_frequency = rebuild_frequency
        p.dynamic_rebuild = int(bool(dynamic_rebuild))
        p.apply_shift = int(bool(apply_shift))
        p.device = device
        p.tile_groups = tile_groups
        p.build_groups = build_groups
        p.tuning_samples = tuning_samples
        p.tuning_interval = tuning_interval
        p.list_mode = list_mode
        p.shared_rebuilds = int(bool(shared_rebuilds))
        p.graphs = int(bool(graphs))
        p.whole_cluster_rows = int(bool(whole_cluster_rows))
        p.persistent = int(bool(persistent))
        p.row_headroom = int(row_headroom)
        self.params = p
        h = _p()
        st = self.lib.ljg_create(ctypes.byref(p), ctypes.byref(h))
        if st != OK:
            raise EngineError(st, "ljg_create")
        self.h = h
        self.n = 0

    def _check(self, st, what):
        if st != OK:
            msg = self.lib.ljg_last_error(self.h)
            raise EngineError(st, f"{what}: {msg.decode() if msg else ''}")

    def close(self):
        if getattr(self, "h", None):
            self.lib.ljg_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def device_name(self):
        buf = ctypes.create_string_buffer(256)
        self._check(self.lib.ljg_device_name(self.h, buf, 256), "ljg_device_name")
        return buf.value.decode()

    def memory_info(self):
        """(free, total) device memory in bytes."""
        f, t = ctypes.c_uint64(0), ctypes.c_uint64(0)
        self._check(self.lib.ljg_memory_info(self.h, ctypes.byref(f), ctypes.byref(t)), "ljg_memory_info")
        return f.value, t.value

    def tuning(self):
        """State of the layout tuning (kernel=KERNEL_AUTO): candidates with their scores and the choice."""
        t = Tuning()
        self._check(self.lib.ljg_get_tuning(self.h, ctypes.byref(t)), "ljg_get_tuning")
        n = t.candidates
        cands = [dict(kernel=t.kernel[i], cluster_size=t.cluster_size[i], tile_groups=t.tile_groups[i],
                      build_groups=t.build_groups[i], effective_ms=t.effective_ms[i]) for i in range(n)]
        return dict(best=t.best, phases=t.phases, tuning=bool(t.tuning), candidates=cands)

    def set_box(self, box_min, box_max):
        """Move the context to a new periodic box; the particles are dropped (upload the next system)."""
        lo = np.ascontiguousarray(box_min, dtype=np.float64)
        hi = np.ascontiguousarray(box_max, dtype=np.float64)
        self._check(self.lib.ljg_set_box(self.h, _ptr(lo), _ptr(hi)), "ljg_set_box")
        self.nboxes = 0
        self.n = 0
        self._types = self._ids = None

    def set_open_box(self, box_min, box_max):
        """Open (non-periodic) box with halo particles: owned particles in [box_min, box_max), halo
        particles (made by the caller) up to rc + skin outside it; no periodic images.  The particles are
        dropped; step() is refused on an open box."""
        lo = np.ascontiguousarray(box_min, dtype=np.float64)
        hi = np.ascontiguousarray(box_max, dtype=np.float64)
        self._check(self.lib.ljg_set_open_box(self.h, _ptr(lo), _ptr(hi)), "ljg_set_open_box")
        self.nboxes = 0
        self.n = 0
        self._types = self._ids = None

    def set_owned(self, n_owned):
        """The first n_owned particles of the last upload are owned, the rest halo: the globals count a pair
        as LJFunctor does for owned particles (owned-halo pairs half, halo-halo pairs not at all)."""
        self._check(self.lib.ljg_set_owned(self.h, int(n_owned)), "ljg_set_owned")

    def set_boxes(self, box_min, box_max, counts):
        """Turn the context into an ensemble of independent periodic boxes: box b spans
        [box_min[b], box_max[b]) (arrays of shape (nboxes, 3)) and holds counts[b] particles; the next
        upload takes the boxes' particles concatenated in box order.  The particles are dropped."""
        lo = np.ascontiguousarray(box_min, dtype=np.float64).reshape(-1, 3)
        hi = np.ascontiguousarray(box_max, dtype=np.float64).reshape(-1, 3)
        cnt = np.ascontiguousarray(counts, dtype=np.int64)
        if not (len(lo) == len(hi) == len(cnt)):
            raise ValueError("box_min, box_max and counts describe different numbers of boxes")
        self._check(self.lib.ljg_set_boxes(self.h, len(cnt), _ptr(lo), _ptr(hi), _ptr(cnt)), "ljg_set_boxes")
        self.nboxes = len(cnt)
        self._counts = cnt.copy()
        self.n = 0
        self._types = self._ids = None

    def edit_boxes(self, keep, box_min=None, box_max=None, counts=None, x=None, y=None, z=None, vx=None, vy=None,
                   vz=None, type=None, id=None):
        """Keep the boxes ``keep`` (indices into the current ensemble, in their new order) with their state,
        then append new boxes (bounds of shape (nadd, 3), counts and their particles concatenated)."""
        k = np.ascontiguousarray(keep, dtype=np.int32)
        nadd = 0 if counts is None else len(counts)
        lo = hi = cnt = None
        arrs = [None] * 6
        t = i = None
        if nadd:
            lo = np.ascontiguousarray(box_min, dtype=np.float64).reshape(-1, 3)
            hi = np.ascontiguousarray(box_max, dtype=np.float64).reshape(-1, 3)
            cnt = np.ascontiguousarray(counts, dtype=np.int64)
            arrs = [None if a is None else np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, z, vx, vy, vz)]
            t = None if type is None else np.ascontiguousarray(type, dtype=np.int32)
            i = None if id is None else np.ascontiguousarray(id, dtype=np.int64)
        old_counts = getattr(self, "_counts", np.zeros(0, dtype=np.int64))
        self._check(self.lib.ljg_edit_boxes(self.h, len(k), _ptr(k) if len(k) else None, nadd, _ptr(lo), _ptr(hi),
                                            _ptr(cnt), *[_ptr(a) for a in arrs], _ptr(t), _ptr(i)), "ljg_edit_boxes")
        self.nboxes = len(k) + nadd
        self.n = int(self._box_counts(k, cnt).sum())
        self._edit_host_types_ids(old_counts, k, cnt, t, i)

    def _edit_host_types_ids(self, old_counts, keep, added, t, i):
        """Host copies of types and ids in the new upload order, as ljg_edit_boxes arranges them: the kept
        boxes' particles keep theirs, added particles take the given ones (type 0 and their new upload
        index by default)."""
        if getattr(self, "_types", None) is None or len(self._types) != int(np.sum(old_counts)):
            self._types = self._ids = None
            return
        start = np.concatenate([[0], np.cumsum(old_counts)])
        parts_t = [self._types[start[b]:start[b + 1]] for b in keep]
        parts_i = [self._ids[start[b]:start[b + 1]] for b in keep]
        kept = int(sum(len(p) for p in parts_t))
        nadd = 0 if added is None else int(np.sum(added))
        parts_t.append(np.zeros(nadd, dtype=np.int32) if t is None else t)
        parts_i.append(np.arange(kept, kept + nadd, dtype=np.int64) if i is None else i)
        self._types = np.concatenate(parts_t).astype(np.int32)
        self._ids = np.concatenate(parts_i).astype(np.int64)

    def _box_counts(self, keep, added):
        old = getattr(self, "_counts", np.zeros(0, dtype=np.int64))
        new = np.concatenate([old[np.asarray(keep, dtype=np.int64)] if len(keep) else np.zeros(0, dtype=np.int64),
                              added if added is not None else np.zeros(0, dtype=np.int64)])
        self._counts = new
        return new

    def download_range(self, first, count):
        """State of particles [first, first + count) of the upload order."""
        out = [np.empty(count) for _ in range(9)]
        self._check(self.lib.ljg_download_range(self.h, first, count, *[_ptr(a) for a in out]), "ljg_download_range")
        keys = ("x", "y", "z", "vx", "vy", "vz", "fx", "fy", "fz")
        return dict(zip(keys, out))

    def box_globals(self):
        """Per-box potential energy, virial and pair count of the last force(globals=True) on an ensemble."""
        nb = getattr(self, "nboxes", 0)
        e, v = np.empty(nb), np.empty(nb)
        pc = np.empty(nb, dtype=np.int64)
        self._check(self.lib.ljg_get_box_globals(self.h, nb, _ptr(e), _ptr(v), _ptr(pc)), "ljg_get_box_globals")
        return dict(potential_energy=e, virial=v, pair_count=pc)

    def set_types(self, epsilon, sigma, mass=None):
        e = np.ascontiguousarray(epsilon, dtype=np.float64)
        s = np.ascontiguousarray(sigma, dtype=np.float64)
        m = None if mass is None else np.ascontiguousarray(mass, dtype=np.float64)
        self._check(self.lib.ljg_set_types(self.h, len(e), _ptr(e), _ptr(s), _ptr(m)), "ljg_set_types")

    def upload(self, x, y, z, vx=None, vy=None, vz=None, type=None, id=None):
        arrs = [np.ascontiguousarray(a, dtype=np.float64) if a is not None else None for a in (x, y, z, vx, vy, vz)]
        t = None if type is None else np.ascontiguousarray(type, dtype=np.int32)
        i = None if id is None else np.ascontiguousarray(id, dtype=np.int64)
        n = len(arrs[0])
        self._types = np.zeros(n, dtype=np.int32) if t is None else t.copy()
        self._ids = np.arange(n, dtype=np.int64) if i is None else i.copy()
        self._check(
            self.lib.ljg_upload(self.h, n, *[_ptr(a) for a in arrs], _ptr(t), _ptr(i)), "ljg_upload"
        )
        self.n = n

    def set_forces(self, fx, fy, fz):
        f = [np.ascontiguousarray(a, dtype=np.float64) for a in (fx, fy, fz)]
        self._check(self.lib.ljg_set_forces(self.h, *[_ptr(a) for a in f]), "ljg_set_forces")

    def update_positions(self, x, y, z, vx=None, vy=None, vz=None):
        """Overwrite positions (and velocities) in upload order, keeping the lists; returns whether a
        particle is now skin/2 or more from its position at the last rebuild."""
        arrs = [None if a is None else np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, z, vx, vy, vz)]
        moved = ctypes.c_int32(0)
        self._check(self.lib.ljg_update_positions(self.h, *[_ptr(a) for a in arrs], ctypes.byref(moved)),
                    "ljg_update_positions")
        return bool(moved.value)

    def rebuild(self, kernel=KERNEL_CLUSTER_MIXED):
        self._check(self.lib.ljg_rebuild(self.h, kernel), "ljg_rebuild")

    def set_layout(self, cluster_size, tile_groups=0, build_groups=0):
        """Cluster size and tile shape of the next build (a tuner switching configurations)."""
        self._check(self.lib.ljg_set_layout(self.h, cluster_size, tile_groups, build_groups), "ljg_set_layout")

    def set_shift(self, apply_shift):
        self._check(self.lib.ljg_set_shift(self.h, int(bool(apply_shift))), "ljg_set_shift")

    def bin(self):
        self._check(self.lib.ljg_bin(self.h), "ljg_bin")

    def build_pairs(self):
        self._check(self.lib.ljg_build_pairs(self.h), "ljg_build_pairs")

    def force(self, kernel=KERNEL_CLUSTER_MIXED, globals=True):
        g = Globals()
        self._check(self.lib.ljg_force(self.h, kernel, int(bool(globals)), ctypes.byref(g)), "ljg_force")
        return g

    def set_timing(self, on=True):
        self._check(self.lib.ljg_set_timing(self.h, int(bool(on))), "ljg_set_timing")

    def step(self, nsteps, dt, kernel=KERNEL_CLUSTER_MIXED):
        rs = RunStats()
        self._check(self.lib.ljg_step(self.h, nsteps, dt, kernel, ctypes.byref(rs)), "ljg_step")
        return rs

    def write_checkpoint(self, directory, name, iteration, digits=None):
        """Write the current state (upload order, uploaded types and ids) as an md-flexible checkpoint."""
        t, i = getattr(self, "_types", None), getattr(self, "_ids", None)
        if t is None or len(t) != self.n or len(i) != self.n:
            raise EngineError(3, "write_checkpoint: the host copies of types and ids do not describe the current "
                                 "particles (upload them again)")
        st = self.download()
        return write_checkpoint(directory, name, iteration, st, type=t, id=i, digits=digits)

    def load_checkpoint(self, path):
        """Upload positions, velocities, types and ids from an md-flexible checkpoint and restore its
        forces as the current force (MDFlexConfig.cpp:173), so the next step's drift uses them."""
        st, t, i = read_checkpoint(path)
        self.upload(st["x"], st["y"], st["z"], st["vx"], st["vy"], st["vz"], type=t, id=i)
        self.set_forces(st["fx"], st["fy"], st["fz"])
        return len(t)

    def download(self):
        out = [np.empty(self.n) for _ in range(9)]
        self._check(self.lib.ljg_download(self.h, *[_ptr(a) for a in out]), "ljg_download")
        keys = ("x", "y", "z", "vx", "vy", "vz", "fx", "fy", "fz")
        return dict(zip(keys, out))

    def cells(self):
        c = np.empty(self.n, dtype=np.int64)
        nh = np.empty(3, dtype=np.int64)
        self._check(self.lib.ljg_get_cells(self.h, _ptr(c), _ptr(nh)), "ljg_get_cells")
        return c, nh

    def clusters(self):
        ns, nc, ne = (np.zeros(1, dtype=np.int64) for _ in range(3))
        tpd = np.zeros(2, dtype=np.int64)
        side = np.zeros(2)
        r = np.zeros(1, dtype=np.int64)
        self._check(
            self.lib.ljg_cluster_sizes(self.h, _ptr(ns), _ptr(nc), _ptr(ne), _ptr(tpd), _ptr(side), _ptr(r)),
            "ljg_cluster_sizes",
        )
        ns, nc, ne = int(ns[0]), int(nc[0]), int(ne[0])
        out = dict(
            slot_particle=np.empty(ns, dtype=np.int64),
            cluster_tower=np.empty(nc, dtype=np.int32),
            list_offsets=np.empty(nc + 1, dtype=np.int64),
            list_j=np.empty(max(ne, 1), dtype=np.int32),
            list_shift=np.empty(max(ne, 1), dtype=np.int32),
            bbox=np.empty(6 * nc),
        )
        self._check(
            self.lib.ljg_get_clusters(
                self.h, *[_ptr(out[k]) for k in ("slot_particle", "cluster_tower", "list_offsets", "list_j",
                                                   "list_shift", "bbox")]
            ),
            "ljg_get_clusters",
        )
        out["list_j"] = out["list_j"][:ne]
        out["list_shift"] = out["list_shift"][:ne]
        out.update(towers_per_dim=tpd, side=side, towers_per_il=int(r[0]))
        return out

    def member_rows(self):
        """Per i-cluster of a cluster-entry build: the j-cluster members the FP32 cluster kernel streams, as
        (offsets, rows) with rows[offsets[c]:offsets[c + 1]] = entry << 3 | member."""
        nc = np.zeros(1, dtype=np.int64)
        self._check(self.lib.ljg_cluster_sizes(self.h, None, _ptr(nc), None, None, None, None), "ljg_cluster_sizes")
        off = np.zeros(int(nc[0]) + 1, dtype=np.int64)
        self._check(self.lib.ljg_get_member_rows(self.h, _ptr(off), None), "ljg_get_member_rows")
        rows = np.empty(max(int(off[-1]), 1), dtype=np.int32)
        self._check(self.lib.ljg_get_member_rows(self.h, _ptr(off), _ptr(rows)), "ljg_get_member_rows")
        return off, rows[:int(off[-1])]

    def verlet_lists(self):
        """Neighbour lists (upload indices) after a Verlet-list build: (offsets[n+1], neighbors)."""
        off = np.empty(self.n + 1, dtype=np.int64)
        self._check(self.lib.ljg_get_verlet_lists(self.h, _ptr(off), None), "ljg_get_verlet_lists")
        nb = np.empty(max(int(off[-1]), 1), dtype=np.int32)
        self._check(self.lib.ljg_get_verlet_lists(self.h, _ptr(off), _ptr(nb)), "ljg_get_verlet_lists")
        return off, nb[: int(off[-1])]
