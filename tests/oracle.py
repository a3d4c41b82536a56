"""ctypes binding of the CPU oracle (oracle/libljoracle.so) -- test infrastructure only.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this module.
"""
import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_LIB = os.path.join(ORACLE_DIR, "libljoracle.so")

_d = ctypes.c_double
_i64 = ctypes.c_int64
_p = ctypes.c_void_p


class CellBlock(ctypes.Structure):
    _fields_ = [
        ("box_min", _d * 3), ("box_max", _d * 3), ("cell_len", _d * 3), ("recip", _d * 3),
        ("halo_min", _d * 3), ("halo_max", _d * 3), ("cells_per_dim_halo", _i64 * 3),
        ("cells_per_il", _i64), ("first_owned", _i64), ("last_owned", _i64),
    ]


_lib = None


def build_oracle():
    r = subprocess.run(["make", "-s", "-C", ORACLE_DIR], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{r.stdout}\n{r.stderr}")


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(ORACLE_LIB):
        build_oracle()
    L = ctypes.CDLL(ORACLE_LIB)
    L.ljo_lj_pair.restype = ctypes.c_int
    L.ljo_lj_pair.argtypes = [_p, _d, _d, _d, _d, _p, _p]
    L.ljo_calc_shift6.restype = _d
    L.ljo_calc_shift6.argtypes = [_d, _d, _d]
    L.ljo_mixing.restype = None
    L.ljo_mixing.argtypes = [ctypes.c_int, _p, _p, _d, ctypes.c_int, _p, _p, _p]
    L.ljo_cellblock_build.restype = None
    L.ljo_cellblock_build.argtypes = [_p, _p, _d, _d, ctypes.POINTER(CellBlock)]
    L.ljo_cell_index3d.restype = None
    L.ljo_cell_index3d.argtypes = [ctypes.POINTER(CellBlock), _p, _p]
    L.ljo_cell_index_batch.restype = None
    L.ljo_cell_index_batch.argtypes = [ctypes.POINTER(CellBlock), _i64, _p, _p, _p, _p]
    L.ljo_c08_offsets.restype = ctypes.c_int
    L.ljo_c08_offsets.argtypes = [_p, _p, _d, _p, _p, ctypes.c_int]
    L.ljo_wrap_periodic.restype = None
    L.ljo_wrap_periodic.argtypes = [_i64, _p, _p, _p, _p, _p]
    L.ljo_direct_sum.restype = _i64
    L.ljo_direct_sum.argtypes = [_i64, _p, _p, _p, _p, ctypes.c_int, _p, _p, _p, _d, _p, ctypes.c_int,
                                 _p, _p, _p, _p, _p]
    L.ljo_lc_forces.restype = _i64
    L.ljo_lc_forces.argtypes = [_i64, _p, _p, _p, _p, ctypes.c_int, _p, _p, _p, _d, _d, _p, _p, ctypes.c_int,
                                _p, _p, _p, _p, _p]
    L.ljo_c08_forces_halo.restype = _i64
    L.ljo_c08_forces_halo.argtypes = [_i64, _p, _p, _p, _p, ctypes.c_int, _p, _p, _p, _d, _d, _p, _p,
                                      ctypes.c_int, ctypes.c_int, _p, _p, _p, _p, _p, _p]
    L.ljo_vcl_build_periodic.restype = _p
    L.ljo_vcl_build_periodic.argtypes = [_i64, _p, _p, _p, _p, _p, _p, _d, ctypes.c_int]
    L.ljo_vcl_sizes.restype = None
    L.ljo_vcl_sizes.argtypes = [_p, _p, _p, _p, _p, _p, _p]
    L.ljo_vcl_get.restype = None
    L.ljo_vcl_get.argtypes = [_p] * 7
    L.ljo_vcl_free.restype = None
    L.ljo_vcl_free.argtypes = [_p]
    L.ljo_vcl_forward_only.restype = None
    L.ljo_vcl_forward_only.argtypes = [_p]
    L.ljo_verlet_pairs_fixed.restype = _i64
    L.ljo_verlet_pairs_fixed.argtypes = [_i64, _p, _p, _p, _p, _p, _d, _p, _p, _i64]
    L.ljo_verlet_positions.restype = None
    L.ljo_verlet_positions.argtypes = [_i64] + [_p] * 14 + [_d]
    L.ljo_verlet_velocities.restype = None
    L.ljo_verlet_velocities.argtypes = [_i64] + [_p] * 11 + [_d]
    _lib = L
    return L


def P(a):
    return None if a is None else a.ctypes.data_as(_p)


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def lj_pair(dr, eps24, sig2, shift6, rc2):
    f = np.zeros(3)
    p6 = np.zeros(1)
    hit = lib().ljo_lj_pair(P(f64(dr)), eps24, sig2, shift6, rc2, P(f), P(p6))
    return bool(hit), f, p6[0]


def mixing(eps, sig, cutoff, apply_shift=True):
    n = len(eps)
    e24, s2, sh = np.zeros(n * n), np.zeros(n * n), np.zeros(n * n)
    lib().ljo_mixing(n, P(f64(eps)), P(f64(sig)), cutoff, int(apply_shift), P(e24), P(s2), P(sh))
    return e24, s2, sh


def cellblock(box_min, box_max, ilen, csf=1.0):
    cb = CellBlock()
    lib().ljo_cellblock_build(P(f64(box_min)), P(f64(box_max)), ilen, csf, ctypes.byref(cb))
    return cb


def cell_index3d(cb, pos):
    out = np.zeros(3, dtype=np.int64)
    lib().ljo_cell_index3d(ctypes.byref(cb), P(f64(pos)), P(out))
    return out


def cell_index_batch(cb, x, y, z):
    out = np.zeros(len(x), dtype=np.int64)
    lib().ljo_cell_index_batch(ctypes.byref(cb), len(x), P(f64(x)), P(f64(y)), P(f64(z)), P(out))
    return out


def c08_offsets(cells_per_dim, cell_len, ilen):
    o2 = np.zeros(256, dtype=np.int64)
    o1 = np.zeros(256, dtype=np.int64)
    n = lib().ljo_c08_offsets(P(np.asarray(cells_per_dim, dtype=np.int64)), P(f64(cell_len)), ilen, P(o2), P(o1),
                              256)
    assert n >= 0
    return list(zip(o2[:n].tolist(), o1[:n].tolist()))


def wrap(x, y, z, box_min, box_max):
    x, y, z = f64(x).copy(), f64(y).copy(), f64(z).copy()
    lib().ljo_wrap_periodic(len(x), P(x), P(y), P(z), P(f64(box_min)), P(f64(box_max)))
    return x, y, z


def _tables(eps, sig, cutoff, apply_shift):
    return mixing(eps, sig, cutoff, apply_shift)


def direct_sum(x, y, z, cutoff, box_len=None, types=None, eps=(1.0,), sig=(1.0,), apply_shift=True, threads=0):
    n = len(x)
    e24, s2, sh = _tables(eps, sig, cutoff, apply_shift)
    fx, fy, fz = np.zeros(n), np.zeros(n), np.zeros(n)
    ep, vi = np.zeros(1), np.zeros(1)
    t = None if types is None else np.ascontiguousarray(types, dtype=np.int32)
    bl = None if box_len is None else f64(box_len)
    pairs = lib().ljo_direct_sum(n, P(f64(x)), P(f64(y)), P(f64(z)), P(t), len(eps), P(e24), P(s2), P(sh), cutoff,
                                 P(bl), threads, P(fx), P(fy), P(fz), P(ep), P(vi))
    return dict(fx=fx, fy=fy, fz=fz, epot=ep[0], virial=vi[0], pairs=int(pairs))


def lc_forces(x, y, z, cutoff, skin, box_min, box_max, types=None, eps=(1.0,), sig=(1.0,), apply_shift=True,
              threads=0):
    n = len(x)
    e24, s2, sh = _tables(eps, sig, cutoff, apply_shift)
    fx, fy, fz = np.zeros(n), np.zeros(n), np.zeros(n)
    ep, vi = np.zeros(1), np.zeros(1)
    t = None if types is None else np.ascontiguousarray(types, dtype=np.int32)
    pairs = lib().ljo_lc_forces(n, P(f64(x)), P(f64(y)), P(f64(z)), P(t), len(eps), P(e24), P(s2), P(sh), cutoff,
                                skin, P(f64(box_min)), P(f64(box_max)), threads, P(fx), P(fy), P(fz), P(ep), P(vi))
    return dict(fx=fx, fy=fy, fz=fz, epot=ep[0], virial=vi[0], pairs=int(pairs))


def c08_forces_halo(x, y, z, cutoff, skin, box_min, box_max, types=None, eps=(1.0,), sig=(1.0,), apply_shift=True,
                    threads=0, repeats=1):
    n = len(x)
    e24, s2, sh = _tables(eps, sig, cutoff, apply_shift)
    fx, fy, fz = np.zeros(n), np.zeros(n), np.zeros(n)
    ep, vi, ts = np.zeros(1), np.zeros(1), np.zeros(1)
    t = None if types is None else np.ascontiguousarray(types, dtype=np.int32)
    pairs = lib().ljo_c08_forces_halo(n, P(f64(x)), P(f64(y)), P(f64(z)), P(t), len(eps), P(e24), P(s2), P(sh),
                                      cutoff, skin, P(f64(box_min)), P(f64(box_max)), threads, repeats, P(fx), P(fy),
                                      P(fz), P(ep), P(vi), P(ts))
    return dict(fx=fx, fy=fy, fz=fz, epot=ep[0], virial=vi[0], pairs=int(pairs), traverse_seconds=ts[0])


def vcl_periodic(x, y, z, ids, box_min, box_max, ilen, cluster_size, newton3=False):
    L = lib()
    idarr = None if ids is None else np.ascontiguousarray(ids, dtype=np.int64)
    h = L.ljo_vcl_build_periodic(len(x), P(f64(x)), P(f64(y)), P(f64(z)), P(idarr), P(f64(box_min)),
                                 P(f64(box_max)), ilen, cluster_size)
    try:
        if newton3:
            L.ljo_vcl_forward_only(h)
        ns, nc, ne = (np.zeros(1, dtype=np.int64) for _ in range(3))
        tpd = np.zeros(2, dtype=np.int64)
        side = np.zeros(2)
        r = np.zeros(1, dtype=np.int64)
        L.ljo_vcl_sizes(h, P(ns), P(nc), P(ne), P(tpd), P(side), P(r))
        ns, nc, ne = int(ns[0]), int(nc[0]), int(ne[0])
        out = dict(
            slot_particle=np.empty(ns, dtype=np.int64),
            cluster_tower=np.empty(nc, dtype=np.int32),
            list_offsets=np.empty(nc + 1, dtype=np.int64),
            list_j=np.empty(max(ne, 1), dtype=np.int32),
            list_shift=np.empty(max(ne, 1), dtype=np.int32),
            bbox=np.empty(max(6 * nc, 1)),
        )
        L.ljo_vcl_get(h, *[P(out[k]) for k in ("slot_particle", "cluster_tower", "list_offsets", "list_j",
                                                "list_shift", "bbox")])
        out["list_j"] = out["list_j"][:ne]
        out["list_shift"] = out["list_shift"][:ne]
        out["bbox"] = out["bbox"][: 6 * nc]
        out.update(towers_per_dim=tpd, side=side, towers_per_il=int(r[0]))
        return out
    finally:
        L.ljo_vcl_free(h)


def verlet_positions(x, y, z, vx, vy, vz, fx, fy, fz, mass_of_type, types, dt):
    ofx, ofy, ofz = np.zeros_like(fx), np.zeros_like(fy), np.zeros_like(fz)
    t = None if types is None else np.ascontiguousarray(types, dtype=np.int32)
    lib().ljo_verlet_positions(len(x), P(x), P(y), P(z), P(vx), P(vy), P(vz), P(fx), P(fy), P(fz), P(ofx), P(ofy),
                               P(ofz), P(f64(mass_of_type)), P(t), dt)
    return ofx, ofy, ofz


def verlet_velocities(vx, vy, vz, fx, fy, fz, ofx, ofy, ofz, mass_of_type, types, dt):
    t = None if types is None else np.ascontiguousarray(types, dtype=np.int32)
    lib().ljo_verlet_velocities(len(vx), P(vx), P(vy), P(vz), P(fx), P(fy), P(fz), P(ofx), P(ofy), P(ofz),
                                P(f64(mass_of_type)), P(t), dt)


def verlet_pairs_fixed(x, y, z, box_min, box_max, ilen):
    n = len(x)
    off = np.zeros(n + 1, dtype=np.int64)
    cap = 200 * n
    nb = np.zeros(max(cap, 1), dtype=np.int32)
    tot = lib().ljo_verlet_pairs_fixed(n, P(f64(x)), P(f64(y)), P(f64(z)), P(f64(box_min)), P(f64(box_max)), ilen,
                                       P(off), P(nb), cap)
    assert tot >= 0
    return off, nb[:tot]
