"""Open (non-periodic) boxes with halo particles, AutoPas's own model (LinkedCells.h:87, VerletClusterLists.h:194):
the caller uploads the owned particles followed by halo particles, as md-flexible's halo exchange makes them.

Checks, per kernel:
  * forces on every uploaded particle against the FP64 direct sum over the same particles without periodic
    images (FP64 kernels 1e-10 relative, FP32-pair kernels 1e-5 RMS), and the pair count exact;
  * with the periodic images of a box as halo (every copy within rc + skin of a face) and the owned count
    set, the owned particles' forces and the globals equal the periodic system's (the direct sum with
    minimum image): an owned-halo pair counts half, as in LJFunctor.h:174-189, so the two halves of each pair
    across the boundary add up to the pair;
  * the entry points refuse what an open box does not do (device steps, particles beyond the halo region,
    Newton-3 globals with halo particles).
"""
import itertools

import numpy as np
import pytest

import autopas_amd as ap
from conftest import announce

pytestmark = pytest.mark.gpu

RC = 2.5
SKIN = 0.3
ILEN = RC + SKIN
FP64 = (ap.KERNEL_LC_FP64, ap.KERNEL_CLUSTER_FP64)
KERNELS = (ap.KERNEL_LC_FP64, ap.KERNEL_CLUSTER_FP64, ap.KERNEL_CLUSTER_MIXED, ap.KERNEL_CLUSTER_JROWS_MIXED,
           ap.KERNEL_VERLET_MIXED)


def lattice(dims, density, seed):
    a = density ** (-1.0 / 3.0)
    n = int(np.prod(dims))
    x, y, z, vx, vy, vz, _ = ap.generate_lattice(seed, n, tuple(dims), a, (0.0, 0.0, 0.0), 0.05, 0.9, 1)
    return np.stack([x, y, z], axis=1), np.array(dims) * a


def periodic_halo(pos, L):
    """md-flexible's periodic halo exchange for one domain: every periodic image of an owned particle that
    lies within rc + skin outside the box."""
    out = []
    for s in itertools.product((-1, 0, 1), repeat=3):
        if s == (0, 0, 0):
            continue
        p = pos + np.array(s) * L
        keep = np.all((p >= -ILEN) & (p <= L + ILEN), axis=1)
        out.append(p[keep])
    return np.concatenate(out)


def rel_max(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def rms_rel(f, ref, sl=slice(None)):
    num = sum(np.sum((f[k][sl] - ref[k][sl]) ** 2) for k in ("fx", "fy", "fz"))
    den = sum(np.sum(ref[k][sl] ** 2) for k in ("fx", "fy", "fz"))
    return float(np.sqrt(num / den))


def open_engine(L, M=8):
    e = ap.Engine((0.0, 0.0, 0.0), tuple(L), RC, SKIN, cluster_size=M)
    e.set_open_box((0.0, 0.0, 0.0), tuple(L))
    return e


@pytest.mark.parametrize("kernel", KERNELS)
@pytest.mark.parametrize("dims", [(12, 12, 12), (10, 11, 14)])
def test_open_box_forces_vs_direct_sum(oracle, kernel, dims):
    pos, L = lattice(dims, 0.75, seed=21)
    halo = periodic_halo(pos, L)
    allp = np.concatenate([pos, halo])
    announce(f"open box {dims}, {len(pos)} owned + {len(halo)} halo, kernel={ap.KERNEL_NAMES[kernel]}")
    with open_engine(L) as e:
        e.upload(allp[:, 0], allp[:, 1], allp[:, 2])
        g = e.force(kernel, globals=False)
        f = e.download()
    for k, d in zip(("x", "y", "z"), range(3)):
        assert np.array_equal(f[k], allp[:, d]), k  # an open box does not wrap positions
    ref = oracle.direct_sum(allp[:, 0], allp[:, 1], allp[:, 2], RC, box_len=None, threads=8)
    if kernel in FP64:
        for k in ("fx", "fy", "fz"):
            assert rel_max(f[k], ref[k]) <= 1e-10, k
    else:
        assert rms_rel(f, ref) <= 1e-5
    assert g.pair_count == ref["pairs"]


@pytest.mark.parametrize("kernel", KERNELS)
def test_open_box_with_periodic_halo_equals_periodic_box(oracle, kernel):
    pos, L = lattice((12, 12, 12), 0.75, seed=22)
    n = len(pos)
    halo = periodic_halo(pos, L)
    allp = np.concatenate([pos, halo])
    announce(f"open box + periodic halo vs periodic direct sum, kernel={ap.KERNEL_NAMES[kernel]}")
    with open_engine(L) as e:
        e.upload(allp[:, 0], allp[:, 1], allp[:, 2])
        e.set_owned(n)
        g = e.force(kernel, globals=True)
        f = e.download()
    ref = oracle.direct_sum(pos[:, 0], pos[:, 1], pos[:, 2], RC, box_len=tuple(L), threads=8)
    own = slice(0, n)
    tol = 1e-10 if kernel in FP64 else 1e-5
    if kernel in FP64:
        for k in ("fx", "fy", "fz"):
            assert rel_max(f[k][own], ref[k]) <= tol, k
    else:
        assert rms_rel({k: f[k][own] for k in ("fx", "fy", "fz")}, ref) <= tol
    assert abs(g.potential_energy - ref["epot"]) <= tol * abs(ref["epot"])
    assert abs(g.virial - ref["virial"]) <= tol * abs(ref["virial"])


def test_open_box_owned_count_changes_only_the_globals():
    pos, L = lattice((10, 10, 10), 0.7, seed=23)
    halo = periodic_halo(pos, L)
    allp = np.concatenate([pos, halo])
    announce("open box: ownership masks the globals, not the forces")
    with open_engine(L) as e:
        e.upload(allp[:, 0], allp[:, 1], allp[:, 2])
        g_all = e.force(ap.KERNEL_CLUSTER_JROWS_MIXED, globals=True)
        f_all = e.download()
        e.set_owned(len(pos))
        g_own = e.force(ap.KERNEL_CLUSTER_JROWS_MIXED, globals=True)
        f_own = e.download()
        e.set_owned(0)
        g_none = e.force(ap.KERNEL_CLUSTER_JROWS_MIXED, globals=True)
    for k in ("fx", "fy", "fz"):
        assert np.array_equal(f_all[k], f_own[k]), k
    assert g_all.pair_count == g_own.pair_count == g_none.pair_count
    assert g_none.potential_energy == 0.0 and g_none.virial == 0.0
    assert g_own.potential_energy != g_all.potential_energy


def test_open_box_refusals():
    pos, L = lattice((8, 8, 8), 0.75, seed=24)
    announce("open box refusals")
    with open_engine(L) as e:
        bad = pos.copy()
        bad[3, 1] = L[1] + ILEN + 0.01  # beyond the halo region
        with pytest.raises(ap.EngineError):
            e.upload(bad[:, 0], bad[:, 1], bad[:, 2])
        e.upload(pos[:, 0], pos[:, 1], pos[:, 2])
        with pytest.raises(ap.EngineError):
            e.step(1, 0.002, ap.KERNEL_CLUSTER_MIXED)
        with pytest.raises(ap.EngineError):
            e.set_owned(len(pos) + 1)
        e.set_owned(len(pos) - 10)
        with pytest.raises(ap.EngineError):
            e.force(ap.KERNEL_CLUSTER_FP64_N3, globals=True)
        e.force(ap.KERNEL_CLUSTER_FP64_N3, globals=False)  # forces alone are fine
        # back to a periodic box: steps run again
        e.set_box((0.0, 0.0, 0.0), tuple(L))
        e.upload(pos[:, 0], pos[:, 1], pos[:, 2])
        e.step(2, 0.002, ap.KERNEL_CLUSTER_MIXED)
