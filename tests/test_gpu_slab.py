"""The two-phase, four-type system (liquid slab at rho 0.8 in gas at rho 0.02, LB mixing, rc = 3): a
scaled-down box against the oracle's direct sum for every kernel, and the full 2M-particle box through
size-independent checks (the FP32-pair kernel against the FP64 one, Newton's third law, a short run)."""
import numpy as np
import pytest

import autopas_amd as ap
from conftest import announce
from test_abi import slab_box

pytestmark = pytest.mark.gpu

EPS = [0.8, 1.0, 1.1, 1.2]
SIG = [0.9, 1.0, 1.05, 1.1]
RC = 3.0
SKIN = 0.3


def rms_rel(f, fr):
    num = sum(np.sum((f[k] - fr[k]) ** 2) for k in ("fx", "fy", "fz"))
    den = sum(np.sum(fr[k] ** 2) for k in ("fx", "fy", "fz"))
    return float(np.sqrt(num / den))


def slab_engine(box, M=8):
    e = ap.Engine((0.0, 0.0, 0.0), tuple(box), RC, SKIN, cluster_size=M, rebuild_frequency=10)
    e.set_types(EPS, SIG, [1.0] * 4)
    return e


KERNELS = [ap.KERNEL_LC_FP64, ap.KERNEL_CLUSTER_FP64, ap.KERNEL_CLUSTER_MIXED, ap.KERNEL_VERLET_MIXED,
           ap.KERNEL_CLUSTER_JROWS_MIXED]


@pytest.mark.parametrize("kernel", KERNELS)
def test_slab_small_vs_direct_sum(oracle, kernel):
    box, slab, n_l, n_g = slab_box(8.0)
    announce(f"two-phase slab 1/8 scale ({n_l + n_g} particles), kernel={ap.KERNEL_NAMES[kernel]}")
    x, y, z, vx, vy, vz, t = ap.generate_slab(3, box, slab, n_l, n_g, dmin=0.85, ntypes=4, temperature=0.9)
    with slab_engine(box) as e:
        e.upload(x, y, z, vx, vy, vz, type=t)
        g = e.force(kernel, globals=True)
        f = e.download()
    ref = oracle.direct_sum(f["x"], f["y"], f["z"], RC, box_len=tuple(box), types=t, eps=EPS, sig=SIG)
    fp64 = kernel in (ap.KERNEL_LC_FP64, ap.KERNEL_CLUSTER_FP64)
    tol = 1e-10 if fp64 else 1e-5
    assert rms_rel(f, ref) <= tol
    assert abs(g.potential_energy - ref["epot"]) <= tol * abs(ref["epot"])
    assert abs(g.virial - ref["virial"]) <= tol * abs(ref["virial"])
    if fp64:
        assert g.pair_count == ref["pairs"]
    else:
        assert abs(g.pair_count - ref["pairs"]) <= 1e-4 * ref["pairs"]


def test_slab_full_size():
    box, slab, n_l, n_g = slab_box(1.0)
    announce(f"two-phase slab full size ({n_l + n_g} particles): j-rows vs FP64 cluster kernel, 20 steps")
    x, y, z, vx, vy, vz, t = ap.generate_slab(3, box, slab, n_l, n_g, dmin=0.85, ntypes=4, temperature=0.9)
    res = {}
    for kernel in (ap.KERNEL_CLUSTER_FP64, ap.KERNEL_CLUSTER_JROWS_MIXED):
        with slab_engine(box) as e:
            e.upload(x, y, z, vx, vy, vz, type=t)
            g = e.force(kernel, globals=True)
            res[kernel] = (e.download(), g)
    f64, g64 = res[ap.KERNEL_CLUSTER_FP64]
    f32, g32 = res[ap.KERNEL_CLUSTER_JROWS_MIXED]
    # Newton's third law on the FP64 forces
    tot = np.array([f64[k].sum() for k in ("fx", "fy", "fz")])
    mag = sum(np.abs(f64[k]).sum() for k in ("fx", "fy", "fz"))
    assert np.all(np.abs(tot) <= 1e-10 * mag)
    err = rms_rel(f32, f64)
    print(f"two-phase slab: force RMS relative error of the FP32-pair kernel vs FP64: {err:.3e}")
    assert err <= 1e-5
    assert abs(g32.potential_energy - g64.potential_energy) <= 1e-5 * abs(g64.potential_energy)
    assert abs(g32.pair_count - g64.pair_count) <= 1e-4 * g64.pair_count
    # a short run through rebuilds of the inhomogeneous towers
    with slab_engine(box) as e:
        e.upload(x, y, z, vx, vy, vz, type=t)
        rs = e.step(20, 0.002, ap.KERNEL_CLUSTER_JROWS_MIXED)
        st = e.download()
    assert rs.rebuilds >= 2
    assert all(np.all(np.isfinite(st[k])) for k in st)
    # random sequential packing at d_min 0.85 stores ~10-70 eps in each close pair, which the run turns
    # into heat; momentum stays zero up to the FP32 pair term's rounding
    p = np.array([st[k].sum() for k in ("vx", "vy", "vz")])
    pabs = sum(np.abs(st[k]).sum() for k in ("vx", "vy", "vz"))
    assert np.all(np.abs(p) <= 1e-6 * pabs)
