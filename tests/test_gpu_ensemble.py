"""Ensemble of independent periodic boxes in one context, against each box alone and against the oracle.

Tolerances (stated per test):
  * a box in the ensemble against the same box in a context of its own (same kernel, cluster size and
    tile shape, lists rebuilt at the same steps): bit-identical positions, velocities and forces;
  * per-box forces against the FP64 direct sum over periodic images: RMS relative error <= 1e-5; per-box
    potential energy and virial 1e-5 relative, pair count within 1e-4 relative.
"""
import numpy as np
import pytest

import autopas_amd as ap
from autopas_amd import ensemble as en
from conftest import announce

pytestmark = pytest.mark.gpu


def rms_rel(f, fr):
    num = sum(np.sum((f[k] - fr[k]) ** 2) for k in ("fx", "fy", "fz"))
    den = sum(np.sum(fr[k] ** 2) for k in ("fx", "fy", "fz"))
    return float(np.sqrt(num / den))


def small_boxes():
    # uneven sizes, among them boxes shorter than 2 (rc + skin) where the tower window wraps onto itself
    boxes = en.make_boxes(40, seed=11, n_range=(64, 1500))
    picked = sorted(boxes, key=lambda b: b.length)[:4] + boxes[:8]
    assert min(b.length for b in picked) < 2 * (en.RC + en.SKIN)
    return picked


@pytest.mark.parametrize("M,G", [(4, 8), (8, 2)])
def test_ensemble_matches_boxes_alone(M, G):
    boxes = small_boxes()
    announce(f"ensemble vs boxes alone M={M} G={G}: {len(boxes)} boxes, 25 steps, rebuild every 10")
    arrays = en.ensemble_arrays(boxes)
    lo, hi, counts, st = arrays
    with en.ensemble_engine(boxes, cluster_size=M, dynamic_rebuild=False, tile_groups=G, arrays=arrays) as e:
        rs = e.step(25, en.DT, ap.KERNEL_CLUSTER_MIXED)
        out = e.download()
    pairs = 0
    for b, sl in zip(boxes, en.box_slices(counts)):
        with ap.Engine((0.0, 0.0, 0.0), (b.length,) * 3, en.RC, en.SKIN, cluster_size=M,
                        rebuild_frequency=10, dynamic_rebuild=False, tile_groups=G) as a:
            a.upload(st["x"][sl], st["y"][sl], st["z"][sl], st["vx"][sl], st["vy"][sl], st["vz"][sl])
            ra = a.step(25, en.DT, ap.KERNEL_CLUSTER_MIXED)
            alone = a.download()
        assert ra.rebuilds == rs.rebuilds
        pairs += ra.pair_interactions
        for k in alone:
            assert np.array_equal(out[k][sl], alone[k]), (b.index, b.n, k)
    assert rs.pair_interactions == pairs


def test_ensemble_forces_vs_oracle(oracle):
    boxes = small_boxes()
    announce(f"ensemble forces and per-box globals vs direct sum: {len(boxes)} boxes")
    arrays = en.ensemble_arrays(boxes)
    lo, hi, counts, st = arrays
    with en.ensemble_engine(boxes, arrays=arrays) as e:
        g = e.force(ap.KERNEL_CLUSTER_MIXED, globals=True)
        f = e.download()
        bg = e.box_globals()
    tot_e = tot_p = 0
    for k, (b, sl) in enumerate(zip(boxes, en.box_slices(counts))):
        L = b.length
        ref = oracle.direct_sum(f["x"][sl], f["y"][sl], f["z"][sl], en.RC, box_len=(L, L, L))
        fb = {c: f[c][sl] for c in ("fx", "fy", "fz")}
        assert rms_rel(fb, ref) <= 1e-5, (b.index, b.n, L)
        assert abs(bg["potential_energy"][k] - ref["epot"]) <= 1e-5 * abs(ref["epot"]), k
        assert abs(bg["virial"][k] - ref["virial"]) <= 1e-5 * abs(ref["virial"]), k
        assert abs(bg["pair_count"][k] - ref["pairs"]) <= max(1, 1e-4 * ref["pairs"]), k
        tot_e += bg["potential_energy"][k]
        tot_p += bg["pair_count"][k]
    assert g.pair_count == tot_p
    assert abs(g.potential_energy - tot_e) <= 1e-12 * abs(tot_e)


def test_ensemble_box_globals_match_box_alone():
    boxes = small_boxes()[:5]
    announce("per-box globals in the ensemble vs each box alone")
    arrays = en.ensemble_arrays(boxes)
    lo, hi, counts, st = arrays
    with en.ensemble_engine(boxes, arrays=arrays) as e:
        e.force(ap.KERNEL_CLUSTER_MIXED, globals=True)
        bg = e.box_globals()
    for k, (b, sl) in enumerate(zip(boxes, en.box_slices(counts))):
        with ap.Engine((0.0, 0.0, 0.0), (b.length,) * 3, en.RC, en.SKIN, cluster_size=4) as a:
            a.upload(st["x"][sl], st["y"][sl], st["z"][sl])
            ga = a.force(ap.KERNEL_CLUSTER_MIXED, globals=True)
        assert bg["pair_count"][k] == ga.pair_count
        assert abs(bg["potential_energy"][k] - ga.potential_energy) <= 1e-12 * abs(ga.potential_energy)
        assert abs(bg["virial"][k] - ga.virial) <= 1e-12 * abs(ga.virial)


def test_ensemble_rejects_other_kernels_and_bad_input():
    boxes = small_boxes()[:3]
    announce("ensemble argument checks")
    arrays = en.ensemble_arrays(boxes)
    lo, hi, counts, st = arrays
    with en.ensemble_engine(boxes, arrays=arrays) as e:
        for k in (ap.KERNEL_LC_FP64, ap.KERNEL_CLUSTER_FP64, ap.KERNEL_VERLET_MIXED, ap.KERNEL_CLUSTER_JROWS_MIXED):
            with pytest.raises(ap.EngineError) as ex:
                e.force(k)
            assert ex.value.status == ap.engine.ERR_UNSUPPORTED
        with pytest.raises(ap.EngineError) as ex:
            e.box_globals()  # no force with globals yet
        assert ex.value.status == ap.engine.ERR_STATE
        with pytest.raises(ap.EngineError) as ex:
            e.upload(st["x"][:-1], st["y"][:-1], st["z"][:-1])
        assert ex.value.status == ap.engine.ERR_ARG
        bad = hi.copy()
        bad[1] = 2.0  # shorter than rc + skin
        with pytest.raises(ap.EngineError) as ex:
            e.set_boxes(lo, bad, counts)
        assert ex.value.status == ap.engine.ERR_ARG
        # back to one box: every kernel runs again
        b = boxes[0]
        sl = en.box_slices(counts)[0]
        e.set_box((0.0, 0.0, 0.0), (b.length,) * 3)
        e.upload(st["x"][sl], st["y"][sl], st["z"][sl])
        e.force(ap.KERNEL_LC_FP64)


def test_ensemble_workload_sample_steps():
    # a 256-box sample of the ensemble workload (N in [64, 8192]): momentum stays zero per box and the
    # trajectory is finite; the pair count matches the per-box globals
    boxes = en.make_boxes(256, seed=4)
    announce(f"ensemble workload sample: {len(boxes)} boxes, {sum(b.n for b in boxes)} particles, 20 steps")
    arrays = en.ensemble_arrays(boxes)
    lo, hi, counts, st = arrays
    with en.ensemble_engine(boxes, arrays=arrays) as e:
        rs = e.step(20, en.DT, ap.KERNEL_CLUSTER_MIXED)
        out = e.download()
        g = e.force(ap.KERNEL_CLUSTER_MIXED, globals=True)
        bg = e.box_globals()
    assert rs.steps == 20 and rs.rebuilds >= 2
    for k in out:
        assert np.all(np.isfinite(out[k])), k
    for b, sl in zip(boxes, en.box_slices(counts)):
        p = np.array([out[c][sl].sum() for c in ("vx", "vy", "vz")])
        assert np.all(np.abs(p) <= 1e-5 * b.n), (b.index, p)
    assert g.pair_count == int(bg["pair_count"].sum())
