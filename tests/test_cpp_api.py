"""C++ host API (include/autopas_gpu.hpp) through its test driver tests/cpp/autopas_api_test.

CPU: misuse raises AutoPasException with the reference's messages; init() without a device throws.
GPU: computeInteractions forces and globals against the FP64 direct sum (RMS relative force error <= 1e-5,
energy and virial 1e-5 relative, for the FP32-pair kernels); md-flexible's step sequence through the API
(host integrator, device forces) against AutoPas::step on the device; the same host steps on an open box
with md-flexible's periodic boundary made by the caller (halo particles) against the periodic box.  The two loops build their lists
at different points (the host loop after the first drift, the device loop before it), so they agree to
the rounding of the FP32 pair term: positions within 1e-6, velocities within 1e-4 after 25 steps; with
the FP64 cluster kernel within 1e-10 and 1e-8.
"""
import os
import subprocess

import numpy as np
import pytest

from conftest import announce

HERE = os.path.dirname(os.path.abspath(__file__))
BIN = os.path.join(HERE, "cpp", "autopas_api_test")
RC = 2.5


@pytest.fixture(scope="module")
def driver():
    if not os.path.exists(BIN):
        from autopas_amd import build

        build.build_cpp_test()
    return BIN


def run(driver, *args, timeout=120):
    r = subprocess.run([driver, *args], capture_output=True, text=True, timeout=timeout)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout


def read_states(path):
    out, cur = {}, None
    with open(path) as f:
        for line in f:
            if line.startswith("#"):
                continue
            w = line.split()
            if w[0] in ("globals", "stats"):
                out.setdefault(w[0], []).append([float(v) for v in w[1:]])
            elif len(w) == 2:
                cur = w[0]
                out[cur] = []
            else:
                out[cur].append([float(v) for v in w])
    for k in list(out):
        out[k] = np.array(out[k])
    return out


def test_api_misuse_raises(driver):
    assert "ok errors" in run(driver, "errors")


def test_init_without_device_throws(driver):
    assert "ok nodevice" in run(driver, "nodevice")


def rms_rel(f, fr):
    return float(np.sqrt(np.sum((f - fr) ** 2) / np.sum(fr ** 2)))


@pytest.mark.gpu
def test_compute_interactions_matches_direct_sum(driver, oracle, tmp_path):
    announce("C++ API computeInteractions: 8^3 lattice, single type (j-row kernel) and two types (cluster kernel)")
    path = str(tmp_path / "forces.txt")
    run(driver, "forces", path)
    st = read_states(path)
    n_side, rho = 8, 0.75
    L = n_side * rho ** (-1.0 / 3.0)
    for k, tag in enumerate(("single", "mixing")):
        s = st[tag]
        order = np.argsort(s[:, 0])
        s = s[order]
        x, y, z, f = s[:, 1], s[:, 2], s[:, 3], s[:, 7:10]
        if tag == "single":
            ref = oracle.direct_sum(x, y, z, RC, box_len=(L, L, L))
        else:
            import autopas_amd as ap

            _, _, _, _, _, _, types = ap.generate_lattice(3, n_side ** 3, (n_side,) * 3, rho ** (-1.0 / 3.0),
                                                          (0.0, 0.0, 0.0), 0.05, 0.9, 2)
            ref = oracle.direct_sum(x, y, z, RC, box_len=(L, L, L), types=types, eps=(1.0, 1.2), sig=(1.0, 1.05))
        fr = np.stack([ref["fx"], ref["fy"], ref["fz"]], axis=1)
        assert rms_rel(f, fr) <= 1e-5, tag
        e, v = st["globals"][k][0], st["globals"][k][1]
        assert abs(e - ref["epot"]) <= 1e-5 * abs(ref["epot"]), tag
        assert abs(v - ref["virial"]) <= 1e-5 * abs(ref["virial"]), tag
    hit_rate, flops = st["globals"][0][2], st["globals"][0][3]
    assert 0.05 < hit_rate < 1.0
    assert flops > 0


@pytest.mark.gpu
def test_functor_flop_model_full_and_newton3(driver, oracle, tmp_path):
    # LJFunctor.h:722-789 with shift and globals: 8 per distance, 15 + 9 per kernel call without Newton-3
    # (two calls per pair within rc on full lists), 18 + 13 per call with it (one per pair).  From each
    # line's hit rate h (calls per distance) and FLOPs F the pair count follows; both equal the oracle's.
    announce("C++ API functor FLOP model: full lists and Newton-3, 8^3 lattice")
    path = str(tmp_path / "flops.txt")
    run(driver, "flops", path)
    rows = [ln.split() for ln in open(path).read().splitlines() if ln.startswith("flops")]
    assert len(rows) == 2
    L = 8 * 0.75 ** (-1.0 / 3.0)
    import autopas_amd as ap

    x, y, z, _, _, _, _ = ap.generate_lattice(3, 8 ** 3, (8,) * 3, 0.75 ** (-1.0 / 3.0), (0.0, 0.0, 0.0), 0.05, 0.9, 1)
    pairs = oracle.direct_sum(x, y, z, RC, box_len=(L, L, L))["pairs"]
    for _, kernel, h, flops in rows:
        h, flops = float(h), int(flops)
        per_call = 15 + 9 if int(kernel) == ap.KERNEL_CLUSTER_FP64 else 18 + 13
        dist = flops / (8 + per_call * h)
        calls = h * dist
        assert abs(dist - round(dist)) < 1e-3 * dist
        expect_calls = 2 * pairs if int(kernel) == ap.KERNEL_CLUSTER_FP64 else pairs
        assert round(calls) == expect_calls, (kernel, calls, expect_calls)


@pytest.mark.gpu
def test_compute_interactions_rebuild_policy(driver):
    announce("C++ API computeInteractions: rebuild every 10 calls and on a skin/2 move")
    assert "ok rebuilds" in run(driver, "rebuilds")


@pytest.mark.gpu
def test_compute_interactions_autotuner_picks_gpu_configuration(driver, oracle, tmp_path):
    # LJG_KERNEL_AUTO: one tuning phase over the cartesian product of the allowed GPU options (3 samples per
    # configuration), then the fastest configuration by the tuner's effective time; the forces of the call
    # after the phase match the FP64 direct sum (FP32-pair tolerance: the choice may be an FP32-pair kernel)
    announce("C++ API AutoTuner over the GPU configurations: 12^3 lattice")
    path = str(tmp_path / "tune.txt")
    run(driver, "tune", path, timeout=300)
    lines = open(path).read().splitlines()
    head = lines[0].split()
    space, calls, tuning_calls, seen = int(head[2]), int(head[4]), int(head[6]), int(head[8])
    ev = [ln for ln in lines if ln.startswith("# evidence")]
    print("\n".join(ev + [ln for ln in lines if ln.startswith("# chosen")]), flush=True)
    # LC fp64; Verlet fp32pair; cluster tiles fp64/fp32pair x N3 off/on x M 4/8; j-rows fp32pair x M 4/8
    assert space == 1 + 1 + 8 + 2
    assert len(ev) == space and seen == space
    assert tuning_calls == 3 * space
    st = read_states(path)
    s = st["final"][np.argsort(st["final"][:, 0])]
    x, y, z, f = s[:, 1], s[:, 2], s[:, 3], s[:, 7:10]
    L = 12 * 0.75 ** (-1.0 / 3.0)
    ref = oracle.direct_sum(x, y, z, RC, box_len=(L, L, L))
    fr = np.stack([ref["fx"], ref["fy"], ref["fz"]], axis=1)
    assert rms_rel(f, fr) <= 1e-5
    e, v = st["globals"][0][0], st["globals"][0][1]
    assert abs(e - ref["epot"]) <= 1e-5 * abs(ref["epot"])
    assert abs(v - ref["virial"]) <= 1e-5 * abs(ref["virial"])


@pytest.mark.gpu
@pytest.mark.parametrize("kernel,tol_x,tol_v", [(4, 1e-6, 1e-4), (2, 1e-6, 1e-4), (1, 1e-10, 1e-8)])
def test_md_loop_through_api_matches_device_steps(driver, tmp_path, kernel, tol_x, tol_v):
    announce(f"C++ API md loop vs AutoPas::step: kernel {kernel}, 8^3 lattice, 25 steps")
    path = str(tmp_path / "loop.txt")
    run(driver, "loop", path, str(kernel))
    st = read_states(path)
    h = st["host"][np.argsort(st["host"][:, 0])]
    d = st["device"][np.argsort(st["device"][:, 0])]
    assert st["stats"][0][0] == 25 and st["stats"][0][1] >= 3
    n_side, rho = 8, 0.75
    L = n_side * rho ** (-1.0 / 3.0)
    dx = h[:, 1:4] - d[:, 1:4]
    dx -= L * np.round(dx / L)  # one side may have wrapped a particle the other has not yet
    assert np.max(np.abs(dx)) <= tol_x
    assert np.max(np.abs(h[:, 4:7] - d[:, 4:7])) <= tol_v
    # the loop moved the particles: not a trivially equal pair of initial states
    assert np.max(np.abs(h[:, 4:7])) > 0.1


@pytest.mark.gpu
@pytest.mark.parametrize("kernel,tol_x,tol_e", [(0, 1e-9, 1e-10), (4, 1e-6, 1e-5)])
def test_open_box_with_halo_exchange_matches_periodic_box(driver, tmp_path, kernel, tol_x, tol_e):
    # md-flexible's periodic boundary through AutoPas's own model: an open box, leaving particles wrapped
    # and re-added, periodic images added as halo particles each step.  Same trajectory and globals as the
    # periodic box (the two build different cell grids, so the sums round differently: kernel 0 is the FP64
    # linked-cells kernel, 4 the FP32-pair j-row kernel)
    announce(f"C++ API open box with halo exchange vs periodic box: kernel {kernel}, 8^3 lattice, 30 steps")
    path = str(tmp_path / "halo.txt")
    run(driver, "halo", path, str(kernel))
    s = read_states(path)
    (eo, vo, halo, _), (ep, vp, _, _) = s["globals"]
    assert halo > 1000  # the halo shell of an 8.8-sigma box
    assert abs(eo - ep) <= tol_e * abs(ep) and abs(vo - vp) <= tol_e * abs(vp)
    o = s["open"][np.argsort(s["open"][:, 0])]
    p = s["periodic"][np.argsort(s["periodic"][:, 0])]
    assert np.array_equal(o[:, 0], p[:, 0])
    L = 8 * 0.75 ** (-1.0 / 3.0)
    dx = o[:, 1:4] - p[:, 1:4]
    dx -= L * np.round(dx / L)
    assert np.max(np.abs(dx)) <= tol_x
    assert np.max(np.abs(o[:, 4:7] - p[:, 4:7])) <= 100 * tol_x
    assert np.max(np.abs(p[:, 4:7])) > 0.1  # the particles moved
