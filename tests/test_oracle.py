"""CPU tests pinning the oracle (oracle/ljoracle.c) before it is trusted as the checker.

Each group checks the oracle against the reference's own known answers (tests/golden, copied as data
from the reference's test files) and against what the mathematics fixes independently of the formula the
oracle implements (force = -dU/dr by finite differences, zero net force on a perfect lattice, shell
counts, pure-Python brute force on tiny inputs).
"""
import itertools
import json
import math
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "reference_known_answers.json")))


# ---- LJ pair term -----------------------------------------------------------------------------------
def test_lj_known_answer(oracle):
    g = GOLD["lj_two_particles"]
    dr = np.subtract(g["p1"], g["p2"])
    hit, f, _ = oracle.lj_pair(dr, 24.0 * g["epsilon"], g["sigma"] ** 2, 0.0, g["cutoff"] ** 2)
    assert hit
    assert np.allclose(f, g["force_on_p1"], rtol=0, atol=g["abs_tolerance"])


def test_lj_known_answer_mixing(oracle):
    g = GOLD["lj_two_particles"]
    e24, s2, sh = oracle.mixing([g["epsilon"], g["epsilon2"]], [g["sigma"], g["sigma2"]], g["cutoff"], False)
    dr = np.subtract(g["p1"], g["p2"])
    hit, f, _ = oracle.lj_pair(dr, e24[1], s2[1], sh[1], g["cutoff"] ** 2)
    assert hit
    assert np.allclose(f, g["force_on_p1_mixing"], rtol=0, atol=g["abs_tolerance"] * 100)


@pytest.mark.parametrize("case", GOLD["force_calculation_grid"]["cases"])
def test_force_calculation_grid(oracle, case):
    a = 2 ** (1 / 6) if case["spacing"] == "2^(1/6)" else case["spacing"]
    pos = np.array([[0.5, 0.5, 0.5], [0.5 + a, 0.5, 0.5], [0.5, 0.5 + a, 0.5], [0.5 + a, 0.5 + a, 0.5]])
    r = oracle.direct_sum(pos[:, 0], pos[:, 1], pos[:, 2], case["cutoff"], box_len=None, apply_shift=False)
    got = np.stack([r["fx"], r["fy"], r["fz"]], axis=1)
    assert np.allclose(got, np.array(case["forces"], dtype=float), rtol=0, atol=case["tolerance"])


@pytest.mark.parametrize("r", [0.9, 1.0, 2 ** (1 / 6), 1.3, 1.7, 2.4])
def test_force_is_minus_potential_gradient(oracle, r):
    # independent of the force formula: F(r) = -dU/dr with U = pot6 / 6 (shift-free), central differences
    eps, sig = 1.2, 1.05
    e24, s2, _ = oracle.mixing([eps], [sig], 3.0, False)
    h = 1e-6

    def u(rr):
        return oracle.lj_pair(np.array([rr, 0, 0]), e24[0], s2[0], 0.0, 9.0)[2] / 6.0

    dudr = (u(r + h) - u(r - h)) / (2 * h)
    f = oracle.lj_pair(np.array([r, 0, 0]), e24[0], s2[0], 0.0, 9.0)[1]
    assert f[0] == pytest.approx(-dudr, rel=1e-6, abs=1e-6)
    # and the potential itself against 4 eps [(s/r)^12 - (s/r)^6] written with r, not r^2
    assert u(r) == pytest.approx(4 * eps * ((sig / r) ** 12 - (sig / r) ** 6), rel=1e-12)


def test_force_zero_at_potential_minimum(oracle):
    _, f, _ = oracle.lj_pair(np.array([2 ** (1 / 6), 0.0, 0.0]), 24.0, 1.0, 0.0, 4.0)
    assert abs(f[0]) < 1e-13


def test_cutoff_inclusive_and_exclusive(oracle):
    rc = 2.5
    hit, _, _ = oracle.lj_pair(np.array([rc, 0.0, 0.0]), 24.0, 1.0, 0.0, rc * rc)
    assert hit
    hit, f, _ = oracle.lj_pair(np.array([np.nextafter(rc, 3.0), 0.0, 0.0]), 24.0, 1.0, 0.0, rc * rc)
    assert not hit and not f.any()


def test_shift_makes_potential_vanish_at_cutoff(oracle):
    for eps, sig, rc in [(1.0, 1.0, 2.5), (0.8, 0.9, 3.0), (1.2, 1.1, 3.0)]:
        e24, s2, sh = oracle.mixing([eps], [sig], rc, True)
        _, _, p6 = oracle.lj_pair(np.array([rc, 0.0, 0.0]), e24[0], s2[0], sh[0], rc * rc)
        assert abs(p6) < 1e-14 * abs(e24[0])


# ---- mixing -----------------------------------------------------------------------------------------
def test_mixing_reference_values(oracle):
    g = GOLD["ppl_mixing_with_unset_site"]
    e24, s2, _ = oracle.mixing(g["epsilon"], g["sigma"], 1.0, False)
    assert e24[0 * 2 + 1] == g["mixing24Epsilon_0_1"]
    assert s2[1 * 2 + 0] == g["mixingSigmaSquared_1_0"]


def test_lorentz_berthelot_closed_form(oracle):
    eps = [0.8, 1.0, 1.1, 1.2]
    sig = [0.9, 1.0, 1.05, 1.1]
    e24, s2, _ = oracle.mixing(eps, sig, 3.0, False)
    for a, b in itertools.product(range(4), repeat=2):
        assert e24[4 * a + b] == pytest.approx(24 * math.sqrt(eps[a] * eps[b]), rel=1e-15)
        assert s2[4 * a + b] == pytest.approx(((sig[a] + sig[b]) / 2) ** 2, rel=1e-15)
        assert e24[4 * a + b] == e24[4 * b + a]


# ---- CellBlock3D -------------------------------------------------------------------------------------
@pytest.mark.parametrize("blk", GOLD["cellblock_boundaries"]["blocks"])
def test_cellblock_boundaries(oracle, blk):
    cb = oracle.cellblock(blk["box_min"], blk["box_max"], blk["interaction_length"], blk["csf"])
    nh = list(cb.cells_per_dim_halo)
    halo = cb.cells_per_il
    bmin, bmax = blk["box_min"], blk["box_max"]
    for combo in itertools.product(range(4), repeat=3):
        pos = []
        for d, k in enumerate(combo):
            pos.append([np.nextafter(bmin[d], -1.0), bmin[d], np.nextafter(bmax[d], -1.0), bmax[d]][k])
        idx = oracle.cell_index3d(cb, pos)
        for d, k in enumerate(combo):
            want = [halo - 1, halo, nh[d] - halo - 1, nh[d] - halo][k]
            assert idx[d] == want, (combo, d)


@pytest.mark.parametrize("case", GOLD["cellblock_mesh_index"]["cases"])
def test_cellblock_mesh_index(oracle, case):
    blk = GOLD["cellblock_boundaries"]["blocks"][case["block"]]
    cb = oracle.cellblock(blk["box_min"], blk["box_max"], blk["interaction_length"], blk["csf"])
    counter = 0
    for z in range(case["num"][2]):
        for y in range(case["num"][1]):
            for x in range(case["num"][0]):
                pos = [case["start"][0] + x * case["dr"][0], case["start"][1] + y * case["dr"][1],
                       case["start"][2] + z * case["dr"][2]]
                assert oracle.cell_index_batch(cb, [pos[0]], [pos[1]], [pos[2]])[0] == counter
                counter += 1


def test_cellblock_geometry_closed_form(oracle):
    cb = oracle.cellblock((0, 0, 0), (35.2, 35.2, 35.2), 2.8)
    assert list(cb.cells_per_dim_halo) == [14, 14, 14]  # floor(35.2 / 2.8) = 12 owned + 2 halo
    assert cb.cell_len[0] == pytest.approx(35.2 / 12)
    assert cb.first_owned == 14 * 14 + 14 + 1


# ---- C08 offsets -------------------------------------------------------------------------------------
def test_c08_offsets_overlap1_exact_order(oracle):
    g = GOLD["c08_offsets_overlap1"]
    pairs = oracle.c08_offsets(g["cells_per_dimension"], g["cell_length"], g["interaction_length"])
    assert [list(p) for p in pairs] == g["pairs"]
    assert sorted(abs(a - b) for a, b in pairs) == g["sorted_differences"]


def test_c08_offsets_overlap2(oracle):
    g = GOLD["c08_offsets_overlap2"]
    pairs = oracle.c08_offsets(GOLD["c08_offsets_overlap1"]["cells_per_dimension"], [1, 1, 1],
                               g["interaction_length"])
    assert len(pairs) == g["count"]
    assert sorted(abs(a - b) for a, b in pairs) == g["sorted_differences"]


# ---- forces at system scale --------------------------------------------------------------------------
def _lattice(n_side, a, jitter=0.0, seed=1):
    rng = np.random.default_rng(seed)
    g = (np.arange(n_side) + 0.5) * a
    x, y, z = (v.ravel() for v in np.meshgrid(g, g, g, indexing="ij"))
    if jitter:
        x, y, z = (v + rng.uniform(-jitter, jitter, v.size) for v in (x, y, z))
    return x, y, z


def test_perfect_lattice_shells_and_zero_force(oracle):
    # spacing 1.1, rc 2.5: shells n^2 = 1..5 hold 6+12+8+6+24 = 56 neighbours -> 28 unique pairs/particle
    n_side, a = 8, 1.1
    x, y, z = _lattice(n_side, a)
    L = n_side * a
    r = oracle.direct_sum(x, y, z, 2.5, box_len=(L, L, L), apply_shift=False)
    assert r["pairs"] == 28 * n_side ** 3
    assert np.max(np.abs(np.concatenate([r["fx"], r["fy"], r["fz"]]))) < 1e-11
    shells = {1: 6, 2: 12, 3: 8, 4: 6, 5: 24}
    e_pp = sum(cnt * 4 * ((1 / (a * a * k)) ** 6 - (1 / (a * a * k)) ** 3) for k, cnt in shells.items()) / 2
    assert r["epot"] == pytest.approx(e_pp * n_side ** 3, rel=1e-12)


def _python_brute_force(x, y, z, L, rc):
    n = len(x)
    f = np.zeros((n, 3))
    k = int(math.ceil(rc / L))
    for i in range(n):
        for j in range(n):
            for sx, sy, sz in itertools.product(range(-k, k + 1), repeat=3):
                if i == j and sx == sy == sz == 0:
                    continue
                d = np.array([x[i] - x[j] - sx * L, y[i] - y[j] - sy * L, z[i] - z[j] - sz * L])
                r2 = float(d @ d)
                if r2 <= rc * rc:
                    f[i] += d * 24 * (2 / r2 ** 7 - 1 / r2 ** 4)
    return f


def test_direct_sum_small_box_against_python(oracle):
    # L = 3.3 < 2 rc: several images of the same neighbour interact
    rng = np.random.default_rng(3)
    L = 3.3
    x, y, z = _lattice(3, 1.1, jitter=0.05, seed=3)
    r = oracle.direct_sum(x, y, z, 2.5, box_len=(L, L, L), apply_shift=False)
    ref = _python_brute_force(x, y, z, L, 2.5)
    got = np.stack([r["fx"], r["fy"], r["fz"]], axis=1)
    assert np.allclose(got, ref, rtol=1e-11, atol=1e-11)
    del rng


@pytest.mark.parametrize("n_side,rho", [(8, 0.75), (6, 0.5), (4, 0.43)])
def test_force_algorithms_agree(oracle, n_side, rho):
    a = rho ** (-1 / 3)
    L = n_side * a
    x, y, z = _lattice(n_side, a, jitter=0.05, seed=n_side)
    d = oracle.direct_sum(x, y, z, 2.5, box_len=(L, L, L))
    lc = oracle.lc_forces(x, y, z, 2.5, 0.3, (0, 0, 0), (L, L, L))
    c8 = oracle.c08_forces_halo(x, y, z, 2.5, 0.3, (0, 0, 0), (L, L, L))
    scale = np.max(np.abs(d["fx"]))
    for r in (lc, c8):
        for k in ("fx", "fy", "fz"):
            assert np.max(np.abs(r[k] - d[k])) <= 1e-12 * scale
        assert r["pairs"] == d["pairs"]
        assert r["epot"] == pytest.approx(d["epot"], rel=1e-12)
        assert r["virial"] == pytest.approx(d["virial"], rel=1e-12)


def test_newton3_total_force_vanishes(oracle):
    x, y, z = _lattice(6, 1.1, jitter=0.1, seed=5)
    L = 6.6
    r = oracle.c08_forces_halo(x, y, z, 2.5, 0.3, (0, 0, 0), (L, L, L))
    for k in ("fx", "fy", "fz"):
        assert abs(r[k].sum()) < 1e-10


# ---- VerletClusterLists ------------------------------------------------------------------------------
def test_vcl_tower_grid_reference_geometry(oracle):
    g = GOLD["vcl_tower_grid"]
    rng = np.random.default_rng(1)
    n = g["num_particles"]
    x, y, z = (rng.uniform(0, 10, n) for _ in range(3))
    r = oracle.vcl_periodic(x, y, z, None, g["box_min"], g["box_max"], g["cutoff"] + g["skin"], g["cluster_size"])
    assert list(r["towers_per_dim"]) == g["owned_towers_per_dim"]
    assert list(r["side"]) == g["tower_side_length"]
    assert r["towers_per_il"] == g["towers_per_interaction_length"]


@pytest.mark.parametrize("M", [1, 2, 4, 8])
def test_vcl_lists_cover_every_pair_once(oracle, M):
    # every ordered pair (i, j, image) within rc must be covered by exactly one list entry of i's cluster
    n_side, rho = 6, 0.6
    a = rho ** (-1 / 3)
    L = n_side * a
    x, y, z = _lattice(n_side, a, jitter=0.2, seed=M)
    x, y, z = oracle.wrap(x, y, z, (0, 0, 0), (L, L, L))
    rc, ilen = 2.5, 2.8
    r = oracle.vcl_periodic(x, y, z, None, (0, 0, 0), (L, L, L), ilen, M)
    slot = r["slot_particle"]
    cluster_of = {int(p): s // M for s, p in enumerate(slot) if p >= 0}
    entries = {}
    for ci in range(len(r["cluster_tower"])):
        for e in range(r["list_offsets"][ci], r["list_offsets"][ci + 1]):
            key = (ci, int(r["list_j"][e]), int(r["list_shift"][e]))
            entries[key] = entries.get(key, 0) + 1
    assert all(v == 1 for v in entries.values())
    n = len(x)
    for i in range(n):
        for j in range(n):
            for sx, sy, sz in itertools.product((-1, 0, 1), repeat=3):
                if i == j and sx == sy == sz == 0:
                    continue
                d = (x[i] - x[j] - sx * L, y[i] - y[j] - sy * L, z[i] - z[j] - sz * L)
                if d[0] ** 2 + d[1] ** 2 + d[2] ** 2 <= rc * rc:
                    sid = (sx + 1) + 3 * (sy + 1) + 9 * (sz + 1)
                    assert (cluster_of[i], cluster_of[j], sid) in entries


@pytest.mark.parametrize("M", [1, 2, 4, 8])
@pytest.mark.parametrize("n_side,rho", [(6, 0.6), (3, 0.5)])
def test_vcl_newton3_lists_cover_every_pair_once(oracle, M, n_side, rho):
    # Newton-3 lists (forward neighbours only): every unordered interacting pair {(i, 0), (j, image)} within
    # rc is covered by exactly one entry, either (cluster i -> cluster j, image) or (cluster j -> cluster i,
    # -image), as VerletClusterListsTest.cpp:215-256 checks its N3 lists against the non-N3 ones; the entries
    # are the full lists' forward entries in list order.  The 3^3 box (L = 4.05 < 2 rc) has clusters that
    # meet their own images.
    a = rho ** (-1 / 3)
    L = n_side * a
    x, y, z = _lattice(n_side, a, jitter=0.2, seed=M)
    x, y, z = oracle.wrap(x, y, z, (0, 0, 0), (L, L, L))
    rc, ilen = 2.5, 2.8
    full = oracle.vcl_periodic(x, y, z, None, (0, 0, 0), (L, L, L), ilen, M)
    r = oracle.vcl_periodic(x, y, z, None, (0, 0, 0), (L, L, L), ilen, M, newton3=True)
    assert np.array_equal(r["slot_particle"], full["slot_particle"])
    nc = len(r["cluster_tower"])
    keep = []
    for ci in range(nc):
        for e in range(full["list_offsets"][ci], full["list_offsets"][ci + 1]):
            cj, sid = int(full["list_j"][e]), int(full["list_shift"][e])
            if cj > ci or (cj == ci and sid >= 13):
                keep.append((ci, cj, sid))
    got = [(ci, int(r["list_j"][e]), int(r["list_shift"][e]))
           for ci in range(nc) for e in range(r["list_offsets"][ci], r["list_offsets"][ci + 1])]
    assert got == keep
    entries = set(got)
    assert len(entries) == len(got)
    slot = r["slot_particle"]
    cluster_of = {int(p): s // M for s, p in enumerate(slot) if p >= 0}
    n = len(x)
    for i in range(n):
        for j in range(i, n):
            for sx, sy, sz in itertools.product((-1, 0, 1), repeat=3):
                if i == j and sx == sy == sz == 0:
                    continue
                d = (x[i] - x[j] - sx * L, y[i] - y[j] - sy * L, z[i] - z[j] - sz * L)
                if d[0] ** 2 + d[1] ** 2 + d[2] ** 2 > rc * rc:
                    continue
                sid = (sx + 1) + 3 * (sy + 1) + 9 * (sz + 1)
                ci, cj = cluster_of[i], cluster_of[j]
                fwd = (ci, cj, sid) in entries
                back = (cj, ci, 26 - sid) in entries
                if ci == cj and sid == 13:
                    assert fwd  # the cluster's own entry: its member pairs
                else:
                    assert fwd != back, (i, j, sid)


def test_verlet_pairs_fixed_match_fp64_within_ilen(oracle):
    n_side, rho = 7, 0.7
    a = rho ** (-1 / 3)
    L = n_side * a
    x, y, z = _lattice(n_side, a, jitter=0.2, seed=11)
    x, y, z = oracle.wrap(x, y, z, (0, 0, 0), (L, L, L))
    ilen = 2.8
    off, nb = oracle.verlet_pairs_fixed(x, y, z, (0, 0, 0), (L, L, L), ilen)
    n = len(x)
    for i in range(n):
        d = np.stack([x[i] - x, y[i] - y, z[i] - z])
        d -= L * np.round(d / L)
        r2 = (d * d).sum(axis=0)
        want = set(np.nonzero(r2 <= ilen * ilen)[0].tolist()) - {i}
        got = set(nb[off[i]:off[i + 1]].tolist())
        # fixed point differs from FP64 only for pairs within ~1e-7 of ilen
        diff = want ^ got
        assert all(abs(math.sqrt(r2[j]) - ilen) < 1e-6 for j in diff)


# ---- integrator and wrap -----------------------------------------------------------------------------
def test_time_discretization_reference_values(oracle):
    g = GOLD["time_discretization"]
    n = 8
    x, y, z = (np.zeros(n) for _ in range(3))
    vx, vy, vz = np.zeros(n), np.zeros(n), np.full(n, g["initial_velocity"][2], dtype=float)
    fx, fy, fz = np.zeros(n), np.zeros(n), np.full(n, g["initial_force"][2], dtype=float)
    m = np.array([g["mass"]])
    ofx, ofy, ofz = oracle.verlet_positions(x, y, z, vx, vy, vz, fx, fy, fz, m, None, g["dt"])
    assert np.allclose(z, g["position_shift_first_update_z"], rtol=0, atol=1e-15)
    assert not fz.any() and np.all(ofz == g["initial_force"][2])
    # velocities: first update with F = (0,0,1) and oldF = 0, then F = (0,0,2), oldF = (0,0,1)
    vz2 = np.full(n, 1.0)
    oracle.verlet_velocities(np.zeros(n), np.zeros(n), vz2, np.zeros(n), np.zeros(n), np.full(n, 1.0),
                             np.zeros(n), np.zeros(n), np.zeros(n), m, None, g["dt"])
    assert np.allclose(vz2, g["velocity_after_first_update_z"], rtol=0, atol=1e-15)
    oracle.verlet_velocities(np.zeros(n), np.zeros(n), vz2, np.zeros(n), np.zeros(n), np.full(n, 2.0),
                             np.zeros(n), np.zeros(n), np.full(n, 1.0), m, None, g["dt"])
    assert np.allclose(vz2, g["velocity_after_second_update_z"], rtol=0, atol=1e-15)
    # second position update with v = 0.5, F = 2
    z2 = np.zeros(n)
    oracle.verlet_positions(np.zeros(n), np.zeros(n), z2, np.zeros(n), np.zeros(n), np.full(n, 0.5), np.zeros(n),
                            np.zeros(n), np.full(n, 2.0), m, None, g["dt"])
    assert np.allclose(z2, g["position_shift_second_update_z"], rtol=0, atol=1e-15)


def test_wrap_periodic_edges(oracle):
    bmin, bmax = (0.0, 0.0, 0.0), (10.0, 10.0, 10.0)
    x = np.array([-1e-17, -0.5, 10.0, 10.5, 3.0])
    y = np.zeros_like(x)
    z = np.zeros_like(x)
    wx, _, _ = oracle.wrap(x, y, z, bmin, bmax)
    assert wx[0] == np.nextafter(10.0, 0.0)  # -1e-17 + 10 rounds to 10: clamped just inside
    assert wx[1] == 9.5 and wx[2] == 0.0 and wx[3] == 0.5 and wx[4] == 3.0
    assert np.all((wx >= 0.0) & (wx < 10.0))
