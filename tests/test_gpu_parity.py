"""GPU parity: the HIP engine (through its C ABI) against the CPU oracle on the same seeded inputs.

Tolerances (stated per test):
  * cell assignment, tower/cluster layout, cluster-pair lists: bit-exact;
  * FP64 kernels: forces within 1e-10 relative (max |dF| / max |F|), energy/virial 1e-10, pair count exact;
  * FP32-pair / FP64-accumulate kernel: force RMS relative error <= 1e-5 (sqrt(sum|dF|^2 / sum|F|^2)),
    energy/virial 1e-5 relative, pair count exact (the pairs near rc are decided in FP64).
"""
import numpy as np
import pytest

import autopas_amd as ap
from conftest import announce

pytestmark = pytest.mark.gpu

RC = 2.5
SKIN = 0.3


def lattice_box(n_side, density, seed, jitter=0.05, temperature=0.9, ntypes=1):
    a = density ** (-1.0 / 3.0)
    L = n_side * a
    n = n_side ** 3
    x, y, z, vx, vy, vz, t = ap.generate_lattice(seed, n, (n_side,) * 3, a, (0.0, 0.0, 0.0), jitter, temperature,
                                                 ntypes)
    return dict(x=x, y=y, z=z, vx=vx, vy=vy, vz=vz, type=t, L=L, n=n)


def rel_max(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def rms_rel(f, fr):
    num = sum(np.sum((f[k] - fr[k]) ** 2) for k in ("fx", "fy", "fz"))
    den = sum(np.sum(fr[k] ** 2) for k in ("fx", "fy", "fz"))
    return float(np.sqrt(num / den))


def make_engine(L, M=4, rc=RC, skin=SKIN, **kw):
    return ap.Engine((0.0, 0.0, 0.0), (L, L, L), rc, skin, cluster_size=M, **kw)


def test_binning_bit_exact(oracle):
    announce("binning bit-exact: 5000 random + boundary positions, box 20^3, ilen 2.8")
    rng = np.random.default_rng(11)
    L = 20.0
    x, y, z = (rng.uniform(0, L, 5000) for _ in range(3))
    # boundary cases of CellBlock3DTest: at min, just below max, exactly on cell faces
    edge = np.array([0.0, np.nextafter(L, 0.0), L / 7 * 3, 2.8, 5.6, 1e-300, L - 1e-12])
    x[: len(edge)] = edge
    y[len(edge): 2 * len(edge)] = edge
    z[2 * len(edge): 3 * len(edge)] = edge
    with make_engine(L) as e:
        e.upload(x, y, z)
        e.bin()
        cells, nh = e.cells()
    wx, wy, wz = oracle.wrap(x, y, z, (0, 0, 0), (L, L, L))
    cb = oracle.cellblock((0, 0, 0), (L, L, L), RC + SKIN)
    ref = oracle.cell_index_batch(cb, wx, wy, wz)
    assert list(nh) == list(cb.cells_per_dim_halo)
    assert np.array_equal(cells, ref)


@pytest.mark.parametrize("list_mode", [1, 2, 3])
@pytest.mark.parametrize("M", [1, 2, 4, 8])
def test_cluster_lists_bit_exact(oracle, M, list_mode):
    # list_mode 1: a wavefront per tower, 2: a lane per i-cluster, 3: 32 lanes per i-cluster (one per
    # neighbour-tower visit; towers of ~20 particles here, so the
    # towers are sorted in registers)
    announce(f"cluster build bit-exact M={M} list_mode={list_mode}: 10^3 lattice rho 0.75")
    s = lattice_box(10, 0.75, seed=7)
    L = s["L"]
    ids = np.arange(s["n"], dtype=np.int64)[::-1].copy()  # ids differ from upload index
    with make_engine(L, M=M, list_mode=list_mode) as e:
        e.upload(s["x"], s["y"], s["z"], id=ids)
        e.build_pairs()
        g = e.clusters()
    wx, wy, wz = oracle.wrap(s["x"], s["y"], s["z"], (0, 0, 0), (L, L, L))
    r = oracle.vcl_periodic(wx, wy, wz, ids, (0, 0, 0), (L, L, L), RC + SKIN, M)
    assert list(g["towers_per_dim"]) == list(r["towers_per_dim"])
    assert g["towers_per_il"] == r["towers_per_il"]
    assert np.array_equal(g["side"], r["side"])
    for k in ("slot_particle", "cluster_tower", "list_offsets", "list_j", "list_shift", "bbox"):
        assert np.array_equal(g[k], r[k]), k


@pytest.mark.parametrize("list_mode", [1, 2, 3])
@pytest.mark.parametrize("M,n_side", [(1, 10), (2, 10), (4, 10), (8, 10), (4, 4), (8, 4)])
def test_cluster_lists_newton3_bit_exact(oracle, M, n_side, list_mode):
    # Newton-3 (forward-only) cluster-pair lists of the N3 kernels against the oracle's, entry by entry, in
    # every list kernel; n_side 4 (L = 4.4 < 2 rc) has clusters meeting their own images
    announce(f"Newton-3 cluster lists bit-exact M={M} list_mode={list_mode}: {n_side}^3 lattice rho 0.75")
    s = lattice_box(n_side, 0.75, seed=7)
    L = s["L"]
    ids = np.arange(s["n"], dtype=np.int64)[::-1].copy()
    with make_engine(L, M=M, list_mode=list_mode) as e:
        e.upload(s["x"], s["y"], s["z"], id=ids)
        e.rebuild(ap.KERNEL_CLUSTER_MIXED_N3)
        g = e.clusters()
    wx, wy, wz = oracle.wrap(s["x"], s["y"], s["z"], (0, 0, 0), (L, L, L))
    r = oracle.vcl_periodic(wx, wy, wz, ids, (0, 0, 0), (L, L, L), RC + SKIN, M, newton3=True)
    for k in ("slot_particle", "cluster_tower", "list_offsets", "list_j", "list_shift", "bbox"):
        assert np.array_equal(g[k], r[k]), k


MIXED = (ap.KERNEL_CLUSTER_MIXED, ap.KERNEL_VERLET_MIXED, ap.KERNEL_CLUSTER_JROWS_MIXED, ap.KERNEL_CLUSTER_MIXED_N3)
ALL_KERNELS = [ap.KERNEL_LC_FP64, ap.KERNEL_CLUSTER_FP64, ap.KERNEL_CLUSTER_MIXED, ap.KERNEL_VERLET_MIXED,
               ap.KERNEL_CLUSTER_JROWS_MIXED, ap.KERNEL_CLUSTER_FP64_N3, ap.KERNEL_CLUSTER_MIXED_N3]


def test_verlet_lists_exact(oracle):
    announce("verlet lists exact: 10^3 lattice rho 0.75, ilen 2.8")
    s = lattice_box(10, 0.75, seed=8)
    L = s["L"]
    with make_engine(L) as e:
        e.upload(s["x"], s["y"], s["z"])
        e.force(ap.KERNEL_VERLET_MIXED, globals=False)
        off, nb = e.verlet_lists()
    wx, wy, wz = oracle.wrap(s["x"], s["y"], s["z"], (0, 0, 0), (L, L, L))
    roff, rnb = oracle.verlet_pairs_fixed(wx, wy, wz, (0, 0, 0), (L, L, L), RC + SKIN)
    assert np.array_equal(off, roff)
    for p in range(s["n"]):
        assert np.array_equal(np.sort(nb[off[p]:off[p + 1]]), rnb[roff[p]:roff[p + 1]]), p


@pytest.mark.parametrize("list_mode", [1, 2, 3])
@pytest.mark.parametrize("nz,sorter", [(40, "registers, up to 256"), (100, "registers, up to 512"), (150, "LDS")])
def test_cluster_lists_bit_exact_long_towers(oracle, nz, sorter, list_mode):
    # 20 x 20 x nz lattice at rho 0.75 with M = 8: 10 x 10 towers of ~4 nz particles, so each tower sorter
    # (registers for up to 64, 256 and 512 particles, LDS above) gets the towers of one of these boxes
    announce(f"cluster build bit-exact M=8 list_mode={list_mode}: 20x20x{nz} lattice rho 0.75 ({sorter})")
    a = 0.75 ** (-1.0 / 3.0)
    dims = (20, 20, nz)
    n = 20 * 20 * nz
    x, y, z, _, _, _, _ = ap.generate_lattice(17, n, dims, a, (0.0, 0.0, 0.0), 0.05, 0.9)
    box = tuple(d * a for d in dims)
    ids = np.arange(n, dtype=np.int64)
    with ap.Engine((0.0, 0.0, 0.0), box, RC, SKIN, cluster_size=8, list_mode=list_mode) as e:
        e.upload(x, y, z, id=ids)
        e.build_pairs()
        g = e.clusters()
    per_tower = n / (g["towers_per_dim"][0] * g["towers_per_dim"][1])
    assert {40: 64 < per_tower <= 256, 100: 256 < per_tower <= 448, 150: per_tower > 448}[nz], per_tower
    wx, wy, wz = oracle.wrap(x, y, z, (0, 0, 0), box)
    r = oracle.vcl_periodic(wx, wy, wz, ids, (0, 0, 0), box, RC + SKIN, 8)
    for k in ("slot_particle", "cluster_tower", "list_offsets", "list_j", "list_shift", "bbox"):
        assert np.array_equal(g[k], r[k]), k


def test_jrow_rebuild_after_vetoed_force_is_bit_identical():
    # The force after a j-row build is launched before the host reads the list lengths, with a device-side
    # veto on rows longer than their capacity; on a veto the step loop builds again and launches again.
    # With no spare row capacity (row_headroom -1) the rows of a melting lattice outgrow the lengths kept
    # from the previous build, so the vetoed path runs (more builds are counted); the trajectory must
    # equal the one with the default headroom bit for bit.
    kernel = ap.KERNEL_CLUSTER_JROWS_MIXED
    announce("j-row force after a vetoed launch: 16^3 lattice rho 0.75, 200 steps, row_headroom -1 vs 0")
    s = lattice_box(16, 0.75, seed=31, jitter=0.0, temperature=1.5)
    L = s["L"]
    out = {}
    for hr in (0, -1):
        with make_engine(L, M=8, row_headroom=hr) as e:
            e.upload(s["x"], s["y"], s["z"], s["vx"], s["vy"], s["vz"])
            rs = e.step(200, 0.002, kernel)
            out[hr] = (e.download(), rs.rebuilds, rs.pair_interactions)
    (a, ra, pa), (b, rb, pb) = out[0], out[-1]
    print(f"builds: headroom 25% {ra}, none {rb}; pair interactions {pa} {pb}")
    assert rb > ra
    assert pa == pb
    for k in ("x", "y", "z", "vx", "vy", "vz", "fx", "fy", "fz"):
        assert np.array_equal(a[k], b[k]), k


@pytest.mark.parametrize("nz", [12, 40, 100])
def test_tower_sort_ties_in_fp32_z(oracle, nz):
    # The register tower sorts order by z rounded to FP32 and settle equal FP32 z by the exact (z, id).
    # An unjittered lattice holds whole planes of equal z; half of its particles are then moved up by
    # 2^-40 of their z (the same FP32 z, a different FP64 z), and the ids are shuffled, so that the id
    # order disagrees with the z order inside every tie.  nz 12, 40, 100: towers for each register sort.
    announce(f"tower sort ties in FP32 z, M=8: 20x20x{nz} unjittered lattice rho 0.75, shuffled ids")
    a = 0.75 ** (-1.0 / 3.0)
    dims = (20, 20, nz)
    n = 20 * 20 * nz
    x, y, z0, _, _, _, _ = ap.generate_lattice(5, n, dims, a, (0.0, 0.0, 0.0), 0.0, 0.9)
    rng = np.random.default_rng(nz)
    up = rng.random(n) < 0.5
    z = np.where(up, z0 + z0 * 2.0 ** -40, z0)
    moved = up & (z != z0)
    assert np.count_nonzero(moved & (z.astype(np.float32) == z0.astype(np.float32))) > n // 4
    box = tuple(d * a for d in dims)
    ids = rng.permutation(n).astype(np.int64)
    with ap.Engine((0.0, 0.0, 0.0), box, RC, SKIN, cluster_size=8) as e:
        e.upload(x, y, z, id=ids)
        e.build_pairs()
        g = e.clusters()
    wx, wy, wz = oracle.wrap(x, y, z, (0, 0, 0), box)
    r = oracle.vcl_periodic(wx, wy, wz, ids, (0, 0, 0), box, RC + SKIN, 8)
    for k in ("slot_particle", "cluster_tower", "bbox"):
        assert np.array_equal(g[k], r[k]), k


@pytest.mark.parametrize("M,n_side", [(1, 10), (2, 10), (4, 10), (8, 10), (4, 6), (8, 6)])
def test_cluster_jrows_exact(oracle, M, n_side):
    # Rows of the j-particle cluster kernel = union of the members' fixed-point Verlet lists (members
    # included), each j once.  n_side 6 (L = 6.6, just above 2 ilen): the tower window wraps onto itself
    # and clusters pass the box test under several images, which the build must not expand twice.
    announce(f"cluster j-rows exact M={M}: {n_side}^3 lattice rho 0.75")
    s = lattice_box(n_side, 0.75, seed=13)
    L = s["L"]
    ids = np.arange(s["n"], dtype=np.int64)
    with make_engine(L, M=M) as e:
        e.upload(s["x"], s["y"], s["z"], id=ids)
        e.force(ap.KERNEL_CLUSTER_JROWS_MIXED, globals=False)
        g = e.clusters()
    wx, wy, wz = oracle.wrap(s["x"], s["y"], s["z"], (0, 0, 0), (L, L, L))
    r = oracle.vcl_periodic(wx, wy, wz, ids, (0, 0, 0), (L, L, L), RC + SKIN, M)
    for k in ("slot_particle", "cluster_tower", "bbox"):
        assert np.array_equal(g[k], r[k]), k
    roff, rnb = oracle.verlet_pairs_fixed(wx, wy, wz, (0, 0, 0), (L, L, L), RC + SKIN)
    sp, off, lj = g["slot_particle"], g["list_offsets"], g["list_j"]
    assert np.all(g["list_shift"] == -1)
    for ci in range(len(off) - 1):
        members = [int(p) for p in sp[ci * M:(ci + 1) * M] if p >= 0]
        expect = set(members)
        for p in members:
            expect.update(int(q) for q in rnb[roff[p]:roff[p + 1]])
        row = lj[off[ci]:off[ci + 1]]
        assert len(row) == len(set(row.tolist())), ci
        assert set(row.tolist()) == expect, ci


@pytest.mark.parametrize("M,n_side", [(4, 8), (8, 8), (8, 5)])
def test_member_rows_are_the_ilen_members(M, n_side):
    # The FP32 cluster kernel's member rows: per i-cluster exactly the members of its listed j-clusters
    # that lie within ilen (FP64, with the entry's periodic image) of one of its real members, up to the
    # FP32 test's slack (2^-16 relative on ilen^2).  n_side 5 (L = 5.5 < 2 ilen): entries of one j-cluster
    # under several images.
    announce(f"member rows M={M}: {n_side}^3 lattice rho 0.75")
    s = lattice_box(n_side, 0.75, seed=21)
    L = s["L"]
    with make_engine(L, M=M) as e:
        e.upload(s["x"], s["y"], s["z"])
        e.force(ap.KERNEL_CLUSTER_MIXED, globals=False)
        g = e.clusters()
        off, rows = e.member_rows()
        st = e.download()
    P = np.stack([st["x"], st["y"], st["z"]], axis=1)
    sp, loff, lj, ls = g["slot_particle"], g["list_offsets"], g["list_j"], g["list_shift"]
    il2 = (RC + SKIN) ** 2
    assert len(off) == len(loff)
    for ci in range(len(loff) - 1):
        mem = np.array([p for p in sp[ci * M:(ci + 1) * M] if p >= 0])
        row = rows[off[ci]:off[ci + 1]].tolist()
        assert len(row) == len(set(row)), ci
        must, may = set(), set()
        for k in range(int(loff[ci + 1] - loff[ci])):
            J, sid = int(lj[loff[ci] + k]), int(ls[loff[ci] + k])
            shift = np.array([sid % 3 - 1, (sid // 3) % 3 - 1, sid // 9 - 1]) * L
            for m in range(M):
                pj = int(sp[J * M + m])
                if pj < 0:
                    continue
                d2 = float(np.min(np.sum((P[mem] - (P[pj] + shift)) ** 2, axis=1)))
                if d2 <= il2:
                    must.add((k << 3) | m)
                if d2 <= il2 * (1 + 2.0 ** -15):
                    may.add((k << 3) | m)
        assert must <= set(row) <= may, ci


def test_member_rows_forces_match_whole_rows():
    # The member rows leave out only pairs beyond rc + skin: forces, energy, virial and the pair count of
    # the FP32 cluster kernel are those of the whole entries, up to the order of the FP32 pair sums
    announce("member rows vs whole cluster rows: 12^3 lattice rho 0.75, 20 steps")
    s = lattice_box(12, 0.75, seed=5)
    L = s["L"]
    res = {}
    for whole in (True, False):
        with make_engine(L, M=8, tile_groups=4, whole_cluster_rows=whole) as e:
            e.upload(s["x"], s["y"], s["z"], s["vx"], s["vy"], s["vz"])
            g = e.force(ap.KERNEL_CLUSTER_MIXED, globals=True)
            f = e.download()
            rs = e.step(20, 0.002, ap.KERNEL_CLUSTER_MIXED)
            res[whole] = (g, f, rs, e.download())
    (g1, f1, r1, t1), (g0, f0, r0, t0) = res[True], res[False]
    assert g1.pair_count == g0.pair_count
    assert rms_rel(f0, f1) <= 1e-6
    assert abs(g0.potential_energy - g1.potential_energy) <= 1e-6 * abs(g1.potential_energy)
    assert abs(g0.virial - g1.virial) <= 1e-6 * abs(g1.virial)
    # (the two trajectories part by FP32 rounding: a pair crossing rc may be counted in one step only)
    assert abs(r0.pair_interactions - r1.pair_interactions) <= 1e-6 * r1.pair_interactions
    assert abs(r0.rebuilds - r1.rebuilds) <= 1
    assert max(np.max(np.abs(t0[k] - t1[k])) for k in ("x", "y", "z")) <= 1e-6


@pytest.mark.parametrize("kernel", ALL_KERNELS)
@pytest.mark.parametrize("M", [4, 8])
def test_forces_vs_direct_sum(oracle, kernel, M):
    if kernel in (ap.KERNEL_LC_FP64, ap.KERNEL_VERLET_MIXED) and M == 8:
        pytest.skip("cluster size does not apply to this kernel")
    announce(f"forces vs direct sum kernel={ap.KERNEL_NAMES[kernel]} M={M}: 8^3 lattice")
    s = lattice_box(8, 0.75, seed=3)
    L = s["L"]
    with make_engine(L, M=M) as e:
        e.upload(s["x"], s["y"], s["z"])
        g = e.force(kernel, globals=True)
        f = e.download()
    ref = oracle.direct_sum(f["x"], f["y"], f["z"], RC, box_len=(L, L, L))
    if kernel in MIXED:
        assert rms_rel(f, ref) <= 1e-5
        assert abs(g.potential_energy - ref["epot"]) <= 1e-5 * abs(ref["epot"])
        assert abs(g.virial - ref["virial"]) <= 1e-5 * abs(ref["virial"])
        assert g.pair_count == ref["pairs"]  # every FP32-pair kernel decides the pairs near rc in FP64
    else:
        for k in ("fx", "fy", "fz"):
            assert rel_max(f[k], ref[k]) <= 1e-10, k
        assert abs(g.potential_energy - ref["epot"]) <= 1e-10 * abs(ref["epot"])
        assert abs(g.virial - ref["virial"]) <= 1e-10 * abs(ref["virial"])
        assert g.pair_count == ref["pairs"]


@pytest.mark.parametrize("kernel", [ap.KERNEL_VERLET_MIXED, ap.KERNEL_CLUSTER_JROWS_MIXED])
def test_fixed_point_kernels_reject_small_box(kernel):
    # L = 5.29: below 3 ilen (Verlet lists) and below 2 ilen (j-particle rows)
    announce(f"kernel={ap.KERNEL_NAMES[kernel]} on a box too short for minimum-image lists: rejected")
    s = lattice_box(4, 0.43, seed=5, jitter=0.1)
    with make_engine(s["L"]) as e:
        e.upload(s["x"], s["y"], s["z"])
        with pytest.raises(ap.EngineError) as ei:
            e.force(kernel)
        assert ei.value.status == 5
        with pytest.raises(ap.EngineError) as ei:
            e.step(2, 0.002, kernel)
        assert ei.value.status == 5


@pytest.mark.parametrize("kernel", [ap.KERNEL_LC_FP64, ap.KERNEL_CLUSTER_FP64, ap.KERNEL_CLUSTER_MIXED,
                                    ap.KERNEL_CLUSTER_FP64_N3, ap.KERNEL_CLUSTER_MIXED_N3])
def test_small_box_multiple_images(oracle, kernel):
    # L = 4.31 < 2 * (rc + skin): a particle sees several images of the same neighbour.
    announce(f"small box images kernel={ap.KERNEL_NAMES[kernel]}: 4^3 lattice rho 0.75, L=5.29")
    s = lattice_box(4, 0.43, seed=5, jitter=0.1)
    L = s["L"]
    assert L < 2 * (RC + SKIN)
    with make_engine(L, M=4) as e:
        e.upload(s["x"], s["y"], s["z"])
        g = e.force(kernel, globals=True)
        f = e.download()
    ref = oracle.direct_sum(f["x"], f["y"], f["z"], RC, box_len=(L, L, L))
    tol = 1e-5 if kernel in MIXED else 1e-10
    assert rms_rel(f, ref) <= tol
    assert g.pair_count == ref["pairs"]


@pytest.mark.parametrize("kernel", ALL_KERNELS)
def test_four_type_mixing(oracle, kernel):
    announce(f"4-type LB mixing kernel={ap.KERNEL_NAMES[kernel]}: rc 3, 10^3 lattice")
    eps = [0.8, 1.0, 1.1, 1.2]
    sig = [0.9, 1.0, 1.05, 1.1]
    s = lattice_box(10, 0.7, seed=9, ntypes=4)
    L = s["L"]
    with make_engine(L, M=4, rc=3.0) as e:
        e.set_types(eps, sig, [1.0] * 4)
        e.upload(s["x"], s["y"], s["z"], type=s["type"])
        g = e.force(kernel, globals=True)
        f = e.download()
    ref = oracle.direct_sum(f["x"], f["y"], f["z"], 3.0, box_len=(L, L, L), types=s["type"], eps=eps, sig=sig)
    tol = 1e-5 if kernel in MIXED else 1e-10
    assert rms_rel(f, ref) <= tol
    assert abs(g.potential_energy - ref["epot"]) <= tol * abs(ref["epot"])


@pytest.mark.parametrize("mixing", [False, True])
@pytest.mark.parametrize("kernel", [ap.KERNEL_LC_FP64, ap.KERNEL_CLUSTER_FP64, ap.KERNEL_CLUSTER_MIXED])
def test_known_answer_two_particles(kernel, mixing):
    # LJFunctorTestNoGlobals.h:27-28: p1 = (0,0,0), p2 = (0.1,0.2,0.3), rc = 1, eps = sigma = 1
    # (mixing: p2 has eps = sigma = 2).
    announce(f"known answer two particles kernel={ap.KERNEL_NAMES[kernel]} mixing={mixing}")
    expected = np.array([-4547248.8989645941, -9094497.7979291882, -13641746.696893783])
    expected_mix = np.array([-835415983.7676939964294, -1670831967.5353879928588, -2506247951.3030819892883])
    e = ap.Engine((-1.0, -1.0, -1.0), (4.0, 4.0, 4.0), 1.0, 0.3, cluster_size=2, apply_shift=False)
    with e:
        if mixing:
            e.set_types([1.0, 2.0], [1.0, 2.0], [1.0, 1.0])
        e.upload(np.array([0.0, 0.1]), np.array([0.0, 0.2]), np.array([0.0, 0.3]),
                 type=np.array([0, 1 if mixing else 0], dtype=np.int32))
        e.force(kernel, globals=False)
        f = e.download()
    exp = expected_mix if mixing else expected
    got1 = np.array([f["fx"][0], f["fy"][0], f["fz"][0]])
    got2 = np.array([f["fx"][1], f["fy"][1], f["fz"][1]])
    if kernel == ap.KERNEL_CLUSTER_MIXED:
        assert np.allclose(got1, exp, rtol=1e-5, atol=0)
        assert np.allclose(got2, -exp, rtol=1e-5, atol=0)
    else:
        assert np.allclose(got1, exp, rtol=0, atol=1e-7 * max(1.0, np.abs(exp).max() / 1e7))
        assert np.allclose(got2, -exp, rtol=0, atol=1e-7 * max(1.0, np.abs(exp).max() / 1e7))


@pytest.mark.parametrize("kernel", [ap.KERNEL_LC_FP64, ap.KERNEL_CLUSTER_FP64, ap.KERNEL_CLUSTER_MIXED])
@pytest.mark.parametrize("spacing,cutoff,fexp", [(1.0, 1.1, 24.0), (2 ** (1 / 6), 1.3, 0.0)])
def test_force_calculation_grid(kernel, spacing, cutoff, fexp):
    # ForceCalculationTest.cpp:43-86: 2x2x1 grid (offset spacing/2... here spacing/2), eps24 = 24, sigma 1.
    announce(f"2x2x1 grid kernel={ap.KERNEL_NAMES[kernel]} spacing={spacing:.4f} rc={cutoff}")
    L = 3.0 * max(spacing, 1.0) + cutoff
    e = ap.Engine((0.0, 0.0, 0.0), (L, L, L), cutoff, 0.2, cluster_size=2, apply_shift=False)
    pos = np.array([[0.5, 0.5, 0.5], [0.5 + spacing, 0.5, 0.5], [0.5, 0.5 + spacing, 0.5],
                    [0.5 + spacing, 0.5 + spacing, 0.5]])
    with e:
        e.upload(pos[:, 0], pos[:, 1], pos[:, 2])
        e.force(kernel, globals=False)
        f = e.download()
    exp = np.array([[-1, -1, 0], [1, -1, 0], [-1, 1, 0], [1, 1, 0]]) * fexp
    got = np.stack([f["fx"], f["fy"], f["fz"]], axis=1)
    tol = 1e-13 if kernel != ap.KERNEL_CLUSTER_MIXED else 1e-4
    assert np.allclose(got, exp, rtol=0, atol=tol)


def test_md_trajectory_lc_fp64(oracle):
    announce("md trajectory LC FP64: 6^3 lattice, 40 steps dt 0.002, rebuild 10")
    s = lattice_box(6, 0.75, seed=4, temperature=0.9)
    L = s["L"]
    dt = 0.002
    with make_engine(L, M=4, rebuild_frequency=10) as e:
        e.upload(s["x"], s["y"], s["z"], s["vx"], s["vy"], s["vz"])
        e.step(40, dt, ap.KERNEL_LC_FP64)
        g = e.download()
    x, y, z = s["x"].copy(), s["y"].copy(), s["z"].copy()
    vx, vy, vz = s["vx"].copy(), s["vy"].copy(), s["vz"].copy()
    fx, fy, fz = np.zeros_like(x), np.zeros_like(x), np.zeros_like(x)
    m = np.array([1.0])
    for _ in range(40):
        ofx, ofy, ofz = oracle.verlet_positions(x, y, z, vx, vy, vz, fx, fy, fz, m, None, dt)
        x, y, z = oracle.wrap(x, y, z, (0, 0, 0), (L, L, L))
        r = oracle.direct_sum(x, y, z, RC, box_len=(L, L, L))
        fx, fy, fz = r["fx"], r["fy"], r["fz"]
        oracle.verlet_velocities(vx, vy, vz, fx, fy, fz, ofx, ofy, ofz, m, None, dt)
    # the GPU wraps only at rebuilds; compare minimum-image displacement
    d = np.stack([g["x"] - x, g["y"] - y, g["z"] - z])
    d -= L * np.round(d / L)
    assert np.max(np.abs(d)) < 1e-9
    assert np.max(np.abs(g["vx"] - vx)) < 1e-8


@pytest.mark.parametrize("kernel", [ap.KERNEL_LC_FP64, ap.KERNEL_CLUSTER_FP64, ap.KERNEL_CLUSTER_JROWS_MIXED,
                                    ap.KERNEL_CLUSTER_FP64_N3, ap.KERNEL_CLUSTER_MIXED_N3])
def test_dynamic_rebuild_trajectory(oracle, kernel):
    # Hot system, no periodic rebuilds: every rebuild comes from the skin/2 displacement trigger
    # (LogicHandler.h:1000-1015), whose step launches its force before the host has seen the trigger and
    # must skip that launch.  Positions follow the oracle's direct-sum trajectory and the pair count
    # summed over the steps is the oracle's exactly (skipped launches count nothing).
    announce(f"dynamic rebuild trajectory kernel={ap.KERNEL_NAMES[kernel]}: 6^3 lattice T 4, 60 steps dt 0.004")
    s = lattice_box(6, 0.75, seed=21, temperature=4.0)
    L = s["L"]
    dt = 0.004
    with make_engine(L, M=4, rebuild_frequency=100000, dynamic_rebuild=True) as e:
        e.upload(s["x"], s["y"], s["z"], s["vx"], s["vy"], s["vz"])
        rs = e.step(60, dt, kernel)
        g = e.download()
    x, y, z = s["x"].copy(), s["y"].copy(), s["z"].copy()
    vx, vy, vz = s["vx"].copy(), s["vy"].copy(), s["vz"].copy()
    fx, fy, fz = np.zeros_like(x), np.zeros_like(x), np.zeros_like(x)
    m = np.array([1.0])
    pairs = 0
    for _ in range(60):
        ofx, ofy, ofz = oracle.verlet_positions(x, y, z, vx, vy, vz, fx, fy, fz, m, None, dt)
        x, y, z = oracle.wrap(x, y, z, (0, 0, 0), (L, L, L))
        r = oracle.direct_sum(x, y, z, RC, box_len=(L, L, L))
        fx, fy, fz = r["fx"], r["fy"], r["fz"]
        pairs += r["pairs"]
        oracle.verlet_velocities(vx, vy, vz, fx, fy, fz, ofx, ofy, ofz, m, None, dt)
    assert rs.rebuilds >= 3  # the initial build plus displacement-triggered ones
    d = np.stack([g["x"] - x, g["y"] - y, g["z"] - z])
    d -= L * np.round(d / L)
    if kernel in (ap.KERNEL_CLUSTER_JROWS_MIXED, ap.KERNEL_CLUSTER_MIXED_N3):
        # FP32 pair term (the j-row kernel also runs the velocity-Verlet update in its epilogue)
        assert abs(rs.pair_interactions - pairs) <= 1e-4 * pairs
        assert np.max(np.abs(d)) < 2e-5
        assert np.max(np.abs(g["vx"] - vx)) < 1e-3
    else:
        # FP64 (the Newton-3 kernel adds its pair forces atomically, in no fixed order: a few ulps)
        assert rs.pair_interactions == pairs
        assert np.max(np.abs(d)) < 1e-8
        assert np.max(np.abs(g["vx"] - vx)) < 1e-8


@pytest.mark.parametrize("n_side", [8, 16])
@pytest.mark.parametrize("kernel", [ap.KERNEL_CLUSTER_MIXED, ap.KERNEL_VERLET_MIXED, ap.KERNEL_CLUSTER_JROWS_MIXED])
def test_md_trajectory_mixed_kernels(kernel, n_side):
    # 30 velocity-Verlet steps with rebuilds every 10 steps: the FP32-pair kernels (positions refreshed by
    # the drift in their own formats) follow the FP64 trajectory to within the FP32 force error.  The j-row
    # kernel takes its exact per-pair variant in the 8^3 box and its first-member-frame variant in 16^3.
    announce(f"md trajectory kernel={ap.KERNEL_NAMES[kernel]} vs LC FP64: {n_side}^3 lattice, 30 steps")
    s = lattice_box(n_side, 0.75, seed=4, temperature=0.9)
    L = s["L"]
    out = {}
    for k in (ap.KERNEL_LC_FP64, kernel):
        with make_engine(L, M=4, rebuild_frequency=10) as e:
            e.upload(s["x"], s["y"], s["z"], s["vx"], s["vy"], s["vz"])
            rs = e.step(30, 0.002, k)
            out[k] = (e.download(), rs)
    a, b = out[ap.KERNEL_LC_FP64][0], out[kernel][0]
    d = np.stack([b["x"] - a["x"], b["y"] - a["y"], b["z"] - a["z"]])
    d -= L * np.round(d / L)
    # close encounters carry forces ~1e3, whose FP32 error moves a particle by up to a few 1e-6
    assert np.max(np.abs(d)) < 2e-5
    assert np.max(np.abs(b["vx"] - a["vx"])) < 1e-3
    assert out[kernel][1].rebuilds == out[ap.KERNEL_LC_FP64][1].rebuilds


@pytest.mark.parametrize("kernel", [ap.KERNEL_LC_FP64, ap.KERNEL_CLUSTER_MIXED, ap.KERNEL_VERLET_MIXED,
                                    ap.KERNEL_CLUSTER_JROWS_MIXED])
def test_deterministic(kernel):
    announce(f"determinism kernel={ap.KERNEL_NAMES[kernel]}: two identical runs")
    s = lattice_box(8, 0.75, seed=12)
    outs = []
    for _ in range(2):
        with make_engine(s["L"], M=4) as e:
            e.upload(s["x"], s["y"], s["z"], s["vx"], s["vy"], s["vz"])
            e.step(25, 0.002, kernel)
            outs.append(e.download())
    for k in outs[0]:
        assert np.array_equal(outs[0][k], outs[1][k]), k


def test_mixed_precision_accuracy_c2_subsample(oracle):
    # 32^3 particles at rho 0.75 (seed 2): the subsample box on which the FP32-pair kernel's force error is
    # quoted against FP64.
    announce("mixed precision accuracy: 32^3 lattice rho 0.75 seed 2")
    s = lattice_box(32, 0.75, seed=2, temperature=0.9)
    L = s["L"]
    res = {}
    for kernel in (ap.KERNEL_CLUSTER_MIXED, ap.KERNEL_CLUSTER_FP64, ap.KERNEL_LC_FP64, ap.KERNEL_VERLET_MIXED,
                   ap.KERNEL_CLUSTER_JROWS_MIXED):
        with make_engine(L, M=4) as e:
            e.upload(s["x"], s["y"], s["z"])
            g = e.force(kernel, globals=True)
            res[kernel] = (e.download(), g)
    f = res[ap.KERNEL_CLUSTER_MIXED][0]
    ref = oracle.lc_forces(f["x"], f["y"], f["z"], RC, SKIN, (0, 0, 0), (L, L, L), threads=8)
    assert rms_rel(res[ap.KERNEL_CLUSTER_FP64][0], ref) <= 1e-12
    assert rms_rel(res[ap.KERNEL_LC_FP64][0], ref) <= 1e-12
    err = rms_rel(f, ref)
    errv = rms_rel(res[ap.KERNEL_VERLET_MIXED][0], ref)
    errj = rms_rel(res[ap.KERNEL_CLUSTER_JROWS_MIXED][0], ref)
    print(f"force RMS relative error vs FP64: cluster FP32-pair {err:.3e}, verlet FP32-pair {errv:.3e}, "
          f"cluster j-rows FP32-pair {errj:.3e}")
    assert err <= 1e-5
    assert errv <= 1e-5
    assert errj <= 1e-5
    assert res[ap.KERNEL_LC_FP64][1].pair_count == ref["pairs"]
    assert res[ap.KERNEL_CLUSTER_FP64][1].pair_count == ref["pairs"]


@pytest.mark.parametrize("kernel,M,G", [(ap.KERNEL_CLUSTER_JROWS_MIXED, 4, 8), (ap.KERNEL_CLUSTER_JROWS_MIXED, 8, 8),
                                        (ap.KERNEL_VERLET_MIXED, 1, 0), (ap.KERNEL_CLUSTER_MIXED, 4, 8),
                                        (ap.KERNEL_CLUSTER_MIXED, 8, 4), (ap.KERNEL_CLUSTER_MIXED, 8, 1)])
def test_mixed_precision_accuracy_full_liquid(kernel, M, G):
    # 10^6 particles (100^3 lattice, rho 0.75, seed 2): the benchmark system itself.  The FP64 cluster kernel
    # is the reference here (it matches the oracle to 1e-12 on the smaller boxes above).  The FP32-pair
    # kernels decide the pairs near rc in FP64, so at this size, where FP32 cutoff decisions would
    # otherwise add ~1e-5, they stay well within 1e-5 force RMS relative error and count exactly the FP64
    # kernel's pairs; energy and virial within 1e-5.
    announce(f"mixed precision accuracy: 100^3 lattice rho 0.75 seed 2, {ap.KERNEL_NAMES[kernel]} M={M} G={G}")
    s = lattice_box(100, 0.75, seed=2)
    L = s["L"]
    res = {}
    for k, m, g in ((kernel, M, G), (ap.KERNEL_CLUSTER_FP64, 4, 0)):
        with make_engine(L, M=m, tile_groups=g) as e:
            e.upload(s["x"], s["y"], s["z"])
            gl = e.force(k, globals=True)
            f = e.download()
            res[k] = (f, gl)
    (f, g), (fr, gr) = res[kernel], res[ap.KERNEL_CLUSTER_FP64]
    err = rms_rel(f, fr)
    print(f"force RMS relative error vs FP64 on 10^6 particles ({ap.KERNEL_NAMES[kernel]}): {err:.3e}")
    assert err <= 5e-6
    assert g.pair_count == gr.pair_count
    assert abs(g.potential_energy - gr.potential_energy) <= 1e-5 * abs(gr.potential_energy)
    assert abs(g.virial - gr.virial) <= 1e-5 * abs(gr.virial)


@pytest.mark.parametrize("kernel", [ap.KERNEL_CLUSTER_JROWS_MIXED, ap.KERNEL_VERLET_MIXED, ap.KERNEL_CLUSTER_FP64,
                                    ap.KERNEL_CLUSTER_MIXED])
@pytest.mark.parametrize("M", [4, 8])
def test_cutoff_decision_matches_fp64(kernel, M):
    # Isolated pairs at r = rc (1 + eps), |eps| from 3e-8 to 3e-7: below the FP32 resolution of r^2 (~1.2e-7
    # relative) at the small end, above the fixed-point position quantum (L / 2^32 = 9.3e-9 sigma here) at
    # all.  The kernels that decide the pairs near rc in FP64 keep exactly the pairs an FP64 test keeps;
    # a pair decided the other way would leave the whole force F(rc) ~ 0.039 on both particles.
    if kernel == ap.KERNEL_VERLET_MIXED and M == 8:
        pytest.skip("cluster size does not apply to this kernel")
    announce(f"cutoff decision vs FP64: pairs at rc (1 +- 3e-8 .. 3e-7), {ap.KERNEL_NAMES[kernel]} M={M}")
    rng = np.random.default_rng(5)
    L = 40.0
    eps = np.array([-3e-7, -1e-7, -3e-8, 3e-8, 1e-7, 3e-7] * 4)
    pos = []
    for k, e in enumerate(eps):
        base = np.array([4.0 + 8.0 * (k % 4), 4.0 + 8.0 * ((k // 4) % 4), 6.0 + 12.0 * (k // 16)])
        u = rng.normal(size=3)
        u /= np.linalg.norm(u)
        pos += [base, base + RC * (1.0 + e) * u]
    x, y, z = (np.array([p[d] for p in pos]) for d in range(3))
    d = np.stack([x[1::2] - x[0::2], y[1::2] - y[0::2], z[1::2] - z[0::2]], 1)
    expect = int(np.sum(np.sum(d * d, 1) <= RC * RC))
    assert expect == int(np.sum(eps < 0))
    with make_engine(L, M=M) as e:
        e.upload(x, y, z)
        g = e.force(kernel, globals=True)
        f = e.download()
    # FP64 pair forces of the kept pairs (LJFunctor.h:288-320 with sigma = epsilon = 1)
    r2 = np.sum(d * d, 1)
    inv = 1.0 / r2
    l6 = inv ** 3
    fac = np.where(r2 <= RC * RC, 24.0 * (l6 * l6 * 2.0 - l6) * inv, 0.0)
    fi = -d * fac[:, None]  # force on the first particle of each pair
    ref = np.zeros((len(x), 3))
    ref[0::2], ref[1::2] = fi, -fi
    assert g.pair_count == expect
    got = np.stack([f["fx"], f["fy"], f["fz"]], 1)
    assert np.max(np.abs(got - ref)) <= 1e-5 * np.max(np.abs(ref))


@pytest.mark.parametrize("kernel", [ap.KERNEL_CLUSTER_JROWS_MIXED, ap.KERNEL_CLUSTER_MIXED, ap.KERNEL_LC_FP64])
def test_graph_captured_steps_match_stream_launches(kernel):
    # ljg_params.graphs: each chunk of steps between rebuilds runs as one captured hipGraph (updated in place
    # from chunk to chunk); the trajectory, pair count and rebuild steps are those of the plain stream
    # launches, bit for bit, with the displacement trigger active (hot system)
    announce(f"graph-captured step chunks vs stream launches, kernel={ap.KERNEL_NAMES[kernel]}: 12^3 lattice T 3")
    s = lattice_box(12, 0.75, seed=23, temperature=3.0)
    outs = []
    for graphs in (False, True):
        with make_engine(s["L"], M=4, rebuild_frequency=10, dynamic_rebuild=True, graphs=graphs) as e:
            e.upload(s["x"], s["y"], s["z"], s["vx"], s["vy"], s["vz"])
            rs = e.step(75, 0.002, kernel)
            outs.append((rs.pair_interactions, rs.rebuilds, e.download()))
    assert outs[0][0] == outs[1][0] and outs[0][1] == outs[1][1]
    for k in outs[0][2]:
        assert np.array_equal(outs[0][2][k], outs[1][2][k]), k


def test_checkpoint_restart_continues_the_trajectory(tmp_path):
    # 30 steps in one go vs 15 steps, checkpoint, a fresh context restored from it, 15 more steps (FP64)
    announce("checkpoint restart: 6^3 lattice, 15 + 15 steps vs 30 steps, LC FP64")
    s = lattice_box(6, 0.75, seed=17, temperature=0.9)
    L = s["L"]
    with make_engine(L, M=4, rebuild_frequency=5) as e:
        e.upload(s["x"], s["y"], s["z"], s["vx"], s["vy"], s["vz"])
        e.step(30, 0.002, ap.KERNEL_LC_FP64)
        ref = e.download()
    with make_engine(L, M=4, rebuild_frequency=5) as e:
        e.upload(s["x"], s["y"], s["z"], s["vx"], s["vy"], s["vz"])
        e.step(15, 0.002, ap.KERNEL_LC_FP64)
        mid = e.download()
        path = e.write_checkpoint(tmp_path, "lj", 15, digits=4)
    back, _, ids = ap.read_checkpoint(path)
    for k in mid:
        assert np.array_equal(back[k], mid[k]), k
    assert np.array_equal(ids, np.arange(len(ids)))
    with make_engine(L, M=4, rebuild_frequency=5) as e:
        assert e.load_checkpoint(path) == s["n"]
        e.step(15, 0.002, ap.KERNEL_LC_FP64)
        out = e.download()
    d = np.stack([out["x"] - ref["x"], out["y"] - ref["y"], out["z"] - ref["z"]])
    d -= L * np.round(d / L)
    assert np.max(np.abs(d)) < 1e-10
    assert np.max(np.abs(out["vx"] - ref["vx"])) < 1e-9


def test_auto_layout_tuning():
    # KERNEL_AUTO: a tuning phase samples every candidate that fits the box for tuning_samples steps of
    # the real run, picks the lowest score and keeps it; a second phase starts after tuning_interval.
    announce("layout tuning: 20^3 lattice, 8 candidates x (1 + 4) steps, interval 10")
    s = lattice_box(20, 0.75, seed=6, temperature=0.9)
    L = s["L"]
    with make_engine(L, M=4, rebuild_frequency=4, tuning_samples=4, tuning_interval=10) as e:
        e.upload(s["x"], s["y"], s["z"], s["vx"], s["vy"], s["vz"])
        t0 = e.tuning()
        assert t0["best"] == -1 and t0["phases"] == 0
        ncand = len(t0["candidates"])
        rs = e.step(5 * ncand, 0.002, ap.KERNEL_AUTO)  # exactly one phase: every candidate 1 unscored + 4 steps
        assert rs.steps == 5 * ncand
        t1 = e.tuning()
        n = len(t1["candidates"])
        assert not t1["tuning"] and t1["phases"] == 1 and 0 <= t1["best"] < n
        scores = [c["effective_ms"] for c in t1["candidates"]]
        assert all(x > 0 for x in scores)  # the 22-sigma box fits every candidate
        assert scores[t1["best"]] == min(scores)
        rs = e.step(16, 0.002, ap.KERNEL_AUTO)  # 10 steps with the choice reach the interval, then a new phase
        assert rs.steps == 16
        assert e.tuning()["phases"] == 2 and e.tuning()["tuning"]
        st = e.download()
    # the run stays on the physical trajectory: compare with the FP64 linked-cells kernel over as many steps
    with make_engine(L, M=4, rebuild_frequency=4) as e:
        e.upload(s["x"], s["y"], s["z"], s["vx"], s["vy"], s["vz"])
        e.step(5 * ncand + 16, 0.002, ap.KERNEL_LC_FP64)
        ref = e.download()
    d = np.stack([st["x"] - ref["x"], st["y"] - ref["y"], st["z"] - ref["z"]])
    d -= L * np.round(d / L)
    assert np.max(np.abs(d)) < 2e-5


@pytest.mark.parametrize("kernel", [ap.KERNEL_CLUSTER_FP64, ap.KERNEL_CLUSTER_MIXED, ap.KERNEL_CLUSTER_JROWS_MIXED,
                                    ap.KERNEL_LC_FP64])
def test_update_positions_keeps_lists(oracle, kernel):
    # positions moved by the caller between two force evaluations (< skin/2): the kept lists still give
    # the forces of the moved system; a move of skin/2 is reported; the shift switch changes the energy
    # by the pair count times the shift (LJFunctor.h:174-197)
    announce(f"update positions, kernel={ap.KERNEL_NAMES[kernel]}: 8^3 lattice")
    s = lattice_box(8, 0.75, seed=21)
    L = s["L"]
    rng = np.random.default_rng(4)
    with make_engine(L) as e:
        e.upload(s["x"], s["y"], s["z"])
        e.rebuild(kernel)
        moved = {k: s[k] + rng.uniform(-0.05, 0.05, s["n"]) for k in ("x", "y", "z")}
        assert not e.update_positions(moved["x"], moved["y"], moved["z"])
        g = e.force(kernel, globals=True)
        f = e.download()
        e.set_shift(False)
        g0 = e.force(kernel, globals=True)
        far = dict(moved)
        far["x"] = far["x"].copy()
        far["x"][7] += 0.3  # at least 0.25 from its rebuild position, beyond skin/2 = 0.15
        assert e.update_positions(far["x"], far["y"], far["z"])
    ref = oracle.direct_sum(moved["x"], moved["y"], moved["z"], RC, box_len=(L, L, L))
    tol = 1e-10 if kernel in (ap.KERNEL_CLUSTER_FP64, ap.KERNEL_LC_FP64) else 1e-5
    assert rms_rel(f, ref) <= tol
    assert abs(g.potential_energy - ref["epot"]) <= tol * abs(ref["epot"])
    rc2 = RC * RC
    shift = 4.0 * ((1.0 / rc2) ** 3 - (1.0 / rc2) ** 6)  # shift6 / 6 for eps = sigma = 1
    assert abs((g.potential_energy - g0.potential_energy) - g.pair_count * shift) <= 1e-6 * abs(g.potential_energy)


def test_c1_plain_method_trajectory(oracle):
    # The plain-method configuration at full size: 32^3 simple-cubic lattice, spacing 1.1 (periodic 35.2),
    # jitter +-0.05, T 0.7, rc 2.5, skin 0.3, lists rebuilt every 10 steps, 100 velocity-Verlet steps at
    # dt 0.001 with the FP64 linked-cells kernel, against the oracle's linked-cells forces and integrator
    # (FP64 both, summed in other orders: positions within 1e-9, velocities within 1e-8, pair counts equal)
    announce("C1 plain method: 32^3 lattice a 1.1, 100 steps dt 0.001, LC FP64 vs oracle")
    n_side, a, dt, steps = 32, 1.1, 0.001, 100
    L = n_side * a
    n = n_side ** 3
    x, y, z, vx, vy, vz, _ = ap.generate_lattice(1, n, (n_side,) * 3, a, (0.0, 0.0, 0.0), 0.05, 0.7)
    with make_engine(L, rebuild_frequency=10) as e:
        e.upload(x, y, z, vx, vy, vz)
        rs = e.step(steps, dt, ap.KERNEL_LC_FP64)
        g = e.download()
    x, y, z, vx, vy, vz = (a_.copy() for a_ in (x, y, z, vx, vy, vz))
    fx, fy, fz = np.zeros(n), np.zeros(n), np.zeros(n)
    m = np.array([1.0])
    pairs = 0
    for _ in range(steps):
        ofx, ofy, ofz = oracle.verlet_positions(x, y, z, vx, vy, vz, fx, fy, fz, m, None, dt)
        x, y, z = oracle.wrap(x, y, z, (0, 0, 0), (L, L, L))
        r = oracle.lc_forces(x, y, z, RC, SKIN, (0, 0, 0), (L, L, L))
        fx, fy, fz = r["fx"], r["fy"], r["fz"]
        pairs += r["pairs"]
        oracle.verlet_velocities(vx, vy, vz, fx, fy, fz, ofx, ofy, ofz, m, None, dt)
    assert rs.pair_interactions == pairs
    d = np.stack([g["x"] - x, g["y"] - y, g["z"] - z])
    d -= L * np.round(d / L)
    assert np.max(np.abs(d)) < 1e-9
    for k, v in (("vx", vx), ("vy", vy), ("vz", vz)):
        assert np.max(np.abs(g[k] - v)) < 1e-8, k


def total_energy_per_particle(e, n):
    g = e.force(ap.KERNEL_CLUSTER_FP64, globals=True)
    st = e.download()
    return (0.5 * np.sum(st["vx"] ** 2 + st["vy"] ** 2 + st["vz"] ** 2) + g.potential_energy) / n


@pytest.mark.parametrize("kernel,M,G", [(ap.KERNEL_CLUSTER_JROWS_MIXED, 8, 8), (ap.KERNEL_CLUSTER_MIXED, 8, 4)])
def test_energy_drift_fp32_pair_matches_fp64(kernel, M, G):
    # energy drift over 1000 steps of the FP32-pair / FP64-accumulate kernels within 2x of the FP64 run's
    # (20^3 liquid at rho 0.75, T 0.9, dt 0.002; energies evaluated with the FP64 kernel)
    announce(f"energy drift kernel={ap.KERNEL_NAMES[kernel]} vs FP64: 20^3 liquid, 1000 steps")
    s = lattice_box(20, 0.75, seed=2, temperature=0.9)
    drifts = {}
    for k, m_, g_ in ((ap.KERNEL_CLUSTER_FP64, M, 0), (kernel, M, G)):
        with make_engine(s["L"], M=m_, tile_groups=g_, rebuild_frequency=10) as e:
            e.upload(s["x"], s["y"], s["z"], s["vx"], s["vy"], s["vz"])
            e0 = total_energy_per_particle(e, s["n"])
            e.step(1000, 0.002, k)
            drifts[k] = total_energy_per_particle(e, s["n"]) - e0
    d64, d32 = abs(drifts[ap.KERNEL_CLUSTER_FP64]), abs(drifts[kernel])
    assert d32 <= 2.0 * d64 + 1e-6, (d32, d64)
