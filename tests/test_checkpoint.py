"""md-flexible checkpoints (.pvtu index + per-rank .vtu pieces): the reader against a fixture in the
reference writer's layout (ParallelVtkWriter.cpp:60-200,309-358) and FP64 round trips of the writer."""
import os

import numpy as np
import pytest

import autopas_amd as ap

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "mdflex_checkpoint")


def test_reads_reference_layout_two_pieces():
    st, t, i = ap.read_checkpoint(os.path.join(GOLDEN, "demo_Particles_0042.pvtu"))
    # pieces concatenated in rank order
    assert list(i) == [7, 3, 12, 0, 99]
    assert list(t) == [0, 1, 0, 2, 1]
    assert list(st["x"]) == [1.1, 9.9999999, 5.0, 0.000123457, 7.75]
    assert list(st["y"]) == [2.2, 0.5, 5.0, 9.5, 1.25]
    assert list(st["z"]) == [3.3, 4.25, 5.0, 8.0, 2.0]
    assert st["vx"][1] == -1.23457e-05 and st["vz"][0] == 0.333333
    assert list(st["fx"]) == [1.5, 123457.0, 0.0, -1.0, 6.02214e23]
    assert st["fz"][4] == -6.02214e23 and st["fz"][1] == 3.14159


def test_round_trip_is_exact(tmp_path):
    rng = np.random.default_rng(5)
    n = 777
    st = {k: rng.normal(size=n) * 10.0 ** rng.integers(-12, 12, size=n) for k in ap.engine.STATE_KEYS}
    st["x"][0] = np.nextafter(10.0, 0.0)  # just below a box bound
    types = rng.integers(0, 4, size=n).astype(np.int32)
    ids = rng.permutation(n).astype(np.int64) * 1000003
    path = ap.write_checkpoint(tmp_path / "run", "lj", 1234, st, type=types, id=ids, digits=6)
    assert path.endswith("run/lj_Particles_001234.pvtu")
    assert os.path.exists(tmp_path / "run" / "data" / "lj_Particles_0_001234.vtu")
    back, t, i = ap.read_checkpoint(path)
    for k in ap.engine.STATE_KEYS:
        assert np.array_equal(back[k], st[k]), k
    assert np.array_equal(t, types) and np.array_equal(i, ids)


def test_written_piece_has_the_reference_layout(tmp_path):
    st = {k: np.arange(3, dtype=float) + j for j, k in enumerate(ap.engine.STATE_KEYS)}
    ap.write_checkpoint(tmp_path, "s", 7, st, digits=3)
    text = open(tmp_path / "data" / "s_Particles_0_007.vtu").read().splitlines()
    assert text[0] == '<?xml version="1.0" encoding="UTF-8" standalone="no" ?>'
    assert text[3] == '    <Piece NumberOfCells="0" NumberOfPoints="3">'
    names = [l.split('Name="')[1].split('"')[0] for l in text if "<DataArray" in l]
    assert names == ["velocities", "forces", "typeIds", "ids", "positions", "types"]
    assert text[6] == "        3 4 5"  # first velocity row: vx, vy, vz of particle 0
    index = open(tmp_path / "s_Particles_007.pvtu").read()
    assert '<Piece Source="./data/s_Particles_0_007.vtu"/>' in index


def test_missing_piece_is_an_error(tmp_path):
    st = {k: np.ones(2) for k in ap.engine.STATE_KEYS}
    path = ap.write_checkpoint(tmp_path, "m", 1, st)
    os.remove(tmp_path / "data" / "m_Particles_0_1.vtu")
    with pytest.raises(ap.EngineError):
        ap.read_checkpoint(path)
