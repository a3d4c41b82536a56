"""Ensemble workload generator (host side, no GPU): the box distribution and the concatenated layout."""
import math

import numpy as np

from autopas_amd import ensemble as en


def test_boxes_follow_the_workload_distribution():
    boxes = en.make_boxes(4096, seed=4)
    n = np.array([b.n for b in boxes])
    rho = np.array([b.density for b in boxes])
    T = np.array([b.temperature for b in boxes])
    L = np.array([b.length for b in boxes])
    assert n.min() >= 64 and n.max() <= 8192
    assert rho.min() >= 0.3 and rho.max() <= 0.9
    assert T.min() >= 0.7 and T.max() <= 1.5
    # log-uniform N: the median is near sqrt(64 * 8192) = 724
    assert 600 < np.median(n) < 870
    # about 4 % of the boxes are shorter than 2 rc and about 10 % shorter than 2 (rc + skin)
    assert 0.02 < np.mean(L < 2 * en.RC) < 0.06
    assert 0.07 < np.mean(L < 2 * (en.RC + en.SKIN)) < 0.12
    assert L.min() > en.RC + en.SKIN
    assert [b.seed for b in boxes[:3]] == [4000, 4001, 4002]
    again = en.make_boxes(4096, seed=4)
    assert [(b.n, b.density) for b in again] == [(b.n, b.density) for b in boxes]


def test_ensemble_arrays_concatenate_boxes_in_order():
    boxes = en.make_boxes(6, seed=9, n_range=(64, 400))
    lo, hi, counts, st = en.ensemble_arrays(boxes)
    assert lo.shape == hi.shape == (6, 3)
    assert list(counts) == [b.n for b in boxes]
    assert len(st["x"]) == counts.sum()
    for b, sl in zip(boxes, en.box_slices(counts)):
        one = en.box_system(b)
        for k in st:
            assert np.array_equal(st[k][sl], one[k]), k
        L = b.length
        for k in ("x", "y", "z"):
            assert np.all(st[k][sl] > -0.06) and np.all(st[k][sl] < L + 0.06)
        k3 = math.ceil(b.n ** (1 / 3) - 1e-9)
        assert k3 ** 3 >= b.n > (k3 - 1) ** 3
        p = [st[k][sl].sum() for k in ("vx", "vy", "vz")]
        assert np.allclose(p, 0.0, atol=1e-9 * b.n)
