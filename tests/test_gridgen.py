"""Host grid producer (lsg_gridgen_*: AddPoint, AddRecursiveSonsOfPoint,
completeDirichletGrid and the refinement towards a corner) against grids the compiled
reference built (bit-exact index sets, slot order and active flags):

- the adaptive golden cases (refinement near the origin, or near the origin and
  (1, ..., 1), at d = 2 and 3) and the config-3 benchmark grids of sum level 7 to 11
  (tests/golden/grids, the reference's generator needs about 20 minutes on one core for
  level 11);
- 16 seeded random point sets completed by completeDirichletGrid (tests/golden/
  gridgen_sets.npz); in 14 of them addPoints2 adds hanging nodes, which the refined grids
  never need.
No GPU is used.
"""
import json
import os
import time

import numpy as np
import pytest

import loadsg_amd as L

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = json.load(open(os.path.join(GOLD, "cases.json")))
ADAPTIVE = sorted(c for c in CASES if CASES[c]["argv"][0].startswith("adaptive"))


def refined(argv, d):
    mode, start, radius, level = argv[0], int(argv[1]), float(argv[2]), int(argv[3])
    return L.GridGenerator.refined(d, start, radius, level, corners=2 if mode == "adaptive2" else 1)


@pytest.mark.parametrize("case", ADAPTIVE)
def test_adaptive_golden_cases(case):
    g = refined(CASES[case]["argv"], CASES[case]["d"])
    idx, act = g.points()
    assert np.array_equal(idx, np.load(os.path.join(GOLD, case, "idx.npy")))
    assert np.array_equal(act, np.load(os.path.join(GOLD, case, "active.npy")))


@pytest.mark.parametrize("level", [7, 8, 9, 10, 11])
def test_config3_grids(level):
    t0 = time.perf_counter()
    g = L.GridGenerator.refined(3, 4, 0.15, level)
    dt = time.perf_counter() - t0
    idx, act = g.points()
    z = np.load(os.path.join(GOLD, "grids", "cfg3_start4_level%d.npz" % level))
    assert np.array_equal(idx, z["idx"]) and np.array_equal(act, z["active"])
    print("level %d: %d points, %d hanging, %.2f s" % (level, len(act), int((act == 0).sum()), dt))


def test_random_point_sets():
    z = np.load(os.path.join(GOLD, "gridgen_sets.npz"))
    k = 0
    while "pts%d" % k in z:
        pts = z["pts%d" % k]
        g = L.GridGenerator(pts.shape[1])
        for p in pts:
            g.AddPoint(p)
        g.completeDirichletGrid()
        idx, act = g.points()
        assert np.array_equal(idx, z["idx%d" % k]), k
        assert np.array_equal(act, z["active%d" % k]), k
        k += 1
    assert k == 16


def test_generator_feeds_the_operators():
    """The produced grid is accepted as it is (slot order, hanging nodes) and equals the
    grid built from the reference's own index table."""
    g = L.GridGenerator.refined(3, 3, 0.15, 6)
    grid = g.grid()
    idx, act = grid.slot_order()
    ref = np.load(os.path.join(GOLD, "d3_adaptive_cfg3", "idx.npy"))
    assert np.array_equal(idx, ref) and (act == 0).sum() > 0


def test_boundary_points_rejected():
    g = L.GridGenerator(2)
    with pytest.raises(L.LoadsgError):
        g.AddPoint([1, 2])
    with pytest.raises(L.LoadsgError):
        g.AddRecursiveSonsOfPoint([2, 2], 15)
    g.AddRecursiveSonsOfPoint([2, 2], 2)
    assert g.size == 17   # depth sums up to 2 above the centre: 1 + 2 + 2 + 4 + 4 + 4
