"""Host side of the batched cavity sweep (no kernel launches)."""
import math

import numpy as np

import parallellbmcranfield_amd as lbm
from parallellbmcranfield_amd import sweep as SW


def test_random_cavities_follow_the_stated_distributions():
    specs = SW.random_cavities(400, seed=3)
    assert specs == SW.random_cavities(400, seed=3)
    nx = np.array([s["NX"] for s in specs])
    ny = np.array([s["NY"] for s in specs])
    re = np.array([s["Re"] for s in specs])
    assert nx.min() >= 65 and nx.max() <= 513 and ny.min() >= 65 and ny.max() <= 513
    assert 0.45 < (nx / ny).min() and (nx / ny).max() < 2.1
    assert re.min() >= 100 and re.max() <= 3200
    # log-uniform: the median sits near the geometric mean
    assert abs(math.log(np.median(re)) - 0.5 * math.log(100 * 3200)) < 0.3
    for s in specs:
        assert abs(s["nu"] - s["lid"] * s["NX"] / s["Re"]) < 1e-12
        assert len(s["rects"]) <= 3
        for i0, j0, i1, j1 in s["rects"]:
            assert 3 <= i0 <= i1 <= s["NX"] - 2 and 3 <= j0 <= j1 <= s["NY"] - 3
    counts = np.bincount([len(s["rects"]) for s in specs], minlength=4)
    assert (counts > 50).all()


def test_build_batch_counts_fluid_nodes():
    specs = SW.random_cavities(6, seed=1, side=(20, 40))
    packed, fl = SW.build_batch(specs)
    assert packed.instances == 6
    assert sum(fl) == int(packed.cells()["fluid"].sum())
    for s, f in zip(specs, fl):
        blocked = np.zeros((s["NY"], s["NX"]), bool)
        for i0, j0, i1, j1 in s["rects"]:
            blocked[j0 - 1:j1, i0 - 1:i1] = True
        assert f == s["NX"] * s["NY"] - blocked.sum()


def test_sweep_options_rates():
    o = SW.sweep_options()
    assert o.collision == lbm.MRT and o.precision == lbm.PREC_F32 and o.mrt_custom_rates == 1
    assert list(o.mrt_rates)[:7] == [0.0, 1.4, 1.4, 0.0, 1.2, 0.0, 1.2]
    r = SW.mrt_rates(0.128)
    assert abs(1 / r[7] - 0.884) < 1e-6  # tau = 3 nu + 1/2 (the 129^2 Re = 100 cavity)
