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


def _fake_timer(calls):
    """Stands in for sweep._time_steps: a time per plain step from the shape
    (so the best shape is known in advance) and the multi-step layout the
    options would give (two-step tiles only where two_step_rows > 0 and
    three-step tiles only where three_step_rows > 0)."""
    def timer(case, o, instances, steps):
        calls.append((o.block_size, o.xcd_group, o.tune_flags, o.two_step_rows,
                      o.three_step_rows, instances is not None))
        no2 = bool(o.tune_flags & lbm.TUNE_NO_TWO_STEP)
        no3 = bool(o.tune_flags & lbm.TUNE_NO_THREE_STEP)
        r2 = 0 if no2 else o.two_step_rows
        r3 = 0 if no3 or instances is not None else o.three_step_rows  # single lattices only
        two = dict(cells=1000 if r2 else 0, tiles=4, rows=r2, band=10)
        three = dict(cells=1000 if r3 else 0, tiles=4, rows=r3, band=10)
        t = 1e-5 * (1 + o.block_size / 256) + 1e-7 * abs(o.xcd_group - 4)
        if o.tune_flags & lbm.TUNE_NO_PAIRS:
            t += 1e-6
        if r3:
            t *= 0.5 + 0.01 * abs(r3 - 8)
        elif r2:
            t *= 0.7 + 0.01 * abs(r2 - 6)
        return t, two, three
    return timer


def test_autotune_stages_on_a_single_lattice(monkeypatch):
    calls = []
    monkeypatch.setattr(SW, "_time_steps", _fake_timer(calls))
    o = lbm.make_options(collision=lbm.MRT, uavg=0.05, rho_ini=1.0, viscosity=0.03,
                         precision=lbm.PREC_F16, tune_flags=lbm.TUNE_NO_GRAPH)
    best, tried = SW.autotune(None, o, steps=12, blocks=(256, 512), groups=(1, 4),
                              two_rows=(4, 6), three_rows=(8, 16))
    stage1 = [r for r in tried if r["stage"] == 1]
    # FP16: block 256 with and without pairs, block 512 once; multi-step off in stage 1
    assert len(stage1) == 2 * 2 + 2
    single = lbm.TUNE_NO_TWO_STEP | lbm.TUNE_NO_THREE_STEP
    assert all(c[2] & single == single for c in calls[:6])
    assert all(r["tune_flags"] & lbm.TUNE_NO_GRAPH for r in tried)  # caller's flags kept
    b1 = min(stage1, key=lambda r: r["ms_per_step"])
    assert (b1["block_size"], b1["xcd_group"], b1["pairs"]) == (256, 4, True)
    stage2 = [r for r in tried if r["stage"] == 2]
    assert [(r["steps_per_launch"], r["tile_rows"]) for r in stage2] == \
        [(2, 4), (2, 6), (3, 8), (3, 16)]
    assert all((r["block_size"], r["xcd_group"]) == (256, 4) for r in stage2)
    assert all(not r["tune_flags"] & lbm.TUNE_NO_PAIRS for r in stage2)
    assert best["steps_per_launch"] == 3 and best["three_step_rows"] == 8
    assert best == min(tried, key=lambda r: r["ms_per_step"])
    assert o.tune_flags == lbm.TUNE_NO_GRAPH and o.block_size == 0  # caller's options untouched


def test_autotune_batched_lattices_take_two_step_tiles(monkeypatch):
    calls = []
    monkeypatch.setattr(SW, "_time_steps", _fake_timer(calls))
    o = lbm.make_options(collision=lbm.MRT, uavg=0.05, rho_ini=1.0, viscosity=0.03,
                         precision=lbm.PREC_F32)
    best, tried = SW.autotune(None, o, instances=[object()], steps=12, blocks=(256,),
                              groups=(1, 2, 4), two_rows=(4, 6), three_rows=(8,))
    assert [r["stage"] for r in tried] == [1, 1, 1, 2, 2]
    assert all(c[5] for c in calls)  # every solver built with the instances
    assert [(r["steps_per_launch"], r["tile_rows"]) for r in tried if r["stage"] == 2] == \
        [(2, 4), (2, 6)]  # no three-step variant is tried on a batched lattice
    assert best["steps_per_launch"] == 2 and best["two_step_rows"] == 6
    # stage 2 keeps a variant only if the library ran that many steps per launch
    calls.clear()
    best, tried = SW.autotune(None, o, steps=12, blocks=(256,), groups=(4,),
                              two_rows=(0,), three_rows=(0,))
    assert [r["stage"] for r in tried] == [1] and len(calls) == 3
