"""The launch-shape autotuner (parallellbmcranfield_amd.sweep.autotune) on the
GPU: both stages run, every shape it reports is the one that ran, and the
options it returns give the same fields as the library's defaults."""
import numpy as np
import pytest

import parallellbmcranfield_amd as lbm
from parallellbmcranfield_amd import sweep as SW

pytestmark = pytest.mark.gpu


def _opts(prec):
    return lbm.make_options(collision=lbm.MRT, uavg=0.05, rho_ini=1.0, viscosity=0.03,
                            precision=prec, init=lbm.INIT_REST, resid_kind=lbm.RESID_NONE,
                            use_stop_rule=0)


@pytest.mark.parametrize("prec", [lbm.PREC_F32, lbm.PREC_F16])
def test_autotune_times_launch_and_tile_shapes(prec):
    print("autotune", prec, flush=True)
    case = lbm.Case.cavity(513)
    best, tried = SW.autotune(case, _opts(prec), steps=36, blocks=(256, 512), groups=(1, 4),
                              two_rows=(4, 8), three_rows=(4, 8))
    stage1 = [r for r in tried if r["stage"] == 1]
    stage2 = [r for r in tried if r["stage"] == 2]
    # stage 1: every block x group, FP16 block 256 with and without pairs; single steps only
    assert len(stage1) == (6 if prec == lbm.PREC_F16 else 4)
    assert all(r["steps_per_launch"] == 1 for r in stage1)
    assert {r["pairs"] for r in stage1} == ({True, False} if prec == lbm.PREC_F16 else {False})
    # stage 2: the forced tile heights ran as asked, at the best stage-1 shape
    b1 = min(stage1, key=lambda r: r["ms_per_step"])
    assert sorted((r["steps_per_launch"], r["tile_rows"]) for r in stage2) == \
        [(2, 4), (2, 8), (3, 4), (3, 8)]
    assert all((r["block_size"], r["xcd_group"]) == (b1["block_size"], b1["xcd_group"])
               for r in stage2)
    assert best == min(tried, key=lambda r: r["ms_per_step"])
    assert all(r["ms_per_step"] > 0 for r in tried)

    # the picked options change the time, not the result
    fields = []
    for pick in (False, True):
        o = _opts(prec)
        if pick:
            o.block_size, o.xcd_group = best["block_size"], best["xcd_group"]
            o.two_step_rows, o.three_step_rows = best["two_step_rows"], best["three_step_rows"]
            o.tune_flags = best["tune_flags"]
        s = lbm.Solver(case, o)
        s.step(31)
        fields.append(s.macro())
        s.close()
    for a, b in zip(*fields):
        assert np.array_equal(a, b)


def test_autotune_on_a_packed_lattice_times_two_step_tiles():
    print("autotune batched", flush=True)
    specs = SW.random_cavities(6, seed=4, side=(65, 129))
    packed, _ = SW.build_batch(specs)
    inst = [(s["nu"], s["lid"]) for s in specs]
    o = SW.sweep_options()
    best, tried = SW.autotune(packed, o, instances=inst, steps=24, blocks=(256,), groups=(1, 4),
                              two_rows=(4, 8), three_rows=(8,))
    stage2 = [r for r in tried if r["stage"] == 2]
    assert sorted((r["steps_per_launch"], r["tile_rows"]) for r in stage2) == [(2, 4), (2, 8)]
    assert best == min(tried, key=lambda r: r["ms_per_step"])
