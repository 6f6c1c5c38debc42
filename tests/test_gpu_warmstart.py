"""Learned warm start end to end on the GPU, at a reduced scale: steady states
of seeded small cavities from the batched solver train the network with the
library's own kernels; on held-out cavities the predicted start reaches the
residual criterion in fewer steps than the rest start and converges to the
same field.
"""
import numpy as np
import pytest

import parallellbmcranfield_amd as lbm
from parallellbmcranfield_amd import sweep as SW
from parallellbmcranfield_amd import warmstart as WS

pytestmark = pytest.mark.gpu


def test_trained_warm_start_beats_rest_start_on_held_out_cavities():
    print("warm start: steady states", flush=True)
    specs = SW.random_cavities(76, seed=11, side=(24, 48), re=(100.0, 1000.0))
    train_specs, test_specs = specs[:64], specs[64:]
    opts = SW.sweep_options(crit=1e-6, resid_every=50, max_steps=150000)
    fields, it, _ = WS.steady_states(train_specs, opts)
    ok = it > 0
    assert ok.sum() >= 56
    G = 49
    X = np.stack([WS.encode(s, G) for s, k in zip(train_specs, ok) if k])
    T = np.stack([WS.field_to_grid(s, *f, G) for s, f, k in zip(train_specs, fields, ok) if k])
    print("warm start: training on", len(X), flush=True)
    net = lbm.Net(G, G, batch=8, base=32, levels=4, seed=0, in_channels=WS.IN_CH)
    hist = WS.train(net, X, T, epochs=160, lr=3e-3)
    print("loss", hist[0], hist[-1], flush=True)
    assert hist[-1] < 0.3 * hist[0]
    r = WS.compare_starts(net, test_specs, opts)
    print("steps rest", r["steps_rest"], "warm", r["steps_warm"], flush=True)
    print("median steps ratio", r["median_steps_ratio"], "field rel L2",
          r["field_rel_l2_median"], flush=True)
    assert r["both_converged"] >= 10
    # the gain at this scale (64 small training cavities) is modest; the
    # full-scale comparison is bench.py --workload warmstart
    assert r["median_steps_ratio"] > 1.08
    assert (r["steps_warm"] < r["steps_rest"]).mean() >= 0.6
    assert r["field_rel_l2_median"] < 2e-3
