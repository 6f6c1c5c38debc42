"""Pins of the warm-start network's CPU oracle (oracle/nn_oracle.py).

The network has no counterpart in the reference, so the oracle is pinned by
mathematics: central finite differences of the float64 loss for the
backward, adjoint identities for each layer's hand-written transpose, and
closed forms of the forward.
"""
import numpy as np
import pytest

from oracle import nn_oracle as NN


def _sample(N, H, W, seed):
    rng = np.random.default_rng(seed)
    inp = np.zeros((N, H, W, 3))
    inp[..., 0] = rng.random((N, H, W)) > 0.25
    inp[..., 1] = rng.random((N, 1, 1))
    inp[..., 2] = rng.random((N, 1, 1))
    tgt = rng.normal(size=(N, H, W, 3))
    return inp, tgt


def test_param_count_matches_the_stated_architecture():
    # 4 levels, 32 -> 256 channels: about two million parameters
    n = NN.num_params(4, 32)
    assert 1.9e6 < n < 2.1e6
    assert n == NN.init_params(4, 32).size


@pytest.mark.parametrize("shape", [(5, 6, 3), (3, 3, 2)])
def test_layer_adjoints(shape):
    rng = np.random.default_rng(1)
    H, W, C = shape
    x = rng.normal(size=(2, H, W, C))
    for d in (1, 2, 4):
        p = NN._patches(x, d)
        g = rng.normal(size=p.shape)
        assert np.isclose((p * g).sum(), (x * NN._col2im(g, d)).sum())
    y, cnt = NN.pool(x)
    g = rng.normal(size=y.shape)
    assert np.isclose((y * g).sum(), (x * NN.pool_bwd(g, cnt, H, W)).sum())
    xc = rng.normal(size=(2, (H + 1) // 2, (W + 1) // 2, C))
    g = rng.normal(size=(2, H, W, C))
    assert np.isclose((NN.up2(xc, H, W) * g).sum(), (xc * NN.up2_bwd(g, *xc.shape[1:3])).sum())


@pytest.mark.parametrize("d", [1, 2, 3])
def test_conv_matches_direct_sum(d):
    rng = np.random.default_rng(2)
    x = rng.normal(size=(1, 6, 7, 3))
    w = rng.normal(size=(2, 9, 3))
    b = rng.normal(size=2)
    y, _ = NN.conv(x, w, b, d)
    for yy in range(6):
        for xx in range(7):
            ref = b.copy()
            for t in range(9):
                dy, dx = d * (t // 3 - 1), d * (t % 3 - 1)
                if 0 <= yy + dy < 6 and 0 <= xx + dx < 7:
                    ref += w[:, t] @ x[0, yy + dy, xx + dx]
            assert np.allclose(y[0, yy, xx], ref)


def test_pool_of_odd_size_averages_the_cells_present():
    x = np.arange(3 * 3, dtype=float).reshape(1, 3, 3, 1)
    y, _ = NN.pool(x)
    assert np.allclose(y[0, :, :, 0], [[(0 + 1 + 3 + 4) / 4, (2 + 5) / 2], [(6 + 7) / 2, 8]])


def test_zero_convs_give_the_head_bias():
    p = NN.init_params(3, 32, seed=3)
    convs, (wh, bh) = NN.unpack(p, 3, 32)
    for w, b in convs:
        w[...] = 0
        b[...] = 0
    inp, _ = _sample(2, 7, 7, 0)
    y = NN.forward(p, inp, 3, 32)
    assert np.allclose(y, bh)


@pytest.mark.parametrize("dil", [1, 2])
def test_backward_matches_finite_differences(dil):
    L, c = 3, 32
    p = NN.init_params(L, c, seed=4)
    inp, tgt = _sample(2, 11, 10, 5)
    loss, g = NN.loss_and_grad(p, inp, tgt, L, c, dil)
    rng = np.random.default_rng(6)
    # coordinates from every layer, the padded input channels excluded
    offs, off = [], 0
    for i, (c0, c1, co, _) in enumerate(NN.layer_shapes(L, c)):
        ci = c0 + c1
        w = np.arange(co * 9 * ci).reshape(co, 9, ci)
        live = w[:, :, :3].ravel() if i == 0 else w.ravel()
        offs += list(off + rng.choice(live, 3, replace=False))
        offs += list(off + co * 9 * ci + rng.choice(co, 1))
        off += co * 9 * ci + co
    offs += list(off + rng.choice(NN.KOUT * c + NN.KOUT, 3, replace=False))
    eps = 1e-6
    for k in offs:
        q = p.copy()
        q[k] += eps
        lp = NN.loss_and_grad(q, inp, tgt, L, c, dil)[0]
        q[k] -= 2 * eps
        lm = NN.loss_and_grad(q, inp, tgt, L, c, dil)[0]
        fd = (lp - lm) / (2 * eps)
        assert abs(fd - g[k]) <= 1e-6 + 1e-4 * abs(fd), (k, fd, g[k])
    # the padded input channels carry no gradient
    w0 = g[:32 * 9 * 32].reshape(32, 9, 32)
    assert not w0[:, :, 3:].any()


def test_divergence_penalty_matches_finite_differences():
    # the adjoint of the discrete divergence, then the whole loss gradient
    rng = np.random.default_rng(8)
    y = rng.normal(size=(2, 6, 7, 3))
    mask = (rng.random((2, 6, 7, 1)) > 0.15).astype(float)
    d, adj = NN.div_penalty(y, mask, 0.7, 1.3)
    g = rng.normal(size=d.shape)
    eps = 1e-6
    num = np.zeros_like(y)
    for idx in np.ndindex(*y.shape):
        if idx[-1] == 2:
            continue
        yp = y.copy()
        yp[idx] += eps
        num[idx] = ((NN.div_penalty(yp, mask, 0.7, 1.3)[0] - d) * g).sum() / eps
    assert np.allclose(adj(g), num, atol=1e-6)
    # an interior point with all four neighbours fluid: the central difference
    m1 = np.ones((1, 5, 5, 1))
    yy = np.zeros((1, 5, 5, 3))
    yy[0, :, :, 0] = np.arange(5)[None, :] * 2.0   # du/dx = 2
    yy[0, :, :, 1] = np.arange(5)[:, None] * -0.5  # dv/dy = -0.5
    d1, _ = NN.div_penalty(yy, m1, 1.0, 2.0)
    assert np.allclose(d1[0, 1:-1, 1:-1], 2.0 - 1.0) and not d1[0, 0].any()
    L, c = 3, 32
    p = NN.init_params(L, c, seed=4)
    inp, tgt = _sample(2, 11, 10, 5)
    loss, gr = NN.loss_and_grad(p, inp, tgt, L, c, 1, div_w=5.0, su=0.2, sv=0.15)
    assert loss > NN.loss_and_grad(p, inp, tgt, L, c, 1)[0]
    rng = np.random.default_rng(9)
    n_head = NN.KOUT * c + NN.KOUT
    for k in list(p.size - n_head + rng.choice(n_head, 4, replace=False)) + list(rng.choice(2000, 3)):
        q = p.copy()
        q[k] += 1e-6
        lp = NN.loss_and_grad(q, inp, tgt, L, c, 1, div_w=5.0, su=0.2, sv=0.15)[0]
        q[k] -= 2e-6
        lm = NN.loss_and_grad(q, inp, tgt, L, c, 1, div_w=5.0, su=0.2, sv=0.15)[0]
        fd = (lp - lm) / 2e-6
        assert abs(fd - gr[k]) <= 1e-6 + 1e-4 * abs(fd), (k, fd, gr[k])
