"""Host side of the learned warm start: grid encoding and resampling (no
kernel launches)."""
import math

import numpy as np
import pytest

from parallellbmcranfield_amd import warmstart as WS


def _spec(NX, NY, rects=(), Re=400.0, lid=0.1):
    return dict(NX=NX, NY=NY, rects=list(rects), Re=Re, nu=lid * NX / Re, lid=lid)


def test_encode_channels_and_mask():
    s = _spec(129, 65, [(10, 10, 20, 30)], Re=1000.0)
    x = WS.encode(s, 257)
    assert x.shape == (257, 257, WS.IN_CH)
    assert np.allclose(x[..., 1], (math.log(1000) - math.log(100)) / math.log(50))
    assert np.allclose(x[..., 2], 1.0)
    # scale 2: the cavity covers grid columns 0..256 and rows 0..128
    m = x[..., 0]
    assert m[:129].sum() > 0 and m[129:].sum() == 0
    assert m[0, 0] == 1 and m[128, 256] == 1
    # obstacle nodes (i, j) = (10..20, 10..30) -> grid (2(i-1), 2(j-1))
    assert m[2 * 15, 2 * 15] == 0 and m[2 * 15, 2 * 40] == 1
    fluid = 129 * 65 - 11 * 21
    assert abs(m.sum() / 4 - fluid) < 0.05 * fluid
    # wall distances over the longer side (128 node spacings): west + east = 1,
    # south + north = 64 / 128 inside the cavity, 0 outside
    assert np.allclose(x[:129, :, 3] + x[:129, :, 4], 1.0)
    assert np.allclose(x[:129, :, 5] + x[:129, :, 6], 0.5)
    assert x[0, 0, 3] == 0 and x[128, 256, 4] == 0 and x[128, 100, 6] == 0
    assert np.isclose(x[100, 50, 3], 25 / 128) and np.isclose(x[100, 50, 5], 50 / 128)
    assert not x[129:, :, 3:].any()
    # distance to the nearest solid node: 1 node next to the west wall, 0 in
    # the obstacle; at node (64, 32) the south and north walls are 33 away
    # (obstacle 45, west/east 65); at node (100, 32) the east wall, 29 away
    assert np.isclose(x[100, 0, 7], 1 / 128)
    assert x[2 * 15, 2 * 15, 7] == 0 and x[2 * 15, 2 * 40, 7] > 0
    assert np.isclose(x[2 * 32, 2 * 64, 7], 33 / 128)
    assert np.isclose(x[2 * 32, 2 * 100, 7], 29 / 128)
    # next to the obstacle's east face (node i = 20 is solid, 1-based)
    assert np.isclose(x[2 * 15, 2 * 20, 7], 1 / 128)


def test_resampling_round_trip_is_exact_on_grid_aligned_nodes():
    # NX = 257: one grid step per node, so every node sits on a grid point
    s = _spec(257, 129, [(100, 40, 120, 60)])
    n, m = 259, 132
    rng = np.random.default_rng(0)
    yy, xx = np.meshgrid(np.arange(m), np.arange(n), indexing="ij")
    u = (0.05 * np.sin(xx / 20.0) * np.cos(yy / 15.0)).ravel()
    v = (0.03 * np.cos(xx / 17.0)).ravel()
    rho = (1 + 1e-3 * rng.standard_normal(n * m))
    t = WS.field_to_grid(s, rho, u, v, 257)
    r2, u2, v2 = WS.grid_to_field(s, t, 257)
    fl = np.zeros((m, n), bool)
    fl[1:130, 1:258] = WS._fluid_mask(s)
    fl = fl.ravel()
    # float32 targets: exact to single precision
    assert np.allclose(u2[fl], u[fl], rtol=0, atol=1e-7)
    assert np.allclose(v2[fl], v[fl], rtol=0, atol=1e-7)
    assert np.allclose(r2[fl], rho[fl], rtol=0, atol=1e-7)
    assert (u2[~fl] == 0).all() and (r2[~fl] == 1).all()


def test_resampling_of_a_smooth_field_at_another_scale():
    s = _spec(90, 140)
    n, m = 92, 143
    yy, xx = np.meshgrid(np.arange(m), np.arange(n), indexing="ij")
    u = (0.1 * (yy / m) ** 2).ravel()
    t = WS.field_to_grid(s, np.ones(n * m), u, np.zeros(n * m), 257)
    _, u2, _ = WS.grid_to_field(s, t, 257)
    inner = np.zeros((m, n), bool)
    inner[1:141, 1:91] = True
    err = np.abs(u2 - u)[inner.ravel()].max()
    # interpolation twice at 0.54-node spacing, within one grid cell of the edges
    assert err < 2e-3


def test_channel_scales_are_masked_rms():
    inp = np.zeros((2, 4, 4, WS.IN_CH), np.float32)
    inp[0, :2, :, 0] = 1  # 8 fluid points in sample 0, none in sample 1
    t = np.zeros((2, 4, 4, 3), np.float32)
    t[0, :2, :, 0] = 3.0
    t[0, :2, :, 1] = np.array([1.0, -1.0, 1.0, -1.0])
    t[0, :2, :, 2] = 0.5
    t[1] = 100.0  # outside the fluid: ignored
    s = WS.channel_scales(inp, t)
    assert np.allclose(s, [3.0, 1.0, 0.5])


def test_separable_bilinear_equals_the_pointwise_formula():
    rng = np.random.default_rng(3)
    f = rng.normal(size=(9, 7, 3))
    gy = np.concatenate([rng.uniform(-1, 9, 11), [0.0, 8.0]])
    gx = np.concatenate([rng.uniform(-1, 7, 5), [0.0, 6.0, 2.5]])
    out = WS._bilinear(f, gy, gx)
    assert out.shape == (gy.size, gx.size, 3)
    for a, y in enumerate(gy):
        for b, x in enumerate(gx):
            y_, x_ = min(max(y, 0), 8), min(max(x, 0), 6)
            y0, x0 = int(np.floor(y_)), int(np.floor(x_))
            y1, x1 = min(y0 + 1, 8), min(x0 + 1, 6)
            ty, tx = y_ - y0, x_ - x0
            ref = ((1 - ty) * (1 - tx) * f[y0, x0] + (1 - ty) * tx * f[y0, x1] +
                   ty * (1 - tx) * f[y1, x0] + ty * tx * f[y1, x1])
            assert np.allclose(out[a, b], ref, rtol=1e-12, atol=1e-12)


def test_density_transform_round_trip():
    class _Net:  # the attributes train() leaves on a network
        target_scales = np.array([0.2, 0.15, 1.3], np.float32)
        rho_s0 = 0.05
    rng = np.random.default_rng(4)
    t = rng.normal(size=(2, 5, 5, 3)).astype(np.float32)
    t[..., 2] *= 10.0  # corner-peak-like magnitudes
    n = WS.to_net(_Net, t)
    assert np.allclose(n[..., :2], t[..., :2] / _Net.target_scales[:2])
    assert np.allclose(n[..., 2], np.arcsinh(t[..., 2] / 0.05) / 1.3, rtol=1e-5)
    assert np.allclose(WS.from_net(_Net, n), t, rtol=1e-4, atol=1e-5)
    _Net.rho_s0 = None
    assert np.allclose(WS.to_net(_Net, t), t / _Net.target_scales)


def test_warm_fields_host_pipeline_with_a_stand_in_network():
    # encode -> predict -> from_net -> resample -> mass shift, threaded over
    # cavities; a stand-in network returns a known prediction in its units
    class _Net:
        H = 33
        target_scales = np.array([0.5, 0.25, 2.0], np.float32)
        rho_s0 = None

        def predict(self, X):
            assert X.shape[1:] == (33, 33, WS.IN_CH) and X.dtype == np.float32
            out = np.zeros(X.shape[:3] + (3,), np.float32)
            out[..., 0] = 0.4 / 0.5   # u / lid = 0.4 everywhere
            out[..., 1] = -0.2 / 0.25  # v / lid = -0.2
            out[..., 2] = X[..., 3] / 2.0  # (rho - 1) / 3 lid^2 = x / Lmax
            return out
    specs = [_spec(17, 9, [(4, 3, 6, 5)], lid=0.05), _spec(12, 20, [], lid=0.1)]
    out = WS.warm_fields(_Net(), specs)
    for sp, (rho, u, v) in zip(specs, out):
        n, m = sp["NX"] + 2, sp["NY"] + 3
        assert rho.shape == u.shape == v.shape == (n * m,)
        fl = np.zeros((m, n), bool)
        fl[1:sp["NY"] + 1, 1:sp["NX"] + 1] = WS._fluid_mask(sp)
        fl = fl.ravel()
        assert np.allclose(u[fl], 0.4 * sp["lid"]) and np.allclose(v[fl], -0.2 * sp["lid"])
        assert (u[~fl] == 0).all() and (rho[~fl] == 1).all()
        # mass of the rest start kept, the predicted gradient along x kept
        assert np.isclose(rho[fl].mean(), 1.0)
        r2 = rho.reshape(m, n)
        j = 1 + sp["NY"] // 2
        row = r2[j, 1:sp["NX"] + 1][WS._fluid_mask(sp)[j - 1]]
        assert (np.diff(row) > 0).all()
    uni = WS.warm_fields(_Net(), specs, density=False)
    assert all((r == 1).all() for r, _, _ in uni)


def test_density_clip_bounds_only_the_density_channel():
    rng = np.random.default_rng(6)
    inp = np.zeros((3, 9, 9, WS.IN_CH), np.float32)
    inp[..., 0] = rng.random((3, 9, 9)) > 0.2
    t = rng.normal(size=(3, 9, 9, 3)).astype(np.float32)
    t[0, 0, 0, 2] = 50.0  # a corner peak
    c = WS.clip_density(inp, t, 0.95)
    x = t[..., 2][inp[..., 0] > 0]
    lo, hi = np.quantile(x, 0.05), np.quantile(x, 0.95)
    assert np.array_equal(c[..., :2], t[..., :2])
    assert c[..., 2].max() <= hi + 1e-6 and c[..., 2].min() >= lo - 1e-6
    inside = (t[..., 2] >= lo) & (t[..., 2] <= hi)
    assert np.array_equal(c[..., 2][inside], t[..., 2][inside])
    assert t[0, 0, 0, 2] == 50.0  # the input set is left as it was


@pytest.mark.parametrize("q", [0.5, 0.3, 0.0001, 1.0, 1.5, -0.9])
def test_density_clip_rejects_quantiles_outside_half_to_one(q):
    """q <= 0.5 would collapse the density channel to one value, q >= 1
    would clip nothing: both are refused, by clip_density and by train()."""
    inp = np.ones((2, 5, 5, WS.IN_CH), np.float32)
    t = np.zeros((2, 5, 5, 3), np.float32)
    with pytest.raises(ValueError):
        WS.clip_density(inp, t, q)
    with pytest.raises(ValueError):
        WS.train(None, inp, t, epochs=1, rho_clip=q)
