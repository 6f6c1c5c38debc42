"""CPU oracle of the warm-start network — test infrastructure only.

A numpy restatement (float64 or float32) of the U-Net the library's
``lbm_net_*`` entry points run on the GPU: same layer order, same canonical
parameter layout, same loss.  Forward and a hand-derived backward; the
backward is pinned by central finite differences in float64
(tests/test_nn_oracle.py).  The network has no counterpart in the reference
program, so there is no reference file to cite; the layout it mirrors is
documented with ``lbm_net_get_params`` in include/lbm.h.

Layers for `levels` L and base c (C_l = c 2^l):
  encoder l < L-1:  a_l = relu(conv(x_l)), s_l = relu(conv(a_l)), x_{l+1} = avgpool2(s_l)
  bottleneck:       a_{L-1}, s_{L-1} likewise, 3x3 stencils dilated by dil and 2 dil
  decoder l = L-2..0: da_l = relu(conv([s_l, up2(prev)])), dd_l = relu(conv(da_l))
  head:             y = dd_0 . Wh^T + bh  (3 channels)
Loss = sum(mask (y - t)^2) / (3 sum(mask)), mask = input channel 0.
"""
from __future__ import annotations

import numpy as np

CIN0 = 32
KOUT = 3


def layer_shapes(levels: int, base: int):
    """[(C0, C1, Cout, level)] in canonical order, then the head."""
    C = [base << l for l in range(levels)]
    convs = []
    for l in range(levels):
        convs.append((CIN0 if l == 0 else C[l - 1], 0, C[l], l))
        convs.append((C[l], 0, C[l], l))
    for l in range(levels - 2, -1, -1):
        convs.append((C[l], C[l + 1], C[l], l))
        convs.append((C[l], 0, C[l], l))
    return convs


def num_params(levels: int, base: int) -> int:
    n = sum(co * 9 * (c0 + c1) + co for c0, c1, co, _ in layer_shapes(levels, base))
    return n + KOUT * base + KOUT


def unpack(params: np.ndarray, levels: int, base: int):
    """Views (W [Cout, 9, Cin], b [Cout]) per conv and (Wh [3, base], bh [3])."""
    out, off = [], 0
    for c0, c1, co, _ in layer_shapes(levels, base):
        ci = c0 + c1
        w = params[off:off + co * 9 * ci].reshape(co, 9, ci)
        off += co * 9 * ci
        b = params[off:off + co]
        off += co
        out.append((w, b))
    wh = params[off:off + KOUT * base].reshape(KOUT, base)
    off += KOUT * base
    bh = params[off:off + KOUT]
    return out, (wh, bh)


def _patches(x, d=1):
    """[N,H,W,C] -> [N,H,W,9,C]; tap t = 3 (dy+1) + (dx+1) reads x[y+d dy, x+d dx]
    (zero outside)."""
    N, H, W, C = x.shape
    xp = np.zeros((N, H + 2 * d, W + 2 * d, C), x.dtype)
    xp[:, d:d + H, d:d + W] = x
    return np.stack([xp[:, d + d * dy:d + d * dy + H, d + d * dx:d + d * dx + W]
                     for dy in (-1, 0, 1) for dx in (-1, 0, 1)], axis=3)


def _col2im(dp, d=1):
    """Adjoint of _patches: [N,H,W,9,C] -> [N,H,W,C]."""
    N, H, W, _, C = dp.shape
    dx_ = np.zeros((N, H + 2 * d, W + 2 * d, C), dp.dtype)
    t = 0
    for dy in (-1, 0, 1):
        for dx in (-1, 0, 1):
            dx_[:, d + d * dy:d + d * dy + H, d + d * dx:d + d * dx + W] += dp[:, :, :, t]
            t += 1
    return dx_[:, d:d + H, d:d + W]


def conv(x, w, b, d=1):
    p = _patches(x, d)
    N, H, W_, _, C = p.shape
    y = p.reshape(-1, 9 * C) @ w.reshape(w.shape[0], -1).T + b
    return y.reshape(N, H, W_, -1), p


def conv_bwd(g, p, w, d=1):
    N, H, W_, _, C = p.shape
    g2 = g.reshape(-1, g.shape[-1])
    dw = (g2.T @ p.reshape(-1, 9 * C)).reshape(w.shape)
    db = g2.sum(0)
    dp = (g2 @ w.reshape(w.shape[0], -1)).reshape(N, H, W_, 9, C)
    return _col2im(dp, d), dw, db


def pool(x):
    N, H, W, C = x.shape
    Ho, Wo = (H + 1) // 2, (W + 1) // 2
    xp = np.zeros((N, 2 * Ho, 2 * Wo, C), x.dtype)
    cnt = np.zeros((2 * Ho, 2 * Wo), x.dtype)
    xp[:, :H, :W] = x
    cnt[:H, :W] = 1
    s = xp.reshape(N, Ho, 2, Wo, 2, C).sum((2, 4))
    c = cnt.reshape(Ho, 2, Wo, 2).sum((1, 3))
    return s / c[None, :, :, None], c


def pool_bwd(g, c, H, W):
    gg = g / c[None, :, :, None]
    up = np.repeat(np.repeat(gg, 2, axis=1), 2, axis=2)
    return up[:, :H, :W]


def up2(x, H, W):
    return np.repeat(np.repeat(x, 2, axis=1), 2, axis=2)[:, :H, :W]


def up2_bwd(g, Hc, Wc):
    N, H, W, C = g.shape
    gp = np.zeros((N, 2 * Hc, 2 * Wc, C), g.dtype)
    gp[:, :H, :W] = g
    return gp.reshape(N, Hc, 2, Wc, 2, C).sum((2, 4))


def relu(x):
    return np.maximum(x, 0)


def forward(params, inp, levels, base, keep=False, dil=1):
    """inp [N,H,W,Ci] (Ci <= 32) -> y [N,H,W,3] (and the cache for backward)."""
    convs, (wh, bh) = unpack(params, levels, base)
    N, H, W, _ = inp.shape
    x = np.zeros((N, H, W, CIN0), params.dtype)
    x[..., :inp.shape[-1]] = inp
    cache = {"x0": x}
    k = 0
    a, s, p_a, p_s, xp, pc = {}, {}, {}, {}, {}, {}
    for l in range(levels):
        (wa, ba), (wb, bb) = convs[k], convs[k + 1]
        k += 2
        da_, db_ = (dil, 2 * dil) if l == levels - 1 else (1, 1)  # dilated bottleneck
        za, p_a[l] = conv(x, wa, ba, da_)
        a[l] = relu(za)
        zs, p_s[l] = conv(a[l], wb, bb, db_)
        s[l] = relu(zs)
        if l < levels - 1:
            xp[l], pc[l] = pool(s[l])
            x = xp[l]
    prev = s[levels - 1]
    da, dd, p_da, p_dd, cat = {}, {}, {}, {}, {}
    for l in range(levels - 2, -1, -1):
        (wa, ba), (wb, bb) = convs[k], convs[k + 1]
        k += 2
        Hl, Wl = s[l].shape[1:3]
        cat[l] = np.concatenate([s[l], up2(prev, Hl, Wl)], axis=-1)
        z, p_da[l] = conv(cat[l], wa, ba)
        da[l] = relu(z)
        z, p_dd[l] = conv(da[l], wb, bb)
        dd[l] = relu(z)
        prev = dd[l]
    y = dd[0] @ wh.T + bh
    if keep:
        cache.update(a=a, s=s, p_a=p_a, p_s=p_s, xp=xp, pc=pc, da=da, dd=dd, p_da=p_da,
                     p_dd=p_dd, cat=cat)
        return y, cache
    return y


def div_penalty(y, mask, su, sv):
    """Discrete divergence of the predicted velocity, d = su (y0[x+1] -
    y0[x-1]) / 2 + sv (y1[y+1] - y1[y-1]) / 2 at points whose four neighbours
    are fluid (0 elsewhere), and the adjoint map g -> d(sum g d) / dy."""
    m = mask[..., 0] > 0
    valid = np.zeros_like(m)
    valid[:, 1:-1, 1:-1] = (m[:, 1:-1, 1:-1] & m[:, 1:-1, 2:] & m[:, 1:-1, :-2] & m[:, 2:, 1:-1]
                            & m[:, :-2, 1:-1])
    d = np.zeros(m.shape)
    d[:, 1:-1, 1:-1] = (su * (y[:, 1:-1, 2:, 0] - y[:, 1:-1, :-2, 0]) / 2 +
                        sv * (y[:, 2:, 1:-1, 1] - y[:, :-2, 1:-1, 1]) / 2)
    d *= valid

    def adjoint(g):
        g = g * valid
        dy = np.zeros_like(y)
        dy[:, :, 1:, 0] += su / 2 * g[:, :, :-1]
        dy[:, :, :-1, 0] -= su / 2 * g[:, :, 1:]
        dy[:, 1:, :, 1] += sv / 2 * g[:, :-1, :]
        dy[:, :-1, :, 1] -= sv / 2 * g[:, 1:, :]
        return dy
    return d, adjoint


def loss_and_grad(params, inp, target, levels, base, dil=1, div_w=0.0, su=1.0, sv=1.0):
    """Masked MSE (+ div_w times the squared divergence of the predicted
    velocity, ``div_penalty``, on the same normalisation) and d loss / d
    params (canonical layout)."""
    y, c = forward(params, inp, levels, base, keep=True, dil=dil)
    convs, (wh, bh) = unpack(params, levels, base)
    grads = np.zeros_like(params)
    gconvs, (gwh, gbh) = unpack(grads, levels, base)
    mask = inp[..., :1]
    nf = max(1.0, float(mask.sum()))
    d = mask * (y - target)
    loss = float((d * d).sum() / (KOUT * nf))
    dy = 2 * d / (KOUT * nf)
    if div_w:
        dv, adjoint = div_penalty(y, mask, su, sv)
        loss += float(div_w * (dv * dv).sum() / (KOUT * nf))
        dy = dy + adjoint(2 * div_w * dv / (KOUT * nf))
    gwh[...] = dy.reshape(-1, KOUT).T @ c["dd"][0].reshape(-1, base)
    gbh[...] = dy.reshape(-1, KOUT).sum(0)
    g = (dy @ wh) * (c["dd"][0] > 0)
    L = levels
    gcat = {}
    gprev_coarse = None
    for l in range(0, L - 1):
        k = 2 * L + 2 * (L - 2 - l)
        (wa, _), (wb, _) = convs[k], convs[k + 1]
        if l > 0:
            g = gprev_coarse * (c["dd"][l] > 0)
        gx, gconvs[k + 1][0][...], gconvs[k + 1][1][...] = conv_bwd(g, c["p_dd"][l], wb)
        gx = gx * (c["da"][l] > 0)
        gcat[l], gconvs[k][0][...], gconvs[k][1][...] = conv_bwd(gx, c["p_da"][l], wa)
        Cl = c["s"][l].shape[-1]
        prev = c["dd"][l + 1] if l + 1 <= L - 2 else c["s"][L - 1]
        gprev_coarse = up2_bwd(gcat[l][..., Cl:], prev.shape[1], prev.shape[2])
    g = gprev_coarse * (c["s"][L - 1] > 0)
    gxp = None
    for l in range(L - 1, -1, -1):
        (wa, _), (wb, _) = convs[2 * l], convs[2 * l + 1]
        if l < L - 1:
            Cl = c["s"][l].shape[-1]
            H, W = c["s"][l].shape[1:3]
            g = (gcat[l][..., :Cl] + pool_bwd(gxp, c["pc"][l], H, W)) * (c["s"][l] > 0)
        da_, db_ = (dil, 2 * dil) if l == L - 1 else (1, 1)
        gx, gconvs[2 * l + 1][0][...], gconvs[2 * l + 1][1][...] = conv_bwd(g, c["p_s"][l], wb,
                                                                             db_)
        gx = gx * (c["a"][l] > 0)
        gxp, gconvs[2 * l][0][...], gconvs[2 * l][1][...] = conv_bwd(gx, c["p_a"][l], wa, da_)
    return loss, grads


def init_params(levels, base, seed=0, dtype=np.float64, in_channels=3):
    rng = np.random.default_rng(seed)
    parts = []
    for i, (c0, c1, co, _) in enumerate(layer_shapes(levels, base)):
        ci = c0 + c1
        real = in_channels if i == 0 else ci
        w = rng.normal(0, np.sqrt(2.0 / (9 * real)), (co, 9, ci))
        w[:, :, real:] = 0
        parts += [w.ravel(), rng.normal(0, 0.05, co)]
    parts += [rng.normal(0, 0.1 / np.sqrt(base), KOUT * base), rng.normal(0, 0.05, KOUT)]
    return np.concatenate(parts).astype(dtype)
