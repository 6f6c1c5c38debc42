"""Learned warm start of batched cavity runs.

The network (``Net``, lbm_net_* in the library: U-Net of 3x3 convolutions,
BF16 on the matrix cores, trained with the library's own backward and Adam)
maps a cavity drawn on a G x G grid to its steady (u, v, rho) there:

  inputs   channel 0  fluid mask: the cavity scaled uniformly so that its
                      longer side spans the grid, anchored at the grid
                      origin; grid points outside it or in an obstacle are 0
           channel 1  (ln Re - ln 100) / ln 50      (Re on the lid length)
           channel 2  lid speed / 0.1
           channels 3-6  distances to the west, east, south and north (lid)
                      walls over the longer side (0 outside the cavity):
                      a 3x3-conv U-Net sees about 100 grid points, less than
                      the cavity, so every point is told where it lies
           channel 7  distance to the nearest solid node (walls and
                      obstacles) over the longer side, 0 outside the fluid
  targets  u / lid, v / lid, (rho - 1) / (3 lid^2)   (0 outside the fluid)

Training data are steady states of seeded cavities run by the batched solver
to the velocity residual criterion; the prediction for a new cavity is
sampled back onto its lattice and seeds f = feq(rho, u, v).  The rest start
(rho = 1, u = 0) stays the fallback and the reference for every comparison.

No counterpart in the reference.
"""
from __future__ import annotations

from concurrent.futures import ThreadPoolExecutor
import math
import os
import time

import numpy as np

from . import INIT_FIELD, INIT_FIELD_NEQ, Case, Net, Solver
from . import sweep as SW

GRID = 257
IN_CH = 8
DILATION = 4  # bottleneck stencils dilated by 4 and 8 at 33^2: the field of view spans the grid
# the density channel's weight in the loss: its corner peaks generalise
# poorly, and a lighter density loss leaves capacity to the velocity (held-out
# seed error 0.13 -> 0.10 at 0.1)
RHO_WEIGHT = 0.1
RHO_ASINH = 0.0  # > 0: density target through asinh(x / s0), s0 this quantile of |x|
DIV_WEIGHT = 0.0  # weight of the predicted velocity's squared divergence in the loss
RHO_CLIP = 0.0  # > 0: density target clipped to these quantiles (e.g. 0.99)
WEIGHT_DECAY = 0.05  # decoupled weight decay of the Adam step (AdamW)
RE_LO, RE_SPAN = math.log(100.0), math.log(50.0)


def _geometry(spec, G):
    NX, NY = spec["NX"], spec["NY"]
    scale = (G - 1) / max(NX - 1, NY - 1)
    return NX, NY, scale


def _fluid_mask(spec):
    """Fluid flags of the cavity's own nodes, [NY, NX] (j-1, i-1)."""
    m = np.ones((spec["NY"], spec["NX"]), bool)
    for i0, j0, i1, j1 in spec["rects"]:
        m[j0 - 1:j1, i0 - 1:i1] = False
    return m


def encode(spec, G: int = GRID) -> np.ndarray:
    """Network input [G, G, IN_CH] of one cavity specification."""
    NX, NY, scale = _geometry(spec, G)
    fl = _fluid_mask(spec)
    g = np.arange(G)
    ii = np.rint(g / scale).astype(int)  # nearest cavity node (0-based)
    inside_x = g / scale <= NX - 1 + 1e-9
    inside_y = g / scale <= NY - 1 + 1e-9
    mask = np.zeros((G, G), np.float32)
    jy, ix = np.meshgrid(np.minimum(ii, NY - 1), np.minimum(ii, NX - 1), indexing="ij")
    inside = inside_y[:, None] & inside_x[None, :]
    mask[inside] = fl[jy[inside], ix[inside]]
    out = np.zeros((G, G, IN_CH), np.float32)
    out[..., 0] = mask
    out[..., 1] = (math.log(spec["Re"]) - RE_LO) / RE_SPAN
    out[..., 2] = spec["lid"] / 0.1
    Lmax = max(NX, NY) - 1
    xs, ys = np.meshgrid(g / scale, g / scale)  # node coordinates (0-based) of grid points
    out[..., 3] = np.where(inside, xs / Lmax, 0.0)
    out[..., 4] = np.where(inside, (NX - 1 - xs) / Lmax, 0.0)
    out[..., 5] = np.where(inside, ys / Lmax, 0.0)
    out[..., 6] = np.where(inside, (NY - 1 - ys) / Lmax, 0.0)
    # Euclidean distance of each fluid node to the nearest solid node of the
    # cavity lattice (its ring of walls included), sampled like the mask
    from scipy.ndimage import distance_transform_edt
    solid = np.zeros((NY + 2, NX + 2), bool)
    solid[0, :] = solid[-1, :] = True
    solid[:, 0] = solid[:, -1] = True
    solid[1:NY + 1, 1:NX + 1] = ~fl
    dist = distance_transform_edt(~solid)[1:NY + 1, 1:NX + 1]
    out[..., 7] = np.where(mask > 0, dist[jy, ix] / Lmax, 0.0)
    return out


def _bilinear(field, gy, gx):
    """field [h, w, ...] sampled at rows gy and columns gx (1-D fractional
    coordinates, clamped) -> [len(gy), len(gx), ...]: the bilinear weights
    separate, so rows are blended first and columns second."""
    h, w = field.shape[:2]
    gy = np.clip(gy, 0, h - 1)
    gx = np.clip(gx, 0, w - 1)
    y0 = np.floor(gy).astype(int)
    x0 = np.floor(gx).astype(int)
    ty, tx = gy - y0, gx - x0
    y1, x1 = np.minimum(y0 + 1, h - 1), np.minimum(x0 + 1, w - 1)
    ext = (slice(None),) + (None,) * (field.ndim - 1)
    rows = field[y0] * (1 - ty)[ext] + field[y1] * ty[ext]
    ext = (None, slice(None)) + (None,) * (field.ndim - 2)
    return rows[:, x0] * (1 - tx)[ext] + rows[:, x1] * tx[ext]


def _pool():
    """Thread pool for the per-cavity host work (numpy and scipy release the
    interpreter lock inside their array loops)."""
    global _POOL
    if _POOL is None:
        _POOL = ThreadPoolExecutor(max_workers=min(16, os.cpu_count() or 1))
    return _POOL


_POOL = None


def field_to_grid(spec, rho, u, v, G: int = GRID) -> np.ndarray:
    """Steady fields on the cavity lattice (n*m arrays of lbm_case_cavity_ex's
    lattice) -> normalised targets [G, G, 3] (0 where the input mask is 0)."""
    NX, NY, scale = _geometry(spec, G)
    n, m = NX + 2, NY + 3
    lid = spec["lid"]
    fl = _fluid_mask(spec)
    chans = []
    for f, off, sc in ((u, 0.0, lid), (v, 0.0, lid), (rho, 1.0, 3 * lid * lid)):
        a = (np.asarray(f).reshape(m, n)[1:NY + 1, 1:NX + 1] - off) / sc
        chans.append(np.where(fl, a, 0.0))
    g = np.arange(G) / scale
    mask = encode(spec, G)[..., 0]
    out = _bilinear(np.stack(chans, -1), g, g)
    return np.where(mask[..., None] > 0, out, 0.0).astype(np.float32)


def grid_to_field(spec, pred, G: int = GRID):
    """Prediction [G, G, 3] -> (rho, u, v) n*m arrays on the cavity lattice
    (fluid nodes from the prediction, other cells at rest)."""
    NX, NY, scale = _geometry(spec, G)
    n, m = NX + 2, NY + 3
    lid = spec["lid"]
    # stay on grid points that lie inside the cavity (the network's output
    # outside it is not trained)
    fy = np.minimum(np.arange(NY) * scale, math.floor((NY - 1) * scale + 1e-9))
    fx = np.minimum(np.arange(NX) * scale, math.floor((NX - 1) * scale + 1e-9))
    fl = _fluid_mask(spec)
    samp = _bilinear(np.asarray(pred, np.float64), fy, fx)
    out = []
    for k, off, sc in ((2, 1.0, 3 * lid * lid), (0, 0.0, lid), (1, 0.0, lid)):
        a = np.full((m, n), off)
        a[1:NY + 1, 1:NX + 1] = np.where(fl, off + sc * samp[..., k], off)
        out.append(a.ravel())
    return tuple(out)  # rho, u, v


def steady_states(specs, opts=None, chunk: int = 5000, progress=None):
    """Run the cavities in one batched solver to the residual criterion and
    return their steady fields [(rho, u, v) per instance] and convergence
    iterations (-1 / -2: not converged / diverged)."""
    r = SW.run_batched(specs, opts, chunk=chunk, progress=progress)
    s, packed = r["solver"], r["case"]
    rho, u, v = s.macro()
    fields = []
    for k in range(len(specs)):
        x0, y0, w, h = packed.rects[k]
        fields.append(tuple(a.reshape(packed.m, packed.n)[y0:y0 + h, x0:x0 + w].ravel().copy()
                            for a in (rho, u, v)))
    s.close()
    return fields, r["conv_iter"], r


def channel_scales(inputs: np.ndarray, targets: np.ndarray) -> np.ndarray:
    """RMS of each target channel over the fluid points of a set: the
    network learns targets / scale, so u, v and rho weigh alike in the loss
    (the density channel's corner peaks would otherwise dominate it)."""
    m = inputs[..., 0:1]
    ms = (m * targets.astype(np.float64) ** 2).sum(axis=(0, 1, 2)) / max(1.0, float(m.sum()))
    return np.sqrt(np.maximum(ms, 1e-12)).astype(np.float32)


def clip_density(inputs: np.ndarray, targets: np.ndarray, q: float) -> np.ndarray:
    """Targets with the density channel clipped to its [1 - q, q] quantiles
    over the fluid points of the set (the other channels unchanged).
    0.5 < q < 1: at q <= 0.5 the interval is empty (the channel would become
    one constant) and at q >= 1 nothing is clipped."""
    if not (0.5 < q < 1.0):
        raise ValueError(f"density clip quantile must lie in (0.5, 1), got {q!r}")
    x = targets[..., 2][inputs[..., 0] > 0]
    lo, hi = np.quantile(x, 1.0 - q), np.quantile(x, q)
    out = np.array(targets, copy=True)
    out[..., 2] = np.clip(out[..., 2], lo, hi)
    return out


def to_net(net: Net, targets: np.ndarray) -> np.ndarray:
    """Normalised targets [..., 3] (``field_to_grid``) -> the network's
    units: the density channel through asinh(x / s0) when ``net.rho_s0`` is
    set (linear in the bulk, logarithmic at the lid corners' pressure
    peaks), then every channel divided by ``net.target_scales``."""
    t = np.array(targets, np.float32, copy=True)
    s0 = getattr(net, "rho_s0", None)
    if s0:
        t[..., 2] = np.arcsinh(t[..., 2] / s0)
    return t / getattr(net, "target_scales", np.ones(3, np.float32))


def from_net(net: Net, pred: np.ndarray) -> np.ndarray:
    """The inverse of ``to_net``."""
    t = pred * getattr(net, "target_scales", np.ones(3, np.float32))
    s0 = getattr(net, "rho_s0", None)
    if s0:
        t[..., 2] = s0 * np.sinh(t[..., 2])
    return t


def train(net: Net, inputs: np.ndarray, targets: np.ndarray, epochs: int = 30,
          lr: float = 1e-3, seed: int = 0, log=None, rho_weight: float = RHO_WEIGHT,
          weight_decay: float = WEIGHT_DECAY, rho_asinh: float = RHO_ASINH,
          div_weight: float = DIV_WEIGHT, rho_clip: float = RHO_CLIP) -> list[float]:
    """Adam over the set with a cosine learning-rate decay on targets scaled
    per channel (``channel_scales``, kept as ``net.target_scales``); returns
    the mean loss of every epoch.  rho_weight: weight of the density channel
    in the loss relative to u and v (its scale grows by 1/sqrt(rho_weight)),
    weight_decay: decoupled weight decay of the Adam step, rho_asinh: if
    > 0, the density target passes through asinh(x / s0) with s0 the
    rho_asinh quantile of |x| over the fluid points (``to_net``), div_weight:
    weight of the predicted velocity's squared divergence in the loss
    (``Net.set_div_weight``; compressive seed errors cost the most),
    rho_clip: if not 0, the density target is clipped to its [1 - rho_clip,
    rho_clip] quantiles over the set's fluid points (``clip_density``), so
    the lid corners' pressure peaks leave the loss to the bulk field."""
    if rho_clip != 0:  # 0 is off; anything else must be a valid quantile
        targets = clip_density(inputs, targets, rho_clip)
    net.rho_s0 = None
    if rho_asinh > 0:
        x = np.abs(targets[..., 2][inputs[..., 0] > 0])
        net.rho_s0 = float(max(np.quantile(x, rho_asinh), 1e-6))
    net.target_scales = np.ones(3, np.float32)
    t = to_net(net, targets)
    sc = channel_scales(inputs, t)
    sc[2] /= math.sqrt(rho_weight)
    net.target_scales = sc
    net.set_weight_decay(weight_decay)
    net.set_div_weight(div_weight, float(sc[0]), float(sc[1]))
    net.set_data(inputs, t / sc)
    del t
    rng = np.random.default_rng(seed)
    N, B = inputs.shape[0], net.batch
    steps_per_epoch = max(1, N // B)
    total = epochs * steps_per_epoch
    hist, k = [], 0
    for ep in range(epochs):
        order = rng.permutation(N)
        ls = []
        for b in range(steps_per_epoch):
            a = lr * 0.5 * (1 + math.cos(math.pi * k / total))
            ls.append(net.train(order[b * B:(b + 1) * B], a))
            k += 1
        hist.append(float(np.mean(ls)))
        if log:
            log(f"epoch {ep + 1}/{epochs}: loss {hist[-1]:.4e}")
    return hist


def warm_fields(net: Net, specs, density: bool = True):
    """Network prediction of every cavity, as (rho, u, v) on its lattice.
    The predicted density is shifted so the cavity holds the mass of the rest
    start (mean rho = 1 over its fluid nodes), as the walls conserve mass;
    density=False seeds rho = 1 with the predicted velocity."""
    X = np.empty((len(specs), net.H, net.H, IN_CH), np.float32)

    def enc(k):
        X[k] = encode(specs[k], net.H)

    list(_pool().map(enc, range(len(specs))))
    pred = from_net(net, net.predict(X))
    return list(_pool().map(lambda sp: _seed(sp[0], sp[1], net.H, density), zip(specs, pred)))


def _seed(spec, pred, G, density):
    """(rho, u, v) on the cavity's lattice from its prediction [G, G, 3]."""
    rho, u, v = grid_to_field(spec, pred, G)
    if not density:
        return np.ones_like(rho), u, v
    NX, NY = spec["NX"], spec["NY"]
    fl = np.zeros((NY + 3, NX + 2), bool)
    fl[1:NY + 1, 1:NX + 1] = _fluid_mask(spec)
    fl = fl.ravel()
    rho = rho.copy()
    rho[fl] -= rho[fl].mean() - 1.0
    return rho, u, v


def packed_field(packed: Case, fields):
    full = [np.ones(packed.n * packed.m), np.zeros(packed.n * packed.m),
            np.zeros(packed.n * packed.m)]
    for k, f in enumerate(fields):
        x0, y0, w, h = packed.rects[k]
        for F, a in zip(full, f):
            F.reshape(packed.m, packed.n)[y0:y0 + h, x0:x0 + w] = np.asarray(a).reshape(h, w)
    return full


def compare_starts(net: Net, specs, opts=None, chunk: int = 5000, log=None,
                   neq: bool = True, density: bool = True, run_rest=None) -> dict:
    """Held-out comparison: the same cavities from rest and from the network's
    prediction, each as one batched run to the criterion.  Per instance:
    steps and wall time to convergence (both include the solver's
    construction, the warm start's also the prediction), and the relative L2 difference of the two converged
    velocity fields and of the seed against the rest start's converged
    velocity (the prediction error).  neq: seed feq plus the Chapman-Enskog non-equilibrium
    part of the predicted velocity gradients (LBM_INIT_FIELD_NEQ).  density:
    seed the predicted density (else rho = 1).  run_rest: the ``rest_run`` of
    an earlier comparison on the same cavities, reused."""
    opts = opts or SW.sweep_options()
    packed, fl = SW.build_batch(specs)
    inst = [(sp["nu"], sp["lid"]) for sp in specs]

    seeds = {}

    def run(field):
        # the clock starts before solver construction for both starts: a
        # field start's construction (seed upload, and for INIT_FIELD_NEQ
        # the host pass over the velocity gradients) is part of its cost
        t0 = time.perf_counter()
        pre = 0.0
        if field == "warm":
            flds = warm_fields(net, specs, density)
            fld = packed_field(packed, flds)
            pre = time.perf_counter() - t0
            seeds["u"], seeds["v"] = fld[1], fld[2]
            o = type(opts).from_buffer_copy(opts)
            o.init = INIT_FIELD_NEQ if neq else INIT_FIELD
            s = Solver(packed, o, field=fld, instances=inst)
        else:
            s = Solver(packed, opts, instances=inst)
        s.sync()
        t1 = time.perf_counter()
        setup = t1 - t0 - pre  # solver construction
        while True:
            done = s.step(chunk)
            if done < chunk or s.iter >= opts.max_iterations:
                break
        s.sync()
        wall = time.perf_counter() - t1
        it, sec, _ = s.batch_status()
        rho, u, v = s.macro()
        s.close()
        return dict(conv_iter=it, conv_s=sec + pre + setup, wall_s=wall + pre + setup,
                    predict_s=pre, setup_s=setup, rho=rho, u=u, v=v)

    rest = run_rest if run_rest is not None else run("rest")
    if log:
        log(f"rest start: {rest['wall_s']:.1f} s")
    warm = run("warm")
    if log:
        log(f"warm start: {warm['wall_s']:.1f} s (prediction {warm['predict_s']:.2f} s, "
            f"solver set-up {warm['setup_s']:.2f} s; rest set-up {rest['setup_s']:.2f} s)")
    both = (rest["conv_iter"] > 0) & (warm["conv_iter"] > 0)
    fluid = packed.cells()["fluid"].reshape(packed.m, packed.n) == 1

    def rel(a_u, a_v, b_u, b_v, sl):  # relative L2 over the instance's fluid nodes
        f = fluid[sl]
        ur, vr = (a.reshape(packed.m, packed.n)[sl][f] for a in (b_u, b_v))
        uw, vw = (a.reshape(packed.m, packed.n)[sl][f] for a in (a_u, a_v))
        num = np.sqrt(((uw - ur) ** 2 + (vw - vr) ** 2).sum())
        den = np.sqrt((ur ** 2 + vr ** 2).sum())
        return num / den if den > 0 else 0.0

    diff, pred = [], []
    for k in range(len(specs)):
        x0, y0, w, h = packed.rects[k]
        sl = (slice(y0, y0 + h), slice(x0, x0 + w))
        diff.append(rel(warm["u"], warm["v"], rest["u"], rest["v"], sl))
        pred.append(rel(seeds["u"], seeds["v"], rest["u"], rest["v"], sl))
    diff, pred = np.asarray(diff), np.asarray(pred)
    # held-out loss per channel, in the network's scaled units, against the
    # rest start's converged fields
    X = np.stack([encode(sp, net.H) for sp in specs])
    P = net.predict(X)
    mse = np.zeros(3)
    wsum = 0.0
    rest_rho = rest["rho"]
    for k, sp in enumerate(specs):
        x0, y0, w, h = packed.rects[k]
        sl = (slice(y0, y0 + h), slice(x0, x0 + w))
        T = to_net(net, field_to_grid(sp, *(a.reshape(packed.m, packed.n)[sl].ravel()
                                            for a in (rest_rho, rest["u"], rest["v"])), net.H))
        m = X[k, ..., 0:1]
        mse += (m * (P[k] - T) ** 2).sum(axis=(0, 1))
        wsum += float(m.sum())
    heldout = mse / max(1.0, wsum)
    return dict(
        n=len(specs), both_converged=int(both.sum()),
        steps_rest=rest["conv_iter"], steps_warm=warm["conv_iter"],
        time_rest=rest["conv_s"], time_warm=warm["conv_s"],
        median_steps_ratio=float(np.median(rest["conv_iter"][both] / warm["conv_iter"][both]))
        if both.any() else None,
        median_time_rest=float(np.median(rest["conv_s"][both])) if both.any() else None,
        median_time_warm=float(np.median(warm["conv_s"][both])) if both.any() else None,
        wall_rest=rest["wall_s"], wall_warm=warm["wall_s"],
        rest_run=rest, warm_run=warm, packed=packed,
        field_rel_l2=diff, field_rel_l2_median=float(np.median(diff[both])) if both.any() else None,
        seed_rel_l2=pred, seed_rel_l2_median=float(np.median(pred[both])) if both.any() else None,
        heldout_mse=heldout,
        predict_s=warm["predict_s"], setup_rest_s=rest["setup_s"], setup_warm_s=warm["setup_s"])


__all__ = ["GRID", "IN_CH", "DILATION", "RHO_WEIGHT", "WEIGHT_DECAY", "RHO_ASINH", "DIV_WEIGHT",
           "RHO_CLIP", "clip_density", "encode",
           "field_to_grid", "to_net", "from_net",
           "grid_to_field", "steady_states", "channel_scales", "train", "warm_fields",
           "packed_field", "compare_starts"]
