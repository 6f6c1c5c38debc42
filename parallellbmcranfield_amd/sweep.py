"""Batched sweeps of lid-driven cavities on one GPU.

A sweep is a set of cavities of uneven size, aspect ratio and Reynolds number,
each with a few solid rectangles, packed into one lattice
(``lbm_case_pack``) and advanced together by the batched solver
(``lbm_solver_create_batched``): every instance runs until its own velocity
residual ||u(t) - u(t-K)|| / (K ||u(t)||) falls below the criterion, then
stops while the others go on.  The same cavities can be run one at a time
with the single-lattice solver for comparison.

No counterpart in the reference (its solver handles one lattice per run);
each instance is the reference's cavity (cases/coarse) with the geometry
generalised.
"""
from __future__ import annotations

import math
import time

import numpy as np

from . import (MRT, PREC_F32, INIT_REST, RESID_VELOCITY, TUNE_NO_GRAPH, Case, Solver,
               make_options)


def random_cavities(count: int = 256, seed: int = 0, side=(65, 513), aspect=(0.5, 2.0),
                    re=(100.0, 3200.0), lid: float = 0.1, max_rects: int = 3,
                    rect_frac=(0.05, 0.25)) -> list[dict]:
    """Seeded cavity specifications: NX uniform in `side` (integers), NY =
    NX / aspect (aspect uniform in `aspect`, NY clipped to `side`), Re
    log-uniform in `re` on the lid length NX (nu = lid NX / Re), 0..max_rects
    solid rectangles with sides rect_frac x the cavity side, kept two nodes
    clear of the walls."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        nx = int(rng.integers(side[0], side[1] + 1))
        a = float(rng.uniform(*aspect))
        ny = int(min(max(round(nx / a), side[0]), side[1]))
        Re = float(math.exp(rng.uniform(math.log(re[0]), math.log(re[1]))))
        rects = []
        for _ in range(int(rng.integers(0, max_rects + 1))):
            w = max(1, int(round(rng.uniform(*rect_frac) * nx)))
            h = max(1, int(round(rng.uniform(*rect_frac) * ny)))
            i0 = int(rng.integers(3, max(4, nx - w - 2)))
            j0 = int(rng.integers(3, max(4, ny - h - 2)))
            rects.append((i0, j0, min(i0 + w - 1, nx - 2), min(j0 + h - 1, ny - 3)))
        out.append(dict(NX=nx, NY=ny, rects=rects, Re=Re, nu=lid * nx / Re, lid=lid))
    return out


def sweep_options(precision=PREC_F32, crit=1e-6, resid_every=100, max_steps=10**6,
                  rho0=1.0, **kw):
    """MRT (Lallemand-Luo rates with 1/tau on the stress moments, textbook
    equilibria), rest start, per-instance velocity residual every
    `resid_every` steps, stop at `crit`."""
    o = make_options(collision=MRT, uavg=0.1, rho_ini=rho0, viscosity=0.01,
                     precision=precision, init=INIT_REST, resid_kind=RESID_VELOCITY,
                     resid_every=resid_every, crit_v=crit, use_stop_rule=1,
                     max_iterations=max_steps, mrt_textbook_meq=1,
                     mrt_rates=mrt_rates(0.01), **kw)
    return o


def mrt_rates(nu: float) -> tuple:
    """Lallemand-Luo rates (0, 1.4, 1.4, 0, 1.2, 0, 1.2, 1/tau, 1/tau); a
    batched solver replaces the last two by each instance's 1/tau."""
    om = 1.0 / (3.0 * float(np.float32(nu)) + 0.5)
    return (0.0, 1.4, 1.4, 0.0, 1.2, 0.0, 1.2, om, om)


def fluid_nodes(case: Case) -> int:
    return int(case.cells()["fluid"].sum())


def build_batch(specs):
    cases = [Case.cavity_ex(s["NX"], s["NY"], s["rects"]) for s in specs]
    fl = [fluid_nodes(c) for c in cases]
    packed = Case.pack(cases)
    return packed, fl


def run_batched(specs, opts=None, chunk: int = 2000, progress=None) -> dict:
    """Advance all instances in one batched solver until each has converged
    (or max_iterations).  Returns per-instance steps and latency and the
    aggregate MLUPS of the useful work (fluid nodes x own steps / wall)."""
    opts = opts or sweep_options()
    packed, fl = build_batch(specs)
    s = Solver(packed, opts, instances=[(sp["nu"], sp["lid"]) for sp in specs])
    s.sync()
    t0 = time.perf_counter()
    while True:
        done = s.step(chunk)
        if progress:
            it, _, _ = s.batch_status()
            progress(s.iter, int((it != -1).sum()), len(specs))
        if done < chunk or s.iter >= opts.max_iterations:
            break
    s.sync()
    wall = time.perf_counter() - t0
    it, sec, res = s.batch_status()
    steps = np.where(it > 0, it, s.iter)
    work = float(np.dot(np.asarray(fl, np.float64), steps))
    return dict(wall_s=wall, iters=int(s.iter), conv_iter=it, conv_s=sec, resid=res,
                fluid=np.asarray(fl), lattice=(packed.n, packed.m),
                mlups=work / wall / 1e6, solver=s, case=packed)


def run_single(spec, opts=None, chunk: int = 5000, fields: bool = False) -> dict:
    """One cavity alone with the single-lattice solver, to the same criterion
    (any storage precision); fields=True adds its final rho, u, v (n*m
    arrays of lbm_case_cavity_ex's lattice) and whether it converged."""
    opts = opts or sweep_options()
    o = make_options(collision=MRT, uavg=spec["lid"], rho_ini=opts.rho_ini,
                     viscosity=spec["nu"], precision=opts.precision, init=INIT_REST,
                     resid_kind=RESID_VELOCITY, resid_every=opts.resid_every,
                     crit_v=opts.crit_v, use_stop_rule=1, max_iterations=opts.max_iterations,
                     mrt_textbook_meq=opts.mrt_textbook_meq, mrt_rates=mrt_rates(spec["nu"]))
    case = Case.cavity_ex(spec["NX"], spec["NY"], spec["rects"])
    s = Solver(case, o)
    s.sync()
    t0 = time.perf_counter()
    while True:
        done = s.step(chunk)
        if done < chunk or s.iter >= o.max_iterations:
            break
    s.sync()
    wall = time.perf_counter() - t0
    fl = fluid_nodes(case)
    out = dict(wall_s=wall, iters=int(s.iter), fluid=fl, mlups=fl * s.iter / wall / 1e6,
               resid=float(s.residuals()[0]))
    if fields:
        out["rho"], out["u"], out["v"] = s.macro()
        out["converged"] = bool(s.iter < o.max_iterations)
    s.close()
    return out


def _time_steps(case: Case, o, instances, steps: int) -> tuple[float, dict, dict]:
    """Seconds per plain step of `steps` steps after 24 untimed ones (graphs
    are captured on first use), and the solver's multi-step layout."""
    s = Solver(case, o, instances=instances) if instances else Solver(case, o)
    try:
        two, three = s.two_step_info(), s.three_step_info()
        s.step(24)
        s.sync()
        t0 = time.perf_counter()
        s.step(steps)
        s.sync()
        return (time.perf_counter() - t0) / steps, two, three
    finally:
        s.close()


def autotune(case: Case, opts, instances=None, steps: int = 198,
             blocks=(256, 512, 1024), groups=(1, 2, 4, 8, 16),
             two_rows=(0,), three_rows=(0,)) -> tuple[dict, list]:
    """Time `steps` plain steps for each launch shape on this lattice and
    return the fastest (and every timing).  Results are identical for every
    shape; only the time differs.

    Stage 1, single steps (multi-step launches off while timing): a bulk tile
    is `block_size` consecutive cells of a storage row (one workgroup);
    `xcd_group` runs of that many adjacent tiles go to one XCD (its L2 serves
    the +-1 pulls the tiles share); FP16 storage also tries one cell per
    thread against the paired-cell kernel.  Block 256 runs the fused step
    kernel, larger blocks split bulk and boundary launches.

    Stage 2: at the best stage-1 shape, the tile shape of multi-step
    launches: two plain steps per launch with `two_rows` output rows per
    256-column tile, three per launch with `three_rows` (single lattices
    only; 0: the library's own pick, which may be single steps where
    multi-step launches lose), each against single steps.  A variant whose
    solver runs no such launches is dropped.  `best` holds the options to
    copy (block_size, xcd_group, two_step_rows, three_step_rows, tune_flags)
    and what ran."""
    from . import PREC_F16, PREC_F16_SR, TUNE_NO_PAIRS, TUNE_NO_THREE_STEP, TUNE_NO_TWO_STEP
    fp16 = opts.precision in (PREC_F16, PREC_F16_SR)
    keep = opts.tune_flags & ~(TUNE_NO_PAIRS | TUNE_NO_TWO_STEP | TUNE_NO_THREE_STEP)

    def variant(bs, g, flags, rows2=0, rows3=0):
        o = type(opts).from_buffer_copy(opts)
        o.block_size, o.xcd_group = bs, g
        o.tune_flags = keep | flags
        o.two_step_rows, o.three_step_rows = rows2, rows3
        o.use_stop_rule = 0
        o.resid_kind = 3
        o.crit_v = -1.0  # no instance stops while timing
        return o

    def record(o, dt, two, three, stage):
        steps_per_launch = 3 if three["cells"] else (2 if two["cells"] else 1)
        return dict(stage=stage, block_size=o.block_size, xcd_group=o.xcd_group,
                    pairs=fp16 and o.block_size == 256 and not (o.tune_flags & TUNE_NO_PAIRS),
                    two_step_rows=o.two_step_rows, three_step_rows=o.three_step_rows,
                    tune_flags=o.tune_flags, steps_per_launch=steps_per_launch,
                    tile_rows=three["rows"] if three["cells"] else (two["rows"] if two["cells"] else 0),
                    ms_per_step=dt * 1e3)

    single = TUNE_NO_TWO_STEP | TUNE_NO_THREE_STEP
    results = []
    for bs in blocks:
        for g in groups:
            for pairs in ((True, False) if fp16 and bs == 256 else (True,)):
                o = variant(bs, g, single | (0 if pairs else TUNE_NO_PAIRS))
                dt, two, three = _time_steps(case, o, instances, steps)
                results.append(record(o, dt, two, three, 1))
    best1 = min(results, key=lambda r: r["ms_per_step"])
    base = best1["tune_flags"] & ~single
    for r2 in two_rows:
        o = variant(best1["block_size"], best1["xcd_group"], base | TUNE_NO_THREE_STEP, rows2=r2)
        dt, two, three = _time_steps(case, o, instances, steps)
        if two["cells"]:
            results.append(record(o, dt, two, three, 2))
    for r3 in (() if instances else three_rows):
        o = variant(best1["block_size"], best1["xcd_group"], base, rows3=r3)
        dt, two, three = _time_steps(case, o, None, steps)
        if three["cells"]:
            results.append(record(o, dt, two, three, 2))
    best = min(results, key=lambda r: r["ms_per_step"])
    return best, results


__all__ = ["random_cavities", "sweep_options", "build_batch", "run_batched", "run_single",
           "autotune", "fluid_nodes", "TUNE_NO_GRAPH"]
