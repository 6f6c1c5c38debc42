#!/usr/bin/env python3
"""Benchmark of the D2Q9 hot path on one MI355X.

Default workload: lid-driven cavity with
4097 x 4097 fluid nodes in the reference mesh convention, Re = 1000 on L = N-1
(nu = 0.4096, tau = 1.7288), MRT with rates s = (0, 1.4, 1.4, 0, 1.2, 0, 1.2,
1/tau, 1/tau) and textbook equilibria, FP32 storage of f - w_k rho0, rest start,
lid u = 0.1.  A step is one fused pass over the lattice (pull-stream,
boundaries, macroscopic update, collision) — the reference loop body.

Prints ONE JSON line: MLUPS = fluid nodes x steps / wall seconds / 1e6, the
HBM-roofline fraction of the dominant (bulk) kernel from HIP events recorded
around each of its launches on the solver's stream, and the CPU oracle timed on
a bounded sample of the same workload (1 core).

  python bench.py [--gpus 1] [--steps 2000] [--warmup 50]
                  [--N 4097] [--collision mrt|bgk|trt] [--precision f32|f64|f16]
                  [--no-cpu-baseline]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "MLUPS (% HBM roofline) at 4097² MRT; time-to-1e-6 residual, learned vs rest"
HBM_PEAK_GBS = 8000.0
PMC_SUMMARY = os.path.join(ROOT, "profiles", "k_step_traffic_pmc.json")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(N, collision, visc, rates, steps=3):
    """The oracle (C restatement of the reference serial solver: 528-byte AoS
    cells, six sweeps per step, FP64) on 1 core, same workload, `steps` timed
    iterations after one untimed one."""
    import numpy as np  # noqa: F401
    sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
    import meshgen
    import oracle as O
    nodes, conn = meshgen.cavity_arrays_fast(N)
    kw = dict(collision=collision, init_rest=1, uavg=0.1, rho_ini=1.0, viscosity=visc,
              resid_kind=3)
    if collision == 3:
        kw.update(mrt_textbook=1, mrt_s=rates)
    if collision == 2:
        kw.update(trt_lambda=0.25, trt_fp64=1)
    p = O.params(**kw)
    o = O.Oracle(nodes, conn, p)
    del nodes, conn
    o.step(1)
    t0 = time.perf_counter()
    o.step(steps)
    dt = time.perf_counter() - t0
    nf = N * N
    return {"value": round(nf * steps / dt / 1e6, 3), "unit": "MLUPS", "cores": 1, "kind": "port",
            "sample": f"{N}x{N} cavity, same collision/rates, FP64, {steps} timed steps after 1 "
                      f"untimed (oracle/lbm_oracle.c, 1 thread, {dt:.1f} s)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--N", type=int, default=4097)
    ap.add_argument("--collision", default="mrt", choices=["bgk", "trt", "mrt"])
    ap.add_argument("--precision", default="f32", choices=["f64", "f32", "f16"])
    ap.add_argument("--block", type=int, default=0)
    ap.add_argument("--xcd-remap", action="store_true")
    ap.add_argument("--split-boundary", action="store_true")
    ap.add_argument("--xcd-group", type=int, default=0)
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--no-pairs", action="store_true",
                    help="FP16: one cell per thread (no paired-cell kernel)")
    ap.add_argument("--sample-every", type=int, default=8,
                    help="HIP events around every N-th step-kernel launch")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=10)
    ap.add_argument("--workload", default="cavity",
                    choices=["cavity", "sweep", "warmstart", "netstep", "matrix", "cavity129",
                             "cylinder", "io"],
                    help="cavity: the single 4097^2 lattice; sweep: the batched cavity sweep; "
                         "warmstart: learned start vs rest start on held-out cavities; "
                         "netstep: training steps of the warm-start network; "
                         "matrix: every collision x storage at --sizes, error vs FP64; "
                         "cavity129: the 129^2 Re=100 cavity beside the reference program; "
                         "cylinder: the 2049x513 cylinder wake, Strouhal and drag")
    ap.add_argument("--count", type=int, default=256, help="sweep: number of cavities")
    ap.add_argument("--side", default="65,513", help="sweep: side-length range")
    ap.add_argument("--crit", type=float, default=1e-6, help="sweep: residual criterion")
    ap.add_argument("--max-steps", type=int, default=400000, help="sweep: step cap")
    ap.add_argument("--single-sample", type=int, default=8,
                    help="sweep: cavities also run one at a time")
    ap.add_argument("--io-final", action="store_true",
                    help="io: also time the final-layout CSV (f0..f8 added, about 2.5x the size)")
    ap.add_argument("--train-count", type=int, default=2000, help="warmstart: training cavities")
    ap.add_argument("--test-count", type=int, default=200, help="warmstart: held-out cavities")
    ap.add_argument("--epochs", type=int, default=120, help="warmstart: training epochs")
    ap.add_argument("--batch", type=int, default=16, help="warmstart: training batch")
    ap.add_argument("--lr", type=float, default=1e-3, help="warmstart: peak learning rate")
    ap.add_argument("--rho-weight", type=float, default=None,
                    help="warmstart: density weight in the loss (default: the package's)")
    ap.add_argument("--weight-decay", type=float, default=None,
                    help="warmstart: decoupled weight decay (default: the package's)")
    ap.add_argument("--rho-clip", type=float, default=None,
                    help="warmstart: clip the density target to these quantiles (0: off; "
                         "default: the package's)")
    ap.add_argument("--sizes", default="1025,4097", help="matrix: cavity sizes")
    ap.add_argument("--save-net", default="", help="warmstart: write the trained weights here")
    ap.add_argument("--data-max-steps", type=int, default=200000,
                    help="warmstart: step cap of the training-data runs (unconverged dropped)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only; --gpus {args.gpus} is not supported "
              "(multi-GPU slab decomposition is a design note, not built).", file=sys.stderr)
        sys.exit(2)

    import parallellbmcranfield_amd as lbm
    if args.workload == "sweep":
        return bench_sweep(args)
    if args.workload == "warmstart":
        return bench_warmstart(args)
    if args.workload == "netstep":
        return bench_netstep(args)
    if args.workload == "matrix":
        return bench_matrix(args)
    if args.workload == "cavity129":
        return bench_cavity129(args)
    if args.workload == "cylinder":
        return bench_cylinder(args)
    if args.workload == "io":
        return bench_io(args)

    N = args.N
    coll = {"bgk": lbm.BGKW, "trt": lbm.TRT, "mrt": lbm.MRT}[args.collision]
    prec = {"f64": lbm.PREC_F64, "f32": lbm.PREC_F32, "f16": lbm.PREC_F16}[args.precision]
    visc = 0.1 * (N - 1) / 1000.0  # Re = 1000 on L = N - 1, lid 0.1
    tau = 3 * float(__import__("numpy").float32(visc)) + 0.5
    rates = [0.0, 1.4, 1.4, 0.0, 1.2, 0.0, 1.2, 1 / tau, 1 / tau]
    kw = dict(collision=coll, uavg=0.1, rho_ini=1.0, viscosity=visc, init=lbm.INIT_REST,
              resid_kind=lbm.RESID_NONE, use_stop_rule=0, block_size=args.block,
              tune_flags=(lbm.TUNE_XCD_REMAP if args.xcd_remap else 0) |
              (lbm.TUNE_SPLIT_BOUNDARY if args.split_boundary else 0) |
              (lbm.TUNE_NO_GRAPH if args.no_graph else 0) |
              (lbm.TUNE_NO_PAIRS if args.no_pairs else 0),
              xcd_group=args.xcd_group)
    if coll == lbm.MRT:
        kw.update(mrt_textbook_meq=1, mrt_rates=rates)
    if coll == lbm.TRT:
        kw.update(trt_lambda=0.25, trt_fp64_parts=1)
    log(f"bench: building case {N}x{N} ...")
    t = time.perf_counter()
    case = lbm.Case.cavity(N)
    opts = lbm.make_options(**kw)
    opts.precision = prec
    s = lbm.Solver(case, opts)
    log(f"bench: set-up {time.perf_counter() - t:.1f} s, device memory "
        f"{s.device_bytes() / 2**30:.2f} GiB, fluid nodes {case.info.num_fluid}, "
        f"boundary-list nodes {case.info.num_boundary}")
    s.step(args.warmup)
    s.sync()
    s.sync()
    t0 = time.perf_counter()
    total_ms, kern_ms = s.step_timed(args.steps, args.sample_every)
    s.sync()
    t1 = time.perf_counter()
    wall = t1 - t0
    nf = case.info.num_fluid
    fused = s.fused
    # nodes one launch of the timed kernel updates: the fused launch covers
    # the boundary-list nodes too
    n_bulk = nf if fused else nf - case.info.num_boundary
    mlups = nf * args.steps / wall / 1e6
    esize = {lbm.PREC_F64: 8, lbm.PREC_F32: 4, lbm.PREC_F16: 2}[prec]
    bytes_per_launch = 2 * 9 * esize * n_bulk
    avg_bulk_s = kern_ms / 1e3
    achieved = bytes_per_launch / avg_bulk_s / 1e9
    traffic = None
    default_shape = (args.block in (0, 256) and not args.xcd_remap and not args.split_boundary
                     and args.xcd_group in (0, 4) and not args.no_pairs)
    if default_shape and os.path.exists(PMC_SUMMARY):
        try:
            pm = json.load(open(PMC_SUMMARY))
            key = f"{args.collision}_{args.precision}_{N}"
            if key in pm.get("kernels", {}):
                traffic = pm["kernels"][key]["hbm_bytes_per_launch"]
        except (OSError, ValueError, KeyError):
            traffic = None
    log(f"bench: {args.steps} steps in {wall * 1e3:.1f} ms (stream {total_ms:.1f} ms), "
        f"bulk kernel avg {avg_bulk_s * 1e6:.1f} us -> {achieved:.0f} GB/s, {mlups:.0f} MLUPS")
    s.close()
    del case
    copy_gbs = lbm.device_copy_gbs(1 << 30, 10)
    log(f"bench: plain device copy {copy_gbs:.0f} GB/s")
    cpu = None
    if not args.no_cpu_baseline:
        log("bench: timing the CPU oracle on a bounded sample (1 core) ...")
        cpu = cpu_baseline(N, coll, visc, rates, steps=args.cpu_steps)
        log(f"bench: cpu baseline {cpu['value']} MLUPS")
    out = {
        "metric": METRIC,
        "value": round(mlups, 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall * 1e3 / args.steps, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"f64": "f64", "f32": "f32", "f16": "f16-storage/f32"}[args.precision],
        "data": "synthetic",
        "config": {
            "workload": f"lid-driven cavity {N}x{N} fluid nodes (reference mesh convention), "
                        f"D2Q9 {args.collision.upper()}, Re=1000, lid u=0.1, rest start, "
                        f"{args.precision} storage" + (" of f - w*rho0" if prec else ""),
            "fluid_nodes": nf,
            "nodes_per_launch": n_bulk,
            "collision": args.collision,
            "precision": args.precision,
            "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "kernel": "k_step (bulk tiles + boundary-list nodes)" if fused else "k_bulk",
            "bytes_per_launch": bytes_per_launch,
            "avg_launch_us": round(avg_bulk_s * 1e6, 2),
            "launches_timed": (args.steps + args.sample_every - 1) // args.sample_every,
            "copy_kernel_gbs": round(copy_gbs, 1),
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def bench_sweep(args):
    """Batched sweep: `count` seeded cavities (sides uniform in --side, aspect
    0.5-2, Re log-uniform 100-3200 on the lid, 0-3 solid rectangles of 5-25 %
    of the side), MRT, FP32 storage of f - w*rho0, rest start, each run to
    its own velocity residual --crit in one batched solver; a sample of the
    same cavities is then run one at a time with the single-lattice solver."""
    import numpy as np
    import parallellbmcranfield_amd as lbm
    from parallellbmcranfield_amd import sweep as SW
    lo, hi = (int(x) for x in args.side.split(","))
    specs = SW.random_cavities(args.count, seed=0, side=(lo, hi))
    opts = SW.sweep_options(crit=args.crit, max_steps=args.max_steps)
    packed, fl = SW.build_batch(specs)
    inst = [(sp["nu"], sp["lid"]) for sp in specs]
    log(f"sweep: {len(specs)} cavities packed into {packed.n}x{packed.m} "
        f"({sum(fl)} fluid nodes, fill {sum(fl) / (packed.n * packed.m):.2f})")
    best, tried = SW.autotune(packed, opts, instances=inst, steps=100)
    log(f"sweep: autotune {tried} -> {best}")
    opts.block_size, opts.xcd_group = best["block_size"], best["xcd_group"]
    del packed

    def progress(it, done, tot):
        log(f"sweep: iteration {it}, {done}/{tot} instances converged")

    r = SW.run_batched(specs, opts, chunk=5000, progress=progress)
    it = r["conv_iter"]
    conv = it > 0
    lat = r["conv_s"][conv]
    log(f"sweep: batched {r['wall_s']:.1f} s, {r['mlups']:.0f} MLUPS aggregate, "
        f"{conv.sum()}/{len(specs)} converged")
    # one at a time: a size-stratified sample of the same cavities
    order = np.argsort(r["fluid"])
    pick = order[np.linspace(0, len(order) - 1, min(args.single_sample, len(order))).astype(int)]
    singles = []
    for k in pick:
        sr = SW.run_single(specs[int(k)], opts)
        singles.append(dict(index=int(k), steps=sr["iters"], batched_steps=int(it[k]),
                            wall_s=round(sr["wall_s"], 3), mlups=round(sr["mlups"], 1)))
        log(f"sweep: single {k}: {sr['iters']} steps (batched {it[k]}), {sr['wall_s']:.2f} s")
    one_work = sum(fl_k * s_["steps"] for fl_k, s_ in zip(r["fluid"][pick], singles))
    one_wall = sum(s_["wall_s"] for s_ in singles)
    one_mlups = one_work / one_wall / 1e6
    bytes_per_update = 2 * 9 * 4
    achieved = r["mlups"] * 1e6 * bytes_per_update / 1e9
    out = {
        "metric": "batched cavity sweep: aggregate MLUPS to per-instance residual 1e-6",
        "value": round(r["mlups"], 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": int(r["iters"]),
        "warmup": 0,
        "ms_per_step": round(r["wall_s"] * 1e3 / max(1, r["iters"]), 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"{len(specs)} lid-driven cavities, sides {lo}-{hi}, aspect 0.5-2, "
                        f"Re 100-3200 log-uniform, lid 0.1, 0-3 solid rectangles, MRT, "
                        f"f32 storage, rest start, residual {args.crit:g} every "
                        f"{opts.resid_every} steps, seed 0",
            "packed_lattice": list(r["lattice"]),
            "fluid_nodes": int(sum(fl)),
            "autotune": best,
            "parallelism": "single GPU",
        },
        "per_instance": {
            "converged": int(conv.sum()),
            "steps_median": int(np.median(it[conv])) if conv.any() else None,
            "steps_max": int(it[conv].max()) if conv.any() else None,
            "latency_median_s": round(float(np.median(lat)), 3) if conv.any() else None,
            "latency_max_s": round(float(lat.max()), 3) if conv.any() else None,
        },
        "one_at_a_time": {
            "sample": singles,
            "mlups": round(one_mlups, 1),
            "batched_speedup": round(r["mlups"] / one_mlups, 2),
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": None,
            "kernel": "k_step over the packed lattice (useful updates only)",
            "bytes_per_update": bytes_per_update,
        },
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def bench_cylinder(args):
    """Channel of 2049 x 513 fluid nodes with a D = 64 cylinder at (512, 260)
    (analytic-q curved links), parabolic inlet u_max = 0.05, rho = 1 pressure
    outlet, MRT, Re = 100 on the mean inlet speed; 100 000 steps in FP64 and
    FP32 with the momentum-exchange force every 10 steps: MLUPS, Strouhal
    number and drag coefficient of each, and the CPU oracle on a bounded
    sample of the same FP64 case (1 core)."""
    import numpy as np
    import parallellbmcranfield_amd as lbm
    NX, NY, D, steps, every = 2049, 513, 64, 100000, 10
    case = lbm.Case.channel(NX, NY, 512.0, 260.0, D / 2)
    umax = 0.05
    uavg = umax / 1.5
    nu = uavg * D / 100.0
    tau = 3 * float(np.float32(nu)) + 0.5
    rates = [0, 1.4, 1.4, 0, 1.2, 0, 1.2, 1 / tau, 1 / tau]
    nf = case.info.num_fluid
    res = {}
    for pname, prec in (("f64", lbm.PREC_F64), ("f32", lbm.PREC_F32)):
        o = lbm.make_options(collision=lbm.MRT, uavg=uavg, rho_ini=1.0, viscosity=nu,
                             inlet_profile=1, curved=1, outlet_profile=lbm.OUTLET_PRESSURE,
                             init=lbm.INIT_REST, resid_kind=lbm.RESID_NONE, mrt_textbook_meq=1,
                             mrt_rates=rates, force_id=5, force_every=every, precision=prec,
                             outlet_rho=1.0, use_stop_rule=0)
        s = lbm.Solver(case, o)
        s.sync()
        t0 = time.perf_counter()
        s.step(steps)
        s.sync()
        wall = time.perf_counter() - t0
        _, f = s.force_history()
        s.close()
        fx, fy = f[:, 0], f[:, 1]
        half = len(fy) // 2
        st = lbm.strouhal(fy, every, D, uavg)
        cd = 2 * fx[half:].mean() / (uavg * uavg * D)
        res[pname] = dict(mlups=round(nf * steps / wall / 1e6, 1), wall_s=round(wall, 2),
                          strouhal=round(float(st), 5), cd=round(float(cd), 5))
        log(f"cylinder {pname}: {res[pname]}")
    cpu = None
    if not args.no_cpu_baseline:
        import shutil
        import tempfile
        import oracle as O
        d = tempfile.mkdtemp(prefix="lbm_cylinder_")
        case.write(os.path.join(d, "n.dat"), os.path.join(d, "c.dat"))
        p = O.params(collision=3, uavg=uavg, rho_ini=1.0, viscosity=nu, inlet_profile=1,
                     curved=1, outlet_profile=4, init_rest=1, mrt_textbook=1, mrt_s=rates,
                     resid_kind=3, outlet_rho=1.0)
        orc = O.Oracle.from_files(os.path.join(d, "n.dat"), os.path.join(d, "c.dat"), p)
        shutil.rmtree(d, ignore_errors=True)
        orc.step(1)
        n_cpu = 10
        t0 = time.perf_counter()
        orc.step(n_cpu)
        dt = time.perf_counter() - t0
        cpu = {"value": round(nf * n_cpu / dt / 1e6, 3), "unit": "MLUPS", "cores": 1,
               "kind": "port", "sample": f"the same FP64 channel, {n_cpu} steps after 1 "
                                          f"(oracle/lbm_oracle.c, 1 thread, {dt:.1f} s)"}
    out = {
        "metric": "2049x513 cylinder wake, MRT, Re=100: MLUPS, Strouhal, drag",
        "value": res["f32"]["mlups"], "unit": "MLUPS", "n_gpus": 1, "steps": steps, "warmup": 0,
        "ms_per_step": round(res["f32"]["wall_s"] * 1e3 / steps, 5), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32 (f64 run beside it)",
        "data": "synthetic",
        "config": {"workload": "channel 2049x513 fluid nodes, cylinder D=64 at (512,260) with "
                               "analytic-q curved links, parabolic inlet u_max=0.05, rho=1 "
                               "pressure outlet, MRT, Re=100 on the mean inlet speed, 100000 "
                               "steps, force every 10", "fluid_nodes": nf,
                   "parallelism": "single GPU"},
        "runs": res,
        "strouhal_rel_diff_f32_vs_f64": abs(res["f32"]["strouhal"] - res["f64"]["strouhal"]) /
        res["f64"]["strouhal"],
        "cd_rel_diff_f32_vs_f64": abs(res["f32"]["cd"] - res["f64"]["cd"]) / abs(res["f64"]["cd"]),
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def bench_io(args):
    """Result files and checkpoints at scale: a 4097^2 cavity (MRT, FP32
    storage) after a few steps; times WriteResults-style CSV and Tecplot
    files (x, y, u, v, |u|, rho, p, fluid per cell), the final-layout CSV
    (plus f0..f8), the binary checkpoint save + load, each on the host of
    the GPU in use (device-to-host copy included), and the mesh readers on
    the cavity's own D2node.dat / BCconnectors.dat.  CPU baseline: the
    oracle's sequential fprintf writer (the reference's WriteResults,
    serial/FileIO.c:298-348) on a 1025^2 cavity, 1 thread."""
    import tempfile
    import parallellbmcranfield_amd as lbm
    N = args.N
    visc = 0.1 * (N - 1) / 1000.0
    tau = 3 * float(__import__("numpy").float32(visc)) + 0.5
    rates = [0.0, 1.4, 1.4, 0.0, 1.2, 0.0, 1.2, 1 / tau, 1 / tau]
    o = lbm.make_options(collision=lbm.MRT, uavg=0.1, rho_ini=1.0, viscosity=visc,
                         init=lbm.INIT_REST, resid_kind=lbm.RESID_NONE, use_stop_rule=0,
                         mrt_textbook_meq=1, mrt_rates=rates)
    o.precision = lbm.PREC_F32
    case = lbm.Case.cavity(N)
    s = lbm.Solver(case, o)
    s.step(100)
    s.sync()
    res = {}
    with tempfile.TemporaryDirectory(dir=os.environ.get("TMPDIR")) as d:
        for name, fmt, final in (("csv", lbm.FORMAT_CSV, False), ("tecplot", lbm.FORMAT_TECPLOT, False),
                                 ("csv_final", lbm.FORMAT_CSV, True)):
            if final and not args.io_final:
                continue
            path = os.path.join(d, name)
            t0 = time.perf_counter()
            s.write_results(path, fmt, final=final)
            dt = time.perf_counter() - t0
            size = os.path.getsize(path)
            os.remove(path)
            res[name] = {"seconds": round(dt, 3), "bytes": size, "MB_per_s": round(size / dt / 1e6, 1)}
            log(f"io: {name} {size / 1e9:.2f} GB in {dt:.2f} s ({size / dt / 1e6:.0f} MB/s)")
        path = os.path.join(d, "ckpt")
        t0 = time.perf_counter()
        s.save(path)
        t_save = time.perf_counter() - t0
        size = os.path.getsize(path)
        t0 = time.perf_counter()
        s.load(path)
        s.sync()
        t_load = time.perf_counter() - t0
        os.remove(path)
        res["checkpoint"] = {"bytes": size, "save_s": round(t_save, 3), "load_s": round(t_load, 3),
                             "save_GB_per_s": round(size / t_save / 1e9, 2),
                             "load_GB_per_s": round(size / t_load / 1e9, 2)}
        log(f"io: checkpoint {size / 1e9:.2f} GB, save {t_save:.2f} s, load {t_load:.2f} s")
    s.close()
    # the mesh readers (ReadNodes + ReadBCconn + CellIni) on the same 4097^2
    # cavity's mesh files in the reference's text format
    with tempfile.TemporaryDirectory(dir=os.environ.get("TMPDIR")) as d:
        pn, pc = os.path.join(d, "D2node.dat"), os.path.join(d, "BCconnectors.dat")
        case.write(pn, pc)
        size = os.path.getsize(pn) + os.path.getsize(pc)
        t0 = time.perf_counter()
        cm = lbm.Case.load(pn, pc)
        dt = time.perf_counter() - t0
        assert cm.info.num_fluid == case.info.num_fluid
        del cm
    res["mesh_read"] = {"bytes": size, "seconds": round(dt, 3), "MB_per_s": round(size / dt / 1e6, 1)}
    log(f"io: mesh read {size / 1e6:.0f} MB in {dt:.2f} s ({size / dt / 1e6:.0f} MB/s)")
    cpu = None
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
        import meshgen
        import oracle as O
        Nc = 1025
        nodes, conn = meshgen.cavity_arrays_fast(Nc)
        orc = O.Oracle(nodes, conn, O.params(collision=3, init_rest=1, uavg=0.1, rho_ini=1.0,
                                             viscosity=0.1 * (Nc - 1) / 1000.0))
        del nodes, conn
        orc.step(1)
        with tempfile.TemporaryDirectory(dir=os.environ.get("TMPDIR")) as d:
            path = os.path.join(d, "o.csv")
            t0 = time.perf_counter()
            orc.write_results(path, final=False)
            dt = time.perf_counter() - t0
            size = os.path.getsize(path)
        cpu = {"value": round(size / dt / 1e6, 1), "unit": "MB/s", "cores": 1, "kind": "port",
               "sample": f"{Nc}x{Nc} cavity results CSV ({size / 1e6:.0f} MB) by the oracle's "
                         f"sequential fprintf writer (oracle/lbm_oracle.c, 1 thread, {dt:.1f} s)"}
        log(f"io: cpu baseline {cpu['value']} MB/s")
    out = {
        "metric": "results CSV writer throughput",
        "value": res["csv"]["MB_per_s"],
        "unit": "MB/s",
        "n_gpus": 1,
        "steps": None,
        "warmup": 0,
        "ms_per_step": None,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "text (f32 lattice)",
        "data": "synthetic",
        "config": {"workload": f"lid-driven cavity {N}x{N} fluid nodes, MRT, FP32 storage, after "
                               f"100 steps; results files and checkpoint on the host's temp dir",
                   "parallelism": "single GPU + host threads"},
        "io": res,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def bench_cavity129(args):
    """The shipped-convention lid-driven cavity with 129 x 129 fluid nodes,
    BGK, FP64, lid u = 0.1, Re = 100 (nu = 0.128), rest start, up to 10 000
    steps or velocity residual 1e-8 (checked every 100 steps) on the GPU;
    beside it the reference program built from its own sources
    (oracle/_ref/LBMSolver, 1 host thread) on the same mesh for the same
    number of steps (its own start and residual; same work per step)."""
    import shutil
    import subprocess
    import tempfile
    import parallellbmcranfield_amd as lbm
    N, steps_cap = 129, 10000
    case = lbm.Case.cavity(N)
    o = lbm.make_options(collision=lbm.BGKW, uavg=0.1, rho_ini=1.0, viscosity=0.128,
                         init=lbm.INIT_REST, resid_kind=lbm.RESID_VELOCITY, resid_every=100,
                         crit_v=1e-8, use_stop_rule=1, max_iterations=steps_cap)
    o.precision = lbm.PREC_F64
    s = lbm.Solver(case, o)
    s.sync()
    t0 = time.perf_counter()
    s.step(steps_cap)
    s.sync()
    wall = time.perf_counter() - t0
    steps = s.iter
    res = float(s.residuals()[0])
    nf = case.info.num_fluid
    # step-kernel duration on the same lattice (plain steps, events)
    o2 = lbm.make_options(collision=lbm.BGKW, uavg=0.1, rho_ini=1.0, viscosity=0.128,
                          init=lbm.INIT_REST, resid_kind=lbm.RESID_NONE, use_stop_rule=0)
    o2.precision = lbm.PREC_F64
    s2 = lbm.Solver(case, o2)
    s2.step(50)
    _, kern_ms = s2.step_timed(2000, 8)
    s2.close()
    achieved = 2 * 9 * 8 * nf / (kern_ms / 1e3) / 1e9
    cpu = None
    ref = os.path.join(ROOT, "oracle", "_ref", "LBMSolver")
    if not args.no_cpu_baseline and os.path.exists(ref):
        d = tempfile.mkdtemp(prefix="lbm_cavity129_")
        try:
            os.makedirs(os.path.join(d, "Mesh"))
            os.makedirs(os.path.join(d, "Results"))
            case.write(os.path.join(d, "Mesh", "D2node.dat"),
                       os.path.join(d, "Mesh", "BCconnectors.dat"))
            with open(os.path.join(d, "SetUpData.ini"), "w") as fh:
                fh.write(f"0.1\n0.0\n1.0\n0.128\n2\n1\n2\n2\n{steps}\n1000000\n1000000\n1\n0\n"
                         "-1.0\n-1.0\n0\n")
            log(f"cavity129: reference program, {steps} steps on 1 host thread ...")
            t1 = time.perf_counter()
            subprocess.run([ref], cwd=d, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL,
                           check=True, timeout=900)
            dt = time.perf_counter() - t1
            cpu = {"value": round(nf * steps / dt / 1e6, 3), "unit": "MLUPS", "cores": 1,
                   "kind": "reference",
                   "sample": f"the reference program (serial/, built from its sources into "
                             f"oracle/_ref) on the same 129^2 mesh, {steps} steps, BGK, "
                             f"whole run incl. set-up and result files: {dt:.1f} s"}
        finally:
            shutil.rmtree(d, ignore_errors=True)
    out = {
        "metric": "129^2 cavity BGK FP64 Re=100: MLUPS to 10 000 steps or residual 1e-8",
        "value": round(nf * steps / wall / 1e6, 1), "unit": "MLUPS", "n_gpus": 1, "steps": steps,
        "warmup": 0, "ms_per_step": round(wall * 1e3 / max(1, steps), 5),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "lid-driven cavity 129x129 fluid nodes (reference mesh convention), "
                               "BGK, FP64, lid u=0.1, nu=0.128 (Re=100), rest start, velocity "
                               "residual every 100 steps", "fluid_nodes": nf,
                   "final_velocity_residual": res, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "avg_launch_us": round(kern_ms * 1e3, 2)},
        "cpu_baseline": cpu,
    }
    s.close()
    print(json.dumps(out), flush=True)


def bench_matrix(args):
    """The cavity at each of --sizes (Re = 1000 on L = N - 1, lid 0.1, rest
    start) for BGK, TRT (magic parameter 1/4) and MRT (Lallemand-Luo rates,
    1/tau on the stress moments) in FP64, FP32 and FP16 storage: --steps
    timed steps each, MLUPS, step-kernel HBM fraction, and the velocity
    relative L2 error of the reduced-precision fields against the FP64 run
    of the same model after the same steps."""
    import numpy as np
    import parallellbmcranfield_amd as lbm
    rows = []
    for N in (int(x) for x in args.sizes.split(",")):
        case = lbm.Case.cavity(N)
        nf = case.info.num_fluid
        visc = 0.1 * (N - 1) / 1000.0
        tau = 3 * float(np.float32(visc)) + 0.5
        rates = [0.0, 1.4, 1.4, 0.0, 1.2, 0.0, 1.2, 1 / tau, 1 / tau]
        for name, coll in (("bgk", lbm.BGKW), ("trt", lbm.TRT), ("mrt", lbm.MRT)):
            ref = None
            for pname, prec, esize in (("f64", lbm.PREC_F64, 8), ("f32", lbm.PREC_F32, 4),
                                       ("f16", lbm.PREC_F16, 2)):
                kw = dict(collision=coll, uavg=0.1, rho_ini=1.0, viscosity=visc,
                          init=lbm.INIT_REST, resid_kind=lbm.RESID_NONE, use_stop_rule=0)
                if coll == lbm.MRT:
                    kw.update(mrt_textbook_meq=1, mrt_rates=rates)
                if coll == lbm.TRT:
                    kw.update(trt_lambda=0.25, trt_fp64_parts=1)
                o = lbm.make_options(**kw)
                o.precision = prec
                s = lbm.Solver(case, o)
                s.step(args.warmup)
                s.sync()
                t0 = time.perf_counter()
                _, kern_ms = s.step_timed(args.steps, args.sample_every)
                s.sync()
                wall = time.perf_counter() - t0
                _, u, v = s.macro()
                s.close()
                if ref is None:
                    ref = (u, v)
                    err = 0.0
                else:
                    err = float(np.sqrt(((u - ref[0]) ** 2 + (v - ref[1]) ** 2).sum() /
                                        (ref[0] ** 2 + ref[1] ** 2).sum()))
                gbs = 2 * 9 * esize * nf / (kern_ms / 1e3) / 1e9
                row = dict(N=N, collision=name, storage=pname,
                           mlups=round(nf * args.steps / wall / 1e6, 1),
                           kernel_us=round(kern_ms * 1e3, 2), hbm_frac=round(gbs / HBM_PEAK_GBS, 4),
                           vel_rel_l2_vs_f64=err)
                rows.append(row)
                log(f"matrix: {row}")
            del ref
        del case
    head = [r for r in rows if r["N"] == max(r_["N"] for r_ in rows) and r["collision"] == "mrt"
            and r["storage"] == "f32"]
    out = {
        "metric": "MLUPS per collision x storage, velocity error vs FP64",
        "value": head[0]["mlups"] if head else None, "unit": "MLUPS", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": None,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64/f32/f16",
        "data": "synthetic",
        "config": {"workload": f"lid-driven cavities {args.sizes}, Re=1000, lid 0.1, rest start, "
                               "BGK / TRT(1/4) / MRT(Lallemand-Luo) x FP64 / FP32 / FP16 storage",
                   "parallelism": "single GPU"},
        "rows": rows, "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def net_flops(H, W, batch, base=32, levels=4):
    """FLOP of one training step (forward, weight and data gradients; the
    first conv needs no data gradient) of the warm-start network."""
    import sys as _s
    _s.path.insert(0, os.path.join(ROOT))
    Hl, Wl = [H], [W]
    for _ in range(levels - 1):
        Hl.append((Hl[-1] + 1) // 2)
        Wl.append((Wl[-1] + 1) // 2)
    C = [base << l for l in range(levels)]
    convs = []
    for l in range(levels):
        convs.append((32 if l == 0 else C[l - 1], C[l], l))
        convs.append((C[l], C[l], l))
    for l in range(levels - 2, -1, -1):
        convs.append((C[l] + C[l + 1], C[l], l))
        convs.append((C[l], C[l], l))
    fwd = sum(2 * 9 * ci * co * batch * Hl[l] * Wl[l] for ci, co, l in convs)
    dgrad = fwd - 2 * 9 * 32 * C[0] * batch * H * W
    return fwd, fwd + fwd + dgrad


def bench_netstep(args):
    """Adam training steps of the warm-start network on a synthetic batch
    (257^2 grid, 4 levels, 32-256 channels, batch --batch), timed over
    --steps after --warmup; FLOP/s of the convolutions against the MI355X
    dense BF16 matrix peak."""
    import numpy as np
    import parallellbmcranfield_amd as lbm
    G, B = 257, args.batch
    net = lbm.Net(G, G, batch=B, base=32, levels=4, seed=0)
    rng = np.random.default_rng(0)
    inp = np.zeros((B, G, G, 3), np.float32)
    inp[..., 0] = 1.0
    inp[..., 1:] = rng.random((B, 1, 1, 2))
    tgt = rng.standard_normal((B, G, G, 3)).astype(np.float32)
    net.set_data(inp, tgt)
    idx = np.arange(B)
    for _ in range(args.warmup):
        net.train(idx, 1e-4)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        net.train(idx, 1e-4)
    dt = (time.perf_counter() - t0) / args.steps
    fwd, tot = net_flops(G, G, B)
    tflops = tot / dt / 1e12
    out = {
        "metric": "warm-start network training step: convolution TFLOP/s",
        "value": round(tflops, 2), "unit": "TFLOP/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(dt * 1e3, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "bf16 (fp32 accumulate)",
        "data": "synthetic",
        "config": {"workload": f"U-Net 257x257, 4 levels, 32-256 channels, batch {B}, "
                               "forward + backward + Adam", "params": net.num_params,
                   "flop_per_step": tot, "parallelism": "single GPU"},
        "roofline": {"bound": "mfma", "achieved": round(tflops, 2), "peak": 2500.0,
                     "unit": "TFLOP/s", "frac": round(tflops / 2500.0, 4), "traffic": None},
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def bench_warmstart(args):
    """Learned warm start: steady states of --train-count seeded cavities
    (sweep geometry, Re log-uniform 100-5000) from the batched solver train
    the network (257^2 grid, 4 levels, 32-256 channels) on the GPU; then
    --test-count held-out cavities run to the residual criterion from rest
    and from the prediction."""
    import numpy as np
    import parallellbmcranfield_amd as lbm
    from parallellbmcranfield_amd import sweep as SW
    from parallellbmcranfield_amd import warmstart as WS
    lo, hi = (int(x) for x in args.side.split(","))
    opts = SW.sweep_options(crit=args.crit, max_steps=args.max_steps)
    G = WS.GRID
    t0 = time.perf_counter()
    specs = SW.random_cavities(args.train_count, seed=1, side=(lo, hi), re=(100.0, 5000.0))
    X, T = [], []
    per = 500
    dopts = SW.sweep_options(crit=args.crit, max_steps=args.data_max_steps)
    for b0 in range(0, len(specs), per):
        part = specs[b0:b0 + per]
        fields, it, r = WS.steady_states(part, dopts)
        for s_, f, k in zip(part, fields, it):
            if k > 0:
                X.append(WS.encode(s_, G))
                T.append(WS.field_to_grid(s_, *f, G))
        log(f"warmstart: data {b0 + len(part)}/{len(specs)}: batch {r['wall_s']:.1f} s, "
            f"{(it > 0).sum()}/{len(part)} converged")
    X, T = np.stack(X), np.stack(T)
    t_data = time.perf_counter() - t0
    net = lbm.Net(G, G, batch=args.batch, base=32, levels=4, seed=0, in_channels=WS.IN_CH,
                  dilation=WS.DILATION)
    t1 = time.perf_counter()
    rho_w = WS.RHO_WEIGHT if args.rho_weight is None else args.rho_weight
    wd = WS.WEIGHT_DECAY if args.weight_decay is None else args.weight_decay
    rho_clip = WS.RHO_CLIP if args.rho_clip is None else args.rho_clip
    hist = WS.train(net, X, T, epochs=args.epochs, lr=args.lr, log=log, rho_weight=rho_w,
                    weight_decay=wd, rho_clip=rho_clip)
    t_train = time.perf_counter() - t1
    if args.save_net:
        net.save(args.save_net)
        np.save(args.save_net + ".scales.npy", net.target_scales)
    del X, T
    test = SW.random_cavities(args.test_count, seed=2, side=(lo, hi), re=(100.0, 5000.0))
    r = WS.compare_starts(net, test, opts, log=log)
    r1 = WS.compare_starts(net, test, opts, log=log, density=False, run_rest=r["rest_run"])
    both = (r["steps_rest"] > 0) & (r["steps_warm"] > 0)
    ratio_t = r["median_time_rest"] / r["median_time_warm"] if both.any() else None
    log(f"warmstart: median steps ratio {r['median_steps_ratio']}, median time ratio {ratio_t}")
    out = {
        "metric": "learned warm start: median wall time to residual 1e-6, rest / warm",
        "value": round(ratio_t, 3) if ratio_t else None,
        "unit": "x",
        "n_gpus": 1,
        "steps": None,
        "warmup": 0,
        "ms_per_step": None,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16 network (fp32 accumulate) / f32 lattice",
        "data": "synthetic",
        "config": {
            "workload": f"train {args.train_count} seeded cavities (sides {lo}-{hi}, aspect "
                        f"0.5-2, Re 100-5000 log-uniform, 0-3 rectangles, seed 1), test "
                        f"{args.test_count} held-out (seed 2); 257^2 grid U-Net, 4 levels, "
                        f"32-256 channels; MRT f32 lattice, residual {args.crit:g}",
            "params": net.num_params,
            "bottleneck_dilation": WS.DILATION,
            "input_channels": WS.IN_CH,
            "rho_loss_weight": rho_w,
            "weight_decay": wd,
            "rho_clip": rho_clip,
            "training_samples": int(len(hist) and args.train_count),
            "epochs": args.epochs,
            "batch": args.batch,
            "parallelism": "single GPU",
        },
        "warmstart": {
            "both_converged": r["both_converged"],
            "median_steps_rest": float(np.median(r["steps_rest"][both])) if both.any() else None,
            "median_steps_warm": float(np.median(r["steps_warm"][both])) if both.any() else None,
            "median_steps_ratio": r["median_steps_ratio"],
            "median_time_rest_s": r["median_time_rest"],
            "median_time_warm_s": r["median_time_warm"],
            "batch_wall_rest_s": round(r["wall_rest"], 2),
            "batch_wall_warm_s": round(r["wall_warm"], 2),
            "predict_s": round(r["predict_s"], 3),
            "field_rel_l2_median": r["field_rel_l2_median"],
            "uniform_density_seed": {"median_steps_ratio": r1["median_steps_ratio"],
                                     "median_time_warm_s": r1["median_time_warm"]},
            "seed_rel_l2_median": r["seed_rel_l2_median"],
            "heldout_mse_u_v_rho": [round(float(x), 5) for x in r["heldout_mse"]],
            "target_scales": [round(float(x), 5) for x in net.target_scales],
            "field_rel_l2_max": float(r["field_rel_l2"][both].max()) if both.any() else None,
            "loss_first_last": [hist[0], hist[-1]],
            "data_s": round(t_data, 1),
            "train_s": round(t_train, 1),
        },
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
