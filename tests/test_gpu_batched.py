"""Batched multi-instance solver on the GPU, through the C ABI.

Bar: an instance of a packed lattice advances exactly as the same cavity run
alone (FP32 / FP16 storage, MRT and BGK: bit-identical fields), converged
instances stop at the iteration the single-lattice solver stops at, and the
single-lattice cavity_ex path is itself bit-exact against the FP64 oracle.
"""
import numpy as np
import pytest

import oracle as O
import parallellbmcranfield_amd as lbm

pytestmark = pytest.mark.gpu

# (NX, NY, rectangles, viscosity, lid speed)
SET = [
    (40, 28, [(5, 5, 10, 9)], 0.02, 0.1),
    (23, 47, [], 0.05, 0.08),
    (64, 33, [(30, 4, 34, 20), (50, 25, 55, 30)], 0.01, 0.1),
    (17, 17, [], 0.1, 0.05),
]


def batch_opts(prec, coll=3, **kw):
    return lbm.make_options(collision=coll, uavg=0.1, rho_ini=1.0, viscosity=0.02,
                            precision=prec, init=lbm.INIT_REST,
                            resid_kind=lbm.RESID_VELOCITY, **kw)


def single(NX, NY, rects, nu, u, prec, coll, steps, **kw):
    case = lbm.Case.cavity_ex(NX, NY, rects)
    o = lbm.make_options(collision=coll, uavg=u, rho_ini=1.0, viscosity=nu, precision=prec,
                         init=lbm.INIT_REST, resid_kind=lbm.RESID_VELOCITY, **kw)
    s = lbm.Solver(case, o)
    s.step(steps)
    return case, s


def crop(a, packed, k):
    x0, y0, w, h = packed.rects[k]
    return a.reshape(packed.m, packed.n)[y0:y0 + h, x0:x0 + w].ravel()


@pytest.mark.parametrize("prec,coll", [(lbm.PREC_F32, 3), (lbm.PREC_F16, 3), (lbm.PREC_F32, 1)])
def test_batched_instances_equal_single_runs(prec, coll):
    print("batched vs single", prec, coll, flush=True)
    cases = [lbm.Case.cavity_ex(nx, ny, r) for nx, ny, r, _, _ in SET]
    packed = lbm.Case.pack(cases)
    s = lbm.Solver(packed, batch_opts(prec, coll, resid_every=50), instances=[(nu, u) for *_, nu, u in SET])
    steps = 301
    assert s.step(steps) == steps
    rho, u, v = s.macro()
    it, _, res = s.batch_status()
    assert (it == -1).all() and np.all(res > 0)
    for k, (nx, ny, r, nu, ul) in enumerate(SET):
        _, s1 = single(nx, ny, r, nu, ul, prec, coll, steps, resid_every=50)
        r1, u1, v1 = s1.macro()
        assert np.array_equal(crop(rho, packed, k), r1), k
        assert np.array_equal(crop(u, packed, k), u1), k
        assert np.array_equal(crop(v, packed, k), v1), k
        # the single run's velocity residual is the instance's residual
        assert abs(s1.residuals()[0] - res[k]) <= 1e-9 * res[k]


def test_batched_convergence_matches_single_runs():
    print("batched convergence", flush=True)
    crit, K = 2e-4, 20
    cases = [lbm.Case.cavity_ex(nx, ny, r) for nx, ny, r, _, _ in SET]
    packed = lbm.Case.pack(cases)
    o = batch_opts(lbm.PREC_F32, resid_every=K, crit_v=crit, use_stop_rule=1,
                   max_iterations=20000)
    s = lbm.Solver(packed, o, instances=[(nu, u) for *_, nu, u in SET])
    s.step(20000)
    it, sec, res = s.batch_status()
    assert (it > 0).all(), it
    assert s.iter == it.max()
    assert np.all(np.diff(sec[np.argsort(it)]) >= 0)
    rho, u, v = s.macro()
    for k, (nx, ny, r, nu, ul) in enumerate(SET):
        case = lbm.Case.cavity_ex(nx, ny, r)
        o1 = lbm.make_options(collision=3, uavg=ul, rho_ini=1.0, viscosity=nu,
                              precision=lbm.PREC_F32, init=lbm.INIT_REST,
                              resid_kind=lbm.RESID_VELOCITY, resid_every=K, crit_v=crit,
                              use_stop_rule=1, max_iterations=20000)
        s1 = lbm.Solver(case, o1)
        s1.step(20000)
        assert s1.iter == it[k], (k, s1.iter, it[k])
        # a stopped instance holds the state it converged at; observed one
        # step on, as the single lattice advanced once more
        o1.use_stop_rule = 0
        s2 = lbm.Solver(case, o1)
        s2.step(int(it[k]) + 1)
        r2, u2, v2 = s2.macro()
        assert np.array_equal(crop(u, packed, k), u2), k
        assert np.array_equal(crop(v, packed, k), v2), k
        assert np.array_equal(crop(rho, packed, k), r2), k


def test_converged_instance_stays_fixed():
    print("frozen instance", flush=True)
    SET2 = [(16, 16, [], 0.1, 0.05), (72, 72, [], 0.004, 0.1), (30, 20, [(4, 4, 8, 8)], 0.02, 0.1)]
    packed = lbm.Case.pack([lbm.Case.cavity_ex(nx, ny, r) for nx, ny, r, _, _ in SET2])
    o = batch_opts(lbm.PREC_F32, resid_every=10, crit_v=1e-3)
    s = lbm.Solver(packed, o, instances=[(nu, u) for *_, nu, u in SET2])
    while not (s.batch_status()[0] >= 0).any():
        s.step(10)
        assert s.iter < 50000
    it0 = s.batch_status()[0].copy()
    frozen = it0 >= 0
    a = [[crop(x, packed, k).copy() for x in s.macro()] for k in range(len(SET2))]
    s.step(7)  # odd count: both buffers of a stopped instance must hold its state
    b = [[crop(x, packed, k) for x in s.macro()] for k in range(len(SET2))]
    it1 = s.batch_status()[0]
    for k in range(len(SET2)):
        same = all(np.array_equal(x, y) for x, y in zip(a[k], b[k]))
        if frozen[k]:
            assert same and it1[k] == it0[k], k
        else:
            assert not same, k  # running instances keep advancing


def test_cavity_ex_fp64_bit_exact_vs_oracle(tmp_path):
    print("cavity_ex vs oracle", flush=True)
    case = lbm.Case.cavity_ex(26, 15, [(6, 4, 9, 8)])
    nodes, bc = str(tmp_path / "n.dat"), str(tmp_path / "c.dat")
    case.write(nodes, bc)
    p = O.params(collision=3, uavg=0.1, rho_ini=1.0, viscosity=0.03, init_rest=1,
                 resid_kind=3)
    ref = O.Oracle.from_files(nodes, bc, p)
    ref.step(150)
    c2 = lbm.Case.load(nodes, bc)
    o = lbm.make_options(collision=3, uavg=0.1, rho_ini=1.0, viscosity=0.03,
                         init=lbm.INIT_REST, resid_kind=lbm.RESID_NONE)
    s = lbm.Solver(c2, o)
    s.step(150)
    for a, b in zip(s.macro(), ref.macro()):
        assert np.array_equal(a, b)


def test_batched_rejects_bad_combinations():
    packed = lbm.Case.pack([lbm.Case.cavity_ex(8, 8), lbm.Case.cavity_ex(9, 9)])
    with pytest.raises(lbm.LbmError):  # one parameter set per instance
        lbm.Solver(packed, batch_opts(lbm.PREC_F32))
    with pytest.raises(lbm.LbmError):  # FP64 batches are not offered
        lbm.Solver(packed, batch_opts(lbm.PREC_F64), instances=[(0.1, 0.1)] * 2)
    s = lbm.Solver(packed, batch_opts(lbm.PREC_F32), instances=[(0.1, 0.1)] * 2)
    with pytest.raises(lbm.LbmError):
        s.save("/tmp/never.ckpt")


def test_batched_field_start_equals_single_field_starts():
    print("batched field start", flush=True)
    cases = [lbm.Case.cavity_ex(nx, ny, r) for nx, ny, r, _, _ in SET]
    packed = lbm.Case.pack(cases)
    rng = np.random.default_rng(7)
    fields = []
    full = [np.ones(packed.n * packed.m), np.zeros(packed.n * packed.m),
            np.zeros(packed.n * packed.m)]
    for k, c in enumerate(cases):
        r = 1.0 + 0.01 * rng.standard_normal(c.n * c.m)
        u = 0.02 * rng.standard_normal(c.n * c.m)
        v = 0.02 * rng.standard_normal(c.n * c.m)
        fields.append((r, u, v))
        x0, y0, w, h = packed.rects[k]
        for F, f in zip(full, (r, u, v)):
            F.reshape(packed.m, packed.n)[y0:y0 + h, x0:x0 + w] = f.reshape(h, w)
    s = lbm.Solver(packed, batch_opts(lbm.PREC_F32, resid_every=25), field=full,
                   instances=[(nu, u) for *_, nu, u in SET])
    s.step(120)
    rho, u, v = s.macro()
    for k, (nx, ny, rr, nu, ul) in enumerate(SET):
        o = lbm.make_options(collision=3, uavg=ul, rho_ini=1.0, viscosity=nu,
                             precision=lbm.PREC_F32, resid_kind=lbm.RESID_VELOCITY,
                             resid_every=25)
        s1 = lbm.Solver(cases[k], o, field=fields[k])
        s1.step(120)
        r1, u1, v1 = s1.macro()
        assert np.array_equal(crop(rho, packed, k), r1), k
        assert np.array_equal(crop(u, packed, k), u1), k
        assert np.array_equal(crop(v, packed, k), v1), k


def test_neq_start_has_the_chapman_enskog_stress_and_converges_sooner():
    """LBM_INIT_FIELD_NEQ: the seeded non-equilibrium part carries no mass or
    momentum and the BGK Chapman-Enskog stress Pi = -rho tau / 3 (grad u +
    grad u^T) of the field (interior nodes, central differences of a
    quadratic field are exact); started from a converged field it re-converges
    in fewer steps than the plain equilibrium seed."""
    print("neq start", flush=True)
    NX, NY, nu = 30, 20, 0.05
    case = lbm.Case.cavity_ex(NX, NY)
    n, m = case.n, case.m
    yy, xx = np.meshgrid(np.arange(m, dtype=float), np.arange(n, dtype=float), indexing="ij")
    u = (1e-3 * (xx * yy + 0.5 * yy ** 2)).ravel()
    v = (1e-3 * (xx ** 2 - 2 * xx * yy)).ravel()
    rho = (1 + 1e-4 * xx).ravel()
    o = lbm.make_options(collision=1, uavg=0.1, rho_ini=1.0, viscosity=nu,
                         precision=lbm.PREC_F64, init=lbm.INIT_FIELD_NEQ,
                         resid_kind=lbm.RESID_NONE)
    s = lbm.Solver(case, o, field=(rho, u, v))
    f_host, _ = s.distributions()  # the seed as the host states it (iteration 0)
    s.step(1)
    _, metaf0 = s.distributions()  # own METAF of the previous buffer: the device seed, collided
    W = np.array([4 / 9] + [1 / 9] * 4 + [1 / 36] * 4)
    CX = np.array([0, 1, 0, -1, 0, 1, -1, -1, 1])
    CY = np.array([0, 0, 1, 0, -1, 1, 1, -1, -1])
    T2 = u[:, None] * CX + v[:, None] * CY
    feq = rho[:, None] * W * (1 + 3 * T2 + 4.5 * T2 ** 2 - 1.5 * (u ** 2 + v ** 2)[:, None])
    tau = 3 * float(np.float32(nu)) + 0.5
    om = 1 / tau
    f = (metaf0 - om * feq) / (1 - om)  # BGK: METAF = om feq + (1 - om) F
    fl = case.cells()["fluid"] == 1
    assert np.allclose(f[fl], f_host[fl], rtol=0, atol=1e-13)
    fn = f - feq
    X, Y = xx.ravel(), yy.ravel()
    dudx, dudy = 1e-3 * Y, 1e-3 * (X + Y)
    dvdx, dvdy = 1e-3 * (2 * X - 2 * Y), -2e-3 * X
    inner = ((X >= 2) & (X <= NX - 1) & (Y >= 2) & (Y <= NY - 1))
    assert np.abs(fn[inner].sum(1)).max() < 1e-13
    assert np.abs((fn * CX)[inner].sum(1)).max() < 1e-13
    pxy = (fn * CX * CY).sum(1)
    assert np.allclose(pxy[inner], (-rho * tau / 3 * (dudy + dvdx))[inner], rtol=1e-7, atol=1e-13)
    pd = (fn * (CX ** 2 - CY ** 2)).sum(1)
    assert np.allclose(pd[inner], (-2 * rho * tau / 3 * (dudx - dvdy))[inner], rtol=1e-7,
                       atol=1e-13)
    # re-convergence from a converged field
    specs = [(40, 28, [(5, 5, 10, 9)], 0.02, 0.1), (33, 33, [], 0.01, 0.1)]
    packed = lbm.Case.pack([lbm.Case.cavity_ex(a, b, r) for a, b, r, _, _ in specs])
    inst = [(nu_, u_) for *_, nu_, u_ in specs]
    ob = batch_opts(lbm.PREC_F32, resid_every=50, crit_v=1e-6, use_stop_rule=1,
                    max_iterations=200000)
    s0 = lbm.Solver(packed, ob, instances=inst)
    s0.step(200000)
    it0 = s0.batch_status()[0]
    assert (it0 > 0).all()
    fld = s0.macro()
    steps = {}
    for kind in (lbm.INIT_FIELD, lbm.INIT_FIELD_NEQ):
        ob.init = kind
        s1 = lbm.Solver(packed, ob, field=fld, instances=inst)
        s1.step(200000)
        steps[kind] = s1.batch_status()[0]
    print("rest", it0, "feq", steps[lbm.INIT_FIELD], "neq", steps[lbm.INIT_FIELD_NEQ], flush=True)
    assert (steps[lbm.INIT_FIELD_NEQ] > 0).all()
    assert (steps[lbm.INIT_FIELD_NEQ] <= steps[lbm.INIT_FIELD]).all()
    assert (steps[lbm.INIT_FIELD_NEQ] < it0 / 2).all()


# Uneven sizes, aspect ratios, Re and lid speeds; rectangles in the bulk, near
# walls and near the lid corners.  (NX, NY, rectangles, viscosity, lid speed)
ORACLE_SET = [
    (48, 30, [(5, 5, 12, 9), (30, 18, 36, 25)], 0.02, 0.1),
    (25, 53, [(3, 40, 20, 45)], 0.05, 0.07),
    (70, 36, [], 0.008, 0.1),
    (19, 19, [(8, 3, 11, 14)], 0.1, 0.05),
    (33, 41, [(3, 3, 6, 6), (25, 30, 30, 36), (10, 20, 18, 24)], 0.015, 0.12),
]


def _oracle_params(kind, nu, lid):
    from parallellbmcranfield_amd import sweep as SW
    if kind == "mrt_sweep":  # the batched sweep's collision: MRT, Lallemand-Luo rates, 1/tau of the instance
        return O.params(collision=3, uavg=lid, rho_ini=1.0, viscosity=nu, init_rest=1,
                        resid_kind=3, mrt_textbook=1, mrt_s=SW.mrt_rates(nu))
    if kind == "mrt_reference":  # MRTInitializer's rates and the reference's rho^2 equilibria
        return O.params(collision=3, uavg=lid, rho_ini=1.0, viscosity=nu, init_rest=1,
                        resid_kind=3)
    return O.params(collision=1, uavg=lid, rho_ini=1.0, viscosity=nu, init_rest=1, resid_kind=3)


@pytest.mark.parametrize("kind", ["mrt_sweep", "mrt_reference", "bgk"])
def test_batched_instances_match_the_fp64_oracle(kind, tmp_path):
    """Each instance of a packed lattice (FP32 storage, its own omega and lid
    speed) against the FP64 oracle run alone on that instance's mesh, written
    in the reference format by lbm_case_write and read back by the oracle's
    own reader: velocity relative L2 < 2e-5 and density within 1e-5 after 400
    steps, per instance.  Anchors: the loop body serial/Iterate.c:263-332,
    BGKW / MRT serial/CellFunctions.c:352-365, 401-438."""
    from parallellbmcranfield_amd import sweep as SW
    print("batched vs oracle", kind, flush=True)
    steps = 400
    cases = [lbm.Case.cavity_ex(nx, ny, r) for nx, ny, r, _, _ in ORACLE_SET]
    packed = lbm.Case.pack(cases)
    if kind == "mrt_sweep":
        o = SW.sweep_options(crit=-1.0, resid_every=50)
        o.use_stop_rule = 0
    else:
        o = batch_opts(lbm.PREC_F32, 3 if kind == "mrt_reference" else 1, resid_every=50)
    s = lbm.Solver(packed, o, instances=[(nu, u) for *_, nu, u in ORACLE_SET])
    assert s.step(steps) == steps
    rho, u, v = s.macro()
    for k, (nx, ny, r, nu, lid) in enumerate(ORACLE_SET):
        nf, cf = str(tmp_path / f"n{k}.dat"), str(tmp_path / f"c{k}.dat")
        cases[k].write(nf, cf)
        ref = O.Oracle.from_files(nf, cf, _oracle_params(kind, nu, lid))
        assert (ref.n, ref.m) == (cases[k].n, cases[k].m)
        ref.step(steps)
        ro, uo, vo = ref.macro()
        fl = cases[k].cells()["fluid"] == 1
        ub, vb, rb = crop(u, packed, k)[fl], crop(v, packed, k)[fl], crop(rho, packed, k)[fl]
        err = np.sqrt(((ub - uo[fl]) ** 2 + (vb - vo[fl]) ** 2).sum() /
                      (uo[fl] ** 2 + vo[fl] ** 2).sum())
        drho = np.abs(rb - ro[fl]).max()
        print(f"instance {k} ({nx}x{ny}, nu {nu}, lid {lid}): velocity rel L2 {err:.2e}, "
              f"max |drho| {drho:.2e}", flush=True)
        assert err < 2e-5, k
        assert drho < 1e-5, k


@pytest.mark.parametrize("prec", [lbm.PREC_F32, lbm.PREC_F16])
def test_batched_launch_shapes_are_bit_identical(prec):
    """The sweep's tuner times tile lengths 256 / 512 / 1024 and XCD runs of
    1..16 tiles on the packed lattice: every shape gives the same fields bit
    for bit, and the same per-instance stop iterations."""
    print("batched shapes", prec, flush=True)
    cases = [lbm.Case.cavity_ex(nx, ny, r) for nx, ny, r, _, _ in SET]
    packed = lbm.Case.pack(cases)
    inst = [(nu, u) for *_, nu, u in SET]
    out = []
    for bs, g in ((256, 4), (256, 1), (512, 2), (1024, 16)):
        s = lbm.Solver(packed, batch_opts(prec, resid_every=10, crit_v=1e-3, use_stop_rule=1,
                                          block_size=bs, xcd_group=g), instances=inst)
        s.step(600)
        it, _, _ = s.batch_status()
        out.append((s.macro(), it, (bs, g)))
        s.close()
    assert (out[0][1] > 0).any(), "no instance stopped: the criterion is not exercised"
    for fields, it, shape in out[1:]:
        assert np.array_equal(it, out[0][1]), shape
        for a, b in zip(out[0][0], fields):
            assert np.array_equal(a, b), shape


@pytest.mark.parametrize("prec,coll", [(lbm.PREC_F32, 3), (lbm.PREC_F32, 1), (lbm.PREC_F16, 3)])
def test_batched_two_step_equals_single_steps(prec, coll):
    """Pairs of plain steps of a packed lattice as two-step launches (each
    node with its instance's rates; nodes of stopped instances not stored;
    the tile list shrinking as instances stop) against every step its own
    launch: same stop iterations, same fields bit for bit."""
    print("batched two-step", prec, coll, flush=True)
    cases = [lbm.Case.cavity_ex(nx, ny, r) for nx, ny, r, _, _ in SET]
    packed = lbm.Case.pack(cases)
    inst = [(nu, u) for *_, nu, u in SET]
    out = []
    for rows, flags in ((0, lbm.TUNE_NO_TWO_STEP), (8, 0), (3, lbm.TUNE_GRAPH)):
        s = lbm.Solver(packed, batch_opts(prec, coll, resid_every=10, crit_v=1e-3,
                                          use_stop_rule=1, two_step_rows=rows,
                                          tune_flags=flags), instances=inst)
        assert (s.two_step_info()["cells"] > 0) == (rows > 0)
        for chunk in (7, 200, 393):
            s.step(chunk)
        it, _, _ = s.batch_status()
        out.append((s.macro(), it, (rows, flags)))
        s.close()
    # instances stop at different iterations, so stopped ones sit beside
    # running ones for part of the run
    assert len(set(out[0][1][out[0][1] > 0])) > 1, out[0][1]
    for fields, it, shape in out[1:]:
        assert np.array_equal(it, out[0][1]), shape
        for a, b in zip(out[0][0], fields):
            assert np.array_equal(a, b), shape
