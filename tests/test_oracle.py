"""Pins the CPU oracle (oracle/lbm_oracle.c).

1. Against the reference program itself: fixtures in tests/golden/ were
   produced by the unmodified reference serial solver (oracle/_ref); the
   oracle must reproduce every FP64 value and flag bit for bit.
2. Where the reference has no behaviour to compare (switches for rest start,
   textbook MRT, TRT magic parameter, full-field / velocity residuals,
   parabolic-inlet extent, first-order outlet), against closed forms and
   invariants chosen so a dropped term, a sign or an index slip fails.
"""
import ctypes as C
import hashlib
import math
import os
import tempfile

import numpy as np
import pytest

import meshgen
import oracle as O
from golden_util import (SETUP_FIELDS, ini_dict, load, medium_cases, residual_lines, sha,
                         shipped_cases, small_cases)

FIELDS_EXACT = ["fluid", "corner", "stream", "id", "boundary", "boundary_id", "bc", "q", "x",
                "y", "uo", "vo", "u", "v", "rho", "f", "metaf", "drag", "lift"]


def oracle_for(g, **extra):
    p = O.params(**{**ini_dict(g["ini"]), **extra})
    return O.Oracle(g["nodes"], g["conn"], p)


@pytest.mark.parametrize("name", small_cases())
def test_oracle_matches_reference_bitwise(name):
    g = load(name)
    o = oracle_for(g)
    assert (o.n, o.m) == (int(g["n"]), int(g["m"]))
    lines = ["Iter L2_norm L2_norm_weighted Drag Lift"]
    iters = int(g["ini"][8])
    for _ in range(iters):
        it = o.iter
        if o.step(1) != 0:
            break
        r = o.residuals()
        lines.append("%d %5.4e %5.4e %f %f" % (it, r[0], r[1], r[2], r[3]))
    cells = o.cells()
    for f in FIELDS_EXACT:
        assert np.array_equal(cells[f], g[f]), (name, f)
    assert lines == residual_lines(g), name


@pytest.mark.parametrize("name", shipped_cases())
def test_oracle_matches_reference_on_coarse(name):
    g = load(name)
    mesh = meshgen.cavity(int(g["N"]))
    nodes, conn = mesh.arrays()
    o = O.Oracle(nodes, conn, O.params(**ini_dict(g["ini"])))
    o.step(int(g["ini"][8]))
    cells = o.cells()
    for f in ("u", "v", "rho"):
        assert np.array_equal(cells[f], g[f]), f
    assert hashlib.sha256(np.ascontiguousarray(cells["f"]).tobytes()).hexdigest() == str(g["f_sha256"])


@pytest.mark.parametrize("name,N", [("coarse", 125), ("medium", 253)])
def test_shipped_mesh_regenerated(name, N):
    """meshgen reproduces cases/coarse and cases/medium: same cell state as the
    shipped files (when present)."""
    path = f"/root/reference/cases/{name}/Mesh"
    if not os.path.isdir(path):
        pytest.skip("reference checkout not present")
    p = O.params()
    a = O.Oracle.from_files(path + "/D2node.dat", path + "/BCconnectors.dat", p).cells()
    nodes, conn = meshgen.cavity(N).arrays()
    b = O.Oracle(nodes, conn, p).cells()
    for f in a.dtype.names:
        assert np.array_equal(a[f], b[f]), f


@pytest.mark.parametrize("name", medium_cases())
def test_oracle_matches_reference_on_medium(name):
    """The shipped medium mesh, 1000 steps: the oracle on the regenerated mesh
    starts from the reference's CellIni state (checksums) and ends on its
    u, v, rho and F (checksums and three full rows)."""
    g = load(name)
    nodes, conn = meshgen.cavity(int(g["N"])).arrays()
    o = O.Oracle(nodes, conn, O.params(**ini_dict(g["ini"])))
    cells = o.cells()
    for f in SETUP_FIELDS:
        assert sha(cells[f]) == str(g[f"setup_{f}_sha256"]), f
    o.step(int(g["ini"][8]))
    cells = o.cells()
    n, m = int(g["n"]), int(g["m"])
    for f in ("u", "v", "rho"):
        assert np.array_equal(cells[f].reshape(m, n)[g["rows"]], g[f"{f}_rows"]), f
    for f in ("u", "v", "rho", "f"):
        assert sha(cells[f]) == str(g[f"{f}_sha256"]), f


def test_read_ini_shipped_layout(tmp_path):
    txt = "0.08\n0.0\n5\n0.01\n2\n1\n2\n2\n5000\n100\n10000\n1\n0\n-1.0\n-1.0\n0\n\n:01: comment\n"
    (tmp_path / "SetUpData.ini").write_text(txt)
    p = O.Params()
    assert O.lib().or_read_ini(str(tmp_path / "SetUpData.ini").encode(), C.byref(p)) == 0
    assert (p.collision, p.iterations, p.autosave_every, p.autosave_after) == (1, 5000, 100, 10000)
    assert abs(p.uavg - 0.08) < 1e-7 and p.rho_ini == 5.0 and p.crit_v == -1.0


# ---------------------------------------------------------------- closed forms
def test_d2q9_constants():
    w = np.zeros(9); cx = np.zeros(9, np.int32); cy = np.zeros(9, np.int32)
    opp = np.zeros(9, np.int32); c = np.zeros(9, np.int32)
    O.lib().or_d2q9(w.ctypes.data, cx.ctypes.data, cy.ctypes.data, opp.ctypes.data,
                    c.ctypes.data, 10)
    assert abs(w.sum() - 1) < 1e-15
    assert np.allclose((w * cx * cx).sum(), 1 / 3) and np.allclose((w * cx * cy).sum(), 0)
    assert np.all(cx[opp] == -cx) and np.all(cy[opp] == -cy)
    assert np.all(c == -cx - 10 * cy)


def test_mrt_matrix_inverse():
    tm = np.zeros((9, 9)); st = np.zeros((9, 9))
    O.lib().or_mrt_init(tm.ctypes.data, st.ctypes.data, np.ones(9).ctypes.data)
    assert np.allclose(st @ tm, np.eye(9), atol=1e-14)
    s = np.arange(1, 10) / 7.0
    O.lib().or_mrt_init(tm.ctypes.data, st.ctypes.data, s.ctypes.data)
    assert np.allclose(st, np.linalg.inv(tm) * s[None, :], atol=1e-15)


def _box(N):
    """Closed box: all four sides plain walls (no lid)."""
    mesh = meshgen.cavity(N, lid_type=1)
    return mesh.arrays()


def test_mass_conserved_in_closed_box():
    """Interpolated bounce-back at q = 1/2 (Delta = 1/16 makes every q exactly
    0.5) is the mass-conserving half-way rule; every collision conserves rho.
    The reference's straight-wall rule copies the post-streaming population
    (serial/BoundaryConditions.c:113) and does not conserve mass."""
    nodes, conn = _box(16)
    for coll in (1, 2, 3):
        o = O.Oracle(nodes, conn, O.params(collision=coll, rho_ini=1.0, uavg=0.05,
                                           viscosity=0.05, curved=1, trt_fp64=1))
        assert np.all(o.cells()["q"][o.cells()["bc"] != 0] == 0.5)
        o.step(1)
        m0 = o.cells()["f"].sum()
        o.step(200)
        m1 = o.cells()["f"].sum()
        assert abs(m1 - m0) < 1e-12 * m0, coll
    o = O.Oracle(nodes, conn, O.params(collision=1, rho_ini=1.0, uavg=0.05, viscosity=0.05))
    o.step(1)
    m0 = o.cells()["f"].sum()
    o.step(200)
    assert abs(o.cells()["f"].sum() - m0) > 1e-6 * m0


def test_equilibrium_and_omega_one():
    nodes, conn = meshgen.cavity(6).arrays()
    o = O.Oracle(nodes, conn, O.params(viscosity=0.1, rho_ini=2.0, uavg=0.0))
    om = o.omegas()[0]
    o.step(1)
    c = o.cells()
    fl = c["fluid"] == 1
    w = np.array([4 / 9] + [1 / 9] * 4 + [1 / 36] * 4)
    # the first collision runs from F = 0 with U = Uo = 0, Rho = rho_ini:
    # Feq(rho, 0) = rho w exactly and METAF = omega Feq + (1 - omega) 0
    assert np.array_equal(c["feq"][fl], np.broadcast_to(2.0 * w, c["feq"][fl].shape))
    assert np.array_equal(c["metaf"][fl], om * c["feq"][fl])


def _state_after(p_kw, steps, N=10, mesh=None):
    nodes, conn = (mesh or meshgen.cavity(N)).arrays()
    o = O.Oracle(nodes, conn, O.params(**p_kw))
    o.step(steps)
    return o.cells()


def test_mrt_reduces_to_bgk():
    """Textbook m_eq with every rate = omega is BGK (up to rounding)."""
    base = dict(uavg=0.1, rho_ini=1.0, viscosity=0.07, init_rest=1)
    nodes, conn = meshgen.cavity(4).arrays()
    om = O.Oracle(nodes, conn, O.params(**base)).omegas()[0]
    a = _state_after({**base, "collision": 1}, 30)
    b = _state_after({**base, "collision": 3, "mrt_textbook": 1, "mrt_s": [om] * 9}, 30)
    assert np.allclose(a["f"], b["f"], rtol=0, atol=1e-13)
    # ... and the reference's rho^2 equilibria are NOT the textbook ones when rho != 1
    c = _state_after({**base, "rho_ini": 3.0, "collision": 3, "mrt_s": [om] * 9}, 5)
    d = _state_after({**base, "rho_ini": 3.0, "collision": 3, "mrt_textbook": 1,
                      "mrt_s": [om] * 9}, 5)
    assert np.abs(c["f"] - d["f"]).max() > 1e-6


def test_trt_reduces_to_bgk():
    """omega_minus == omega (Lambda = (1/omega - 1/2)^2) with FP64 parts is BGK."""
    visc = np.float32(0.07)
    om = 1.0 / (3.0 * np.float64(visc) + 0.5)
    lam = (1 / om - 0.5) ** 2
    base = dict(uavg=0.1, rho_ini=1.0, viscosity=float(visc), init_rest=1)
    a = _state_after({**base, "collision": 1}, 30)
    b = _state_after({**base, "collision": 2, "trt_fp64": 1, "trt_lambda": lam}, 30)
    assert np.allclose(a["f"], b["f"], rtol=0, atol=1e-13)
    # Lambda = 3/16 reproduces the reference's 8(2-w)/(8-w) to rounding
    nodes, conn = meshgen.cavity(4).arrays()
    o1 = O.Oracle(nodes, conn, O.params(viscosity=0.03))
    o2 = O.Oracle(nodes, conn, O.params(viscosity=0.03, trt_lambda=3 / 16))
    assert abs(o1.omegas()[1] - o2.omegas()[1]) < 1e-14
    # the binary32 parts of the reference TRT change the result
    c = _state_after({**base, "collision": 2}, 5)
    d = _state_after({**base, "collision": 2, "trt_fp64": 1}, 5)
    assert 0 < np.abs(c["f"] - d["f"]).max() < 1e-6


def test_rest_start_is_fixed_point_without_lid():
    nodes, conn = _box(8)
    o = O.Oracle(nodes, conn, O.params(collision=3, rho_ini=1.0, init_rest=1, uavg=0.1,
                                       viscosity=0.1, mrt_textbook=1))
    o.step(50)
    c = o.cells()
    fl = c["fluid"] == 1
    assert np.abs(c["u"][fl]).max() < 1e-15 and np.abs(c["rho"][fl] - 1).max() < 1e-14


def test_cavity_mirror_symmetry_first_step():
    """A lid moving +x equals the mirror image of a lid moving -x."""
    N = 9
    a = _state_after(dict(uavg=0.1, rho_ini=1.0, viscosity=0.1), 1, N=N)
    b = _state_after(dict(uavg=-0.1, rho_ini=1.0, viscosity=0.1), 1, N=N)
    n, m = N + 2, N + 3
    ua = a["u"].reshape(m, n)[1:N + 1, 1:N + 1]
    ub = b["u"].reshape(m, n)[1:N + 1, 1:N + 1]
    ra = a["rho"].reshape(m, n)[1:N + 1, 1:N + 1]
    rb = b["rho"].reshape(m, n)[1:N + 1, 1:N + 1]
    assert np.allclose(ua, -ub[:, ::-1], atol=1e-15)
    assert np.allclose(ra, rb[:, ::-1], atol=1e-15)


def test_parabolic_inlet_profile():
    """InletProfile 1: Uo = 6 U (y-lo)(hi-y)/(hi-lo)^2 over the inlet extent
    (min/max inlet-link y -+ Delta/2), peak 1.5 U at mid-height."""
    NX, NY = 6, 10
    nodes, conn = meshgen.channel(NX, NY).arrays()
    o = O.Oracle(nodes, conn, O.params(inlet_profile=1, uavg=0.04, rho_ini=1.0))
    d = 1.0 / NY
    lo, hi = o.min_inlet_y, o.max_inlet_y
    ys = np.float32(np.array([(j - 0.5) * d for j in range(1, NY + 1)])).astype(np.float64)
    assert abs(lo - (min(ys) - d / 2)) < 1e-6 and abs(hi - (max(ys) + d / 2)) < 1e-6
    c = o.cells()
    uo = c["uo"].reshape(NY + 2, NX + 2)[1:NY + 1, 1]
    expect = 6 * np.float32(0.04) * (ys - lo) * (hi - ys) / (hi - lo) ** 2
    assert np.allclose(uo, expect, rtol=1e-12)
    # the inlet spans y in [0, 1]: profile 6U y(1-y), zero at the walls
    assert np.allclose(expect, 6 * 0.04 * ys * (1 - ys), rtol=1e-5)


def test_first_order_outlet_copies_upstream():
    nodes, conn = meshgen.channel(12, 6).arrays()
    o = O.Oracle(nodes, conn, O.params(outlet_profile=3, uavg=0.05, rho_ini=1.0,
                                       viscosity=0.05))
    o.step(20)
    c = o.cells()
    n = 14
    for j in range(1, 7):
        a = c["f"][j * n + 12]
        b = c["f"][j * n + 11]
        assert np.array_equal(a[[1, 5, 8]], b[[1, 5, 8]])


def test_full_field_residual_closed_form():
    """resid_kind 1 sums (F-METAF)^2 over all cells (cuda_step3 intent)."""
    nodes, conn = meshgen.cavity(7).arrays()
    o = O.Oracle(nodes, conn, O.params(resid_kind=1, resid_every=1, uavg=0.1, rho_ini=1.0,
                                       viscosity=0.1))
    o.step(3)
    c = o.cells()
    s = ((c["f"] - c["metaf"]) ** 2).sum()
    r = o.residuals()
    assert abs(r[0] - math.sqrt(s)) < 1e-13 * max(1, r[0])
    assert abs(r[1] - math.sqrt(s / (o.n * o.m))) < 1e-13
    assert r[0] > 0


def test_velocity_residual_definition():
    nodes, conn = meshgen.cavity(7).arrays()
    K = 5
    p = O.params(resid_kind=2, resid_every=K, uavg=0.1, rho_ini=1.0, viscosity=0.1)
    o = O.Oracle(nodes, conn, p)
    o.step(K)
    _, u0, v0 = o.macro()
    o.step(K)
    rho, u1, v1 = o.macro()
    fl = o.cells()["fluid"] == 1
    du = np.sqrt(((u1 - u0)[fl] ** 2 + (v1 - v0)[fl] ** 2).sum())
    uu = np.sqrt((u1[fl] ** 2 + v1[fl] ** 2).sum())
    assert abs(o.residuals()[0] - du / (K * uu)) < 1e-14
    assert o.iter == 2 * K


def test_writers_match_reference_text():
    """The oracle's CSV writer reproduces the reference's Final/Initial layout."""
    import golden_util
    d = os.path.join(golden_util.GOLDEN, "text", "cav8_text_csv")
    values = [float(v) for v in open(os.path.join(d, "ini.txt")).read().split()]
    nodes, conn = meshgen.cavity(8).arrays()
    o = O.Oracle(nodes, conn, O.params(**ini_dict(values)))
    with tempfile.TemporaryDirectory() as t:
        o.write_results(os.path.join(t, "i.csv"), final=False)
        assert open(os.path.join(t, "i.csv")).read() == open(os.path.join(d, "InitialData.csv")).read()
        o.step(10)
        o.write_results(os.path.join(t, "f.csv"), final=True)
        assert open(os.path.join(t, "f.csv")).read() == open(os.path.join(d, "FinalData.csv")).read()


# ------------------------------------------------ channel physics (not in the reference)
def test_pressure_outlet_closed_forms():
    """Zou-He pressure outlet: at every outlet node (BC_ID[1] = 3, not a
    corner) the post-boundary density equals rho_out and j_y = 0."""
    NX, NY = 30, 12
    nodes, conn = meshgen.channel(NX, NY).arrays()
    o = O.Oracle(nodes, conn, O.params(outlet_profile=4, outlet_rho=1.02, uavg=0.05, rho_ini=1.0,
                                       viscosity=0.05, collision=3))
    o.step(40)
    c = o.cells()
    n = NX + 2
    for j in range(2, NY):  # skip corner rows
        cell = c[j * n + NX]
        assert cell["bc"][1] == 3 and cell["corner"] == 0
        F = cell["f"]
        assert abs(F.sum() - 1.02) < 1e-13
        jy = F[2] - F[4] + F[5] + F[6] - F[7] - F[8]
        assert abs(jy) < 1e-15
        assert abs(cell["rho"] - 1.02) < 1e-13


def test_force_closed_forms():
    """Momentum exchange in a closed box at rest: the bottom wall feels the
    lattice pressure rho/3 per unit length downward (N nodes plus the corner
    diagonals), the top wall the opposite, and all walls together nothing."""
    N = 16
    nodes, conn = meshgen.cavity(N, lid_type=1).arrays()
    o = O.Oracle(nodes, conn, O.params(collision=1, rho_ini=1.0, uavg=0.0, init_rest=1,
                                       viscosity=0.1, curved=1))
    o.step(3)
    f = {b: o.force(b) for b in (1, 2, 3, 4)}
    total = sum(f.values())
    assert np.abs(total).max() < 1e-12
    # west / east walls are mirror images (their corners carry the same ids)
    assert abs(f[1][0] + f[3][0]) < 1e-12 and abs(f[1][1] - f[3][1]) < 1e-12
    # bottom pushed down, top pushed up, by rho/3 per node plus corner links
    assert -(N + 3) / 3 < f[4][1] < -(N - 3) / 3
    assert (N - 3) / 3 < f[2][1] < (N + 3) / 3
    assert abs(f[4][0]) < 1e-12 and abs(f[2][0]) < 1e-12


def test_strouhal_of_a_sinusoid():
    import parallellbmcranfield_amd as lbm
    steps, every = 100000, 10
    t = np.arange(0, steps, every)
    period = 1234.5
    lift = 0.3 * np.sin(2 * np.pi * t / period) + 0.01 * np.cos(2 * np.pi * t / 97.0) + 5.0
    st = lbm.strouhal(lift, every, diameter=64.0, velocity=0.04)
    assert abs(st - 64.0 / (period * 0.04)) < 2e-3 * st
