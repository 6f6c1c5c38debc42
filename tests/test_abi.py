"""CPU-side checks of the native library (no kernel launches).

- the C ABI library loads and exports every function include/lbm.h declares;
- the host set-up (ReadNodes/ReadBCconn/CompData*/CellIni restated in O(N))
  is bit-exact against the oracle on every golden mesh and the shipped-case
  equivalents;
- the mesh generators reproduce the reference-format meshes.
"""
import ctypes as C
import os
import re

import numpy as np
import pytest

import meshgen
import oracle as O
import parallellbmcranfield_amd as lbm
from golden_util import ini_dict, load, small_cases

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    txt = open(os.path.join(ROOT, "include", "lbm.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = re.findall(r"\b(lbm_[a-z_0-9]+)\s*\(", txt)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    L = lbm.lib()
    declared = header_functions()
    assert len(declared) >= 25
    for name in declared:
        assert hasattr(L, name), name
    assert sorted(lbm.EXPORTS) == declared
    assert L.lbm_abi_version() == 1


def test_struct_layouts_match_header():
    # sizes the C side sees (via offsets of the last field)
    assert C.sizeof(lbm.Ini) == 64
    assert lbm.Options.block_size.offset + 4 == C.sizeof(lbm.Options) or \
        C.sizeof(lbm.Options) % 8 == 0


def _cells_match(case: lbm.Case, o: O.Oracle):
    a = case.cells()
    b = o.cells()
    for f in ("fluid", "corner", "stream", "id", "boundary", "boundary_id", "bc", "q", "x", "y"):
        assert np.array_equal(a[f], b[f]), f


@pytest.mark.parametrize("name", small_cases())
def test_setup_matches_oracle(name):
    g = load(name)
    case = lbm.Case.from_arrays(g["nodes"], g["conn"])
    o = O.Oracle(g["nodes"], g["conn"], O.params(**ini_dict(g["ini"])))
    assert (case.n, case.m) == (o.n, o.m)
    assert case.info.delta == o.delta and case.info.num_inlet == o.num_inlet
    assert case.info.min_inlet_y == o.min_inlet_y and case.info.max_inlet_y == o.max_inlet_y
    _cells_match(case, o)
    # and against the reference program's own CellIni output (golden)
    a = case.cells()
    for f in ("fluid", "corner", "stream", "id", "boundary", "boundary_id", "bc", "q", "x", "y"):
        assert np.array_equal(a[f], g[f]), f


def test_file_parser_matches_fscanf(tmp_path):
    mesh = meshgen.cavity(13)
    mesh.write(str(tmp_path))
    nf = str(tmp_path / "Mesh" / "D2node.dat")
    bf = str(tmp_path / "Mesh" / "BCconnectors.dat")
    case = lbm.Case.load(nf, bf)
    o = O.Oracle.from_files(nf, bf, O.params())
    assert (case.n, case.m, case.info.num_nodes, case.info.num_conn) == (o.n, o.m, 15 * 16, len(mesh.links))
    _cells_match(case, o)


@pytest.mark.parametrize("variant", ["clean", "joined_lines", "stop_token"])
def test_large_mesh_parser_matches_fscanf(tmp_path, variant):
    # files over 1 MB are parsed on several threads when every line holds one
    # clean record; otherwise (two records on a line) the sequential reader
    # decides.  A token that fails to parse ends the table in both.
    mesh = meshgen.cavity(257)
    mesh.write(str(tmp_path))
    nf = str(tmp_path / "Mesh" / "D2node.dat")
    bf = str(tmp_path / "Mesh" / "BCconnectors.dat")
    assert os.path.getsize(nf) > (1 << 20)
    if variant != "clean":
        lines = open(nf).read().split("\n")
        if variant == "joined_lines":
            lines[40000] = lines[40000] + "  " + lines.pop(40001)
        else:
            lines.insert(50000, "END")
        open(nf, "w").write("\n".join(lines))
    case = lbm.Case.load(nf, bf)
    o = O.Oracle.from_files(nf, bf, O.params())
    nn = 50000 if variant == "stop_token" else 259 * 260
    assert (case.n, case.m, case.info.num_nodes) == (o.n, o.m, nn)
    _cells_match(case, o)


def test_shipped_coarse_case_parses_identically():
    path = "/root/reference/cases/coarse/Mesh"
    if not os.path.isdir(path):
        pytest.skip("reference checkout not present")
    case = lbm.Case.load(path + "/D2node.dat", path + "/BCconnectors.dat")
    o = O.Oracle.from_files(path + "/D2node.dat", path + "/BCconnectors.dat", O.params())
    _cells_match(case, o)


def test_missing_file_is_an_error():
    with pytest.raises(lbm.LbmError) as e:
        lbm.Case.load("/nonexistent/D2node.dat", "/nonexistent/BC.dat")
    assert e.value.code == lbm.EIO


def test_cavity_generator_matches_reference_format(tmp_path):
    """lbm_case_cavity == the shipped convention: same per-cell set-up as the
    test-side generator (which reproduces cases/coarse), and the written files
    parse back to the same cells (coordinates at %g precision)."""
    for N in (5, 16, 125):
        case = lbm.Case.cavity(N)
        nodes, conn = meshgen.cavity(N).arrays()
        ref = lbm.Case.from_arrays(nodes, conn)
        a, b = case.cells(), ref.cells()
        for f in ("fluid", "corner", "stream", "boundary", "boundary_id", "bc"):
            assert np.array_equal(a[f], b[f]), (N, f)
        assert np.allclose(a["q"], b["q"], atol=1e-5)
        assert np.allclose(a["x"], b["x"], atol=1e-6)
    case = lbm.Case.cavity(125)
    case.write(str(tmp_path / "n.dat"), str(tmp_path / "c.dat"))
    nodes, conn = meshgen.cavity(125).arrays()
    assert open(tmp_path / "n.dat").read().splitlines() == list(meshgen.cavity(125).node_lines())
    back = lbm.Case.load(str(tmp_path / "n.dat"), str(tmp_path / "c.dat")).cells()
    ref = lbm.Case.from_arrays(nodes, conn).cells()
    for f in ref:
        assert np.array_equal(back[f], ref[f]), f


def test_channel_generator_matches_test_generator():
    case = lbm.Case.channel(24, 10)
    nodes, conn = meshgen.channel(24, 10).arrays()
    ref = lbm.Case.from_arrays(nodes, conn)
    a, b = case.cells(), ref.cells()
    for f in ("fluid", "corner", "stream", "boundary", "boundary_id", "bc"):
        assert np.array_equal(a[f], b[f]), f
    # obstacle: every fluid->obstacle link is a wall with 0 < q <= 1
    c = lbm.Case.channel(60, 30, 20.3, 15.6, 5.2).cells()
    wall5 = (c["boundary_id"] == 5)
    assert wall5.sum() > 10
    q = c["q"][wall5][c["bc"][wall5] == 1]
    assert np.all(q > 0) and np.all(q <= 1.0 + 1e-6) and q.std() > 0.05


def test_cavity_ex_square_equals_shipped_cavity():
    """An obstacle-free square cavity_ex is the shipped cavity cell for cell."""
    for N in (5, 16, 33):
        a = lbm.Case.cavity_ex(N, N).cells()
        b = lbm.Case.cavity(N).cells()
        for f in a:
            assert np.array_equal(a[f], b[f]), (N, f)


def test_cavity_ex_rectangles_and_aspect():
    NX, NY = 30, 18
    rects = [(4, 3, 9, 7), (20, 10, 25, 16)]
    case = lbm.Case.cavity_ex(NX, NY, rects)
    assert (case.n, case.m) == (NX + 2, NY + 3)
    c = case.cells()
    fl = c["fluid"].reshape(case.m, case.n)
    for i0, j0, i1, j1 in rects:
        assert not fl[j0:j1 + 1, i0:i1 + 1].any()
    interior = fl[1:NY + 1, 1:NX + 1]
    blocked = sum((i1 - i0 + 1) * (j1 - j0 + 1) for i0, j0, i1, j1 in rects)
    assert interior.sum() == NX * NY - blocked
    # fluid nodes next to a rectangle carry wall links (type 1) of boundary id 5
    bc = c["bc"].reshape(case.m, case.n, 9)
    bid = c["boundary_id"].reshape(case.m, case.n)
    assert bc[5, 3, 1] == 1 and bid[5, 3] == 5      # west face of the first rectangle
    assert bc[2, 6, 2] == 1 and bid[2, 6] == 5      # below it: link north into it
    # Delta from the longer side
    assert abs(c["x"].reshape(case.m, case.n)[1, 2] - c["x"].reshape(case.m, case.n)[1, 1]
               - 1.0 / NX) < 1e-6


def test_pack_keeps_every_instance_cell_for_cell():
    cases = [lbm.Case.cavity_ex(20, 12, [(3, 3, 6, 6)]), lbm.Case.cavity_ex(9, 25),
             lbm.Case.cavity_ex(14, 14), lbm.Case.cavity_ex(40, 8, [(10, 2, 12, 5)])]
    packed = lbm.Case.pack(cases)
    assert packed.instances == 4 and cases[0].instances == 0
    pc = packed.cells()
    owner = np.full((packed.m, packed.n), -1)
    for k, (c, (x0, y0, w, h)) in enumerate(zip(cases, packed.rects)):
        assert (w, h) == (c.n, c.m)
        assert (owner[y0:y0 + h, x0:x0 + w] == -1).all(), "rectangles overlap"
        owner[y0:y0 + h, x0:x0 + w] = k
        cc = c.cells()
        for f in ("fluid", "corner", "stream", "boundary", "boundary_id", "bc", "q", "x", "y"):
            a = pc[f].reshape(packed.m, packed.n, -1)[y0:y0 + h, x0:x0 + w]
            b = cc[f].reshape(c.m, c.n, -1)
            assert np.array_equal(a, b), (k, f)
    assert not pc["fluid"].reshape(packed.m, packed.n)[owner < 0].any()
    # fluid nodes of different instances are never neighbours
    fl = pc["fluid"].reshape(packed.m, packed.n).astype(bool)
    for dj, di in ((0, 1), (1, 0), (1, 1), (1, -1)):
        a = owner[max(dj, 0):packed.m, max(di, 0):packed.n + min(di, 0)]
        b = owner[:packed.m - dj, max(-di, 0):packed.n - max(di, 0)]
        fa = fl[max(dj, 0):packed.m, max(di, 0):packed.n + min(di, 0)]
        fb = fl[:packed.m - dj, max(-di, 0):packed.n - max(di, 0)]
        assert not (fa & fb & (a != b)).any()
    with pytest.raises(lbm.LbmError):
        packed.write("/tmp/never_n.dat", "/tmp/never_c.dat")


def test_ini_reader(tmp_path):
    txt = "0.08\n0.0\n5\n0.01\n2\n3\n2\n2\n5000\n100\n10000\n1\n0\n-1.0\n-1.0\n0\n\n:01: Mean U\n"
    (tmp_path / "SetUpData.ini").write_text(txt)
    ini = lbm.read_ini(str(tmp_path / "SetUpData.ini"))
    assert ini.collision_model == 3 and ini.iterations == 5000 and ini.rho_ini == 5.0
    o = lbm.options_from_ini(ini)
    assert o.collision == 3 and o.precision == lbm.PREC_F64 and o.use_stop_rule == 1
    assert o.max_iterations == 5000 and o.resid_kind == lbm.RESID_REFERENCE
    (tmp_path / "bad.ini").write_text("0.1 0.2\n")
    with pytest.raises(lbm.LbmError):
        lbm.read_ini(str(tmp_path / "bad.ini"))


def test_compute_residuals_rejects_bad_arguments_before_any_device_work():
    L = lbm.lib()
    for args in [(None, 8, 8, 10, 10, 0, lbm.RESID_FULL, 8),          # no f
                 (8, 8, 8, 10, 0, 0, lbm.RESID_FULL, 8),              # m = 0
                 (8, 8, 8, 10, 10, 0, lbm.RESID_VELOCITY, 8),         # not a sweep kind
                 (8, 8, 8, 10, 10, 0, lbm.RESID_FULL, None)]:         # no out
        f, metaf, bid, n, m, dl, kind, out = args
        rc = L.lbm_compute_residuals(f, metaf, None, None, None, n, m, dl, kind, out, None)
        assert rc == lbm.EINVAL, args
