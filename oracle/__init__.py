"""CPU oracle for the D2Q9 hot path — test infrastructure only.

Python ctypes binding of ``oracle/liblbm_oracle.so`` (a C restatement of the
reference serial solver, see ``oracle/lbm_oracle.h``) plus helpers that run the
reference program built from its own sources into ``oracle/_ref``.  Only
``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` import this package; the product never does.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblbm_oracle.so")
REF_DIR = os.path.join(HERE, "_ref")
REF_SOLVER = os.path.join(REF_DIR, "LBMSolver_dump")

# numpy view of struct or_cell == struct CellProps (528 bytes)
CELL_DTYPE = np.dtype([
    ("fluid", "<i4"), ("corner", "<i4"), ("stream", "<i4", 9), ("id", "<i4"),
    ("boundary", "<i4"), ("boundary_id", "<i4"), ("bc", "<i4", 9),
    ("pad", "<i4"),
    ("q", "<f8", 9), ("x", "<f8"), ("y", "<f8"), ("u", "<f8"), ("v", "<f8"),
    ("rho", "<f8"), ("uo", "<f8"), ("vo", "<f8"), ("drag", "<f8"), ("lift", "<f8"),
    ("f", "<f8", 9), ("feq", "<f8", 9), ("metaf", "<f8", 9), ("fnb", "<f8", 9),
])
assert CELL_DTYPE.itemsize == 528


class Params(C.Structure):
    _fields_ = [
        ("uavg", C.c_float), ("vavg", C.c_float), ("rho_ini", C.c_float),
        ("viscosity", C.c_float),
        ("inlet_profile", C.c_int), ("collision", C.c_int), ("curved", C.c_int),
        ("outlet_profile", C.c_int), ("iterations", C.c_int),
        ("autosave_every", C.c_int), ("autosave_after", C.c_int),
        ("postproc", C.c_int), ("drag_lift", C.c_int),
        ("crit_v", C.c_float), ("crit_rho", C.c_float), ("read_former", C.c_int),
        ("init_rest", C.c_int), ("trt_fp64", C.c_int), ("trt_lambda", C.c_double),
        ("mrt_textbook", C.c_int), ("mrt_custom_s", C.c_int),
        ("mrt_s", C.c_double * 9), ("resid_kind", C.c_int), ("resid_every", C.c_int),
        ("outlet1_lid", C.c_int), ("outlet_rho", C.c_double),
    ]


_lib = None


def build(quiet: bool = True) -> None:
    """Compile the oracle (and oracle/_ref when the reference is present)."""
    out = subprocess.run(["make", "-C", HERE, "all"], capture_output=True, text=True)
    if out.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + out.stdout + out.stderr)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.or_create.restype = C.c_void_p
        L.or_create.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.POINTER(Params)]
        L.or_create_from_files.restype = C.c_void_p
        L.or_create_from_files.argtypes = [C.c_char_p, C.c_char_p, C.POINTER(Params)]
        L.or_read_ini.argtypes = [C.c_char_p, C.POINTER(Params)]
        L.or_destroy.argtypes = [C.c_void_p]
        L.or_step.argtypes = [C.c_void_p, C.c_int]
        L.or_dims.argtypes = [C.c_void_p] + [C.c_void_p] * 6
        L.or_iter.argtypes = [C.c_void_p]
        L.or_residuals.argtypes = [C.c_void_p, C.c_void_p]
        L.or_omegas.argtypes = [C.c_void_p, C.c_void_p]
        L.or_cells.restype = C.c_void_p
        L.or_cells.argtypes = [C.c_void_p]
        L.or_get_macro.argtypes = [C.c_void_p] * 4
        L.or_get_f.argtypes = [C.c_void_p] * 3
        L.or_write_results.argtypes = [C.c_void_p, C.c_char_p, C.c_int, C.c_int]
        L.or_mrt_init.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.or_d2q9.argtypes = [C.c_void_p] * 5 + [C.c_int]
        L.or_force.argtypes = [C.c_void_p, C.c_int, C.c_void_p]
        _lib = L
    return _lib


def params(**kw) -> Params:
    """Params with the reference's shipped-case defaults, overridden by kw."""
    p = Params()
    defaults = dict(uavg=0.08, vavg=0.0, rho_ini=5.0, viscosity=0.01, inlet_profile=2,
                    collision=1, curved=2, outlet_profile=2, iterations=10**9,
                    autosave_every=100, autosave_after=10**9, postproc=1, drag_lift=0,
                    crit_v=-1.0, crit_rho=-1.0, read_former=0, resid_every=10)
    defaults.update(kw)
    for k, v in defaults.items():
        if k == "mrt_s":
            p.mrt_s = (C.c_double * 9)(*v)
            p.mrt_custom_s = 1
        else:
            setattr(p, k, v)
    return p


class Oracle:
    """One CPU solver instance (struct or_solver)."""

    def __init__(self, nodes: np.ndarray, conn: np.ndarray, p: Params):
        L = lib()
        nodes = np.ascontiguousarray(nodes, dtype=np.float32)
        conn = np.ascontiguousarray(conn, dtype=np.float32).reshape(-1, 7)
        self._p = p
        self.h = L.or_create(nodes.ctypes.data, nodes.shape[0], conn.ctypes.data,
                             conn.shape[0], C.byref(p))
        if not self.h:
            raise RuntimeError("or_create failed")
        n, m = C.c_int(), C.c_int()
        d = C.c_double()
        ni = C.c_int()
        mn, mx = C.c_double(), C.c_double()
        L.or_dims(self.h, C.byref(n), C.byref(m), C.byref(d), C.byref(ni), C.byref(mn), C.byref(mx))
        self.n, self.m, self.delta = n.value, m.value, d.value
        self.num_inlet, self.min_inlet_y, self.max_inlet_y = ni.value, mn.value, mx.value

    @classmethod
    def from_files(cls, node_file: str, bc_file: str, p: Params) -> "Oracle":
        self = cls.__new__(cls)
        L = lib()
        self._p = p
        self.h = L.or_create_from_files(node_file.encode(), bc_file.encode(), C.byref(p))
        if not self.h:
            raise RuntimeError("or_create_from_files failed")
        n, m = C.c_int(), C.c_int()
        d = C.c_double()
        ni = C.c_int()
        mn, mx = C.c_double(), C.c_double()
        L.or_dims(self.h, C.byref(n), C.byref(m), C.byref(d), C.byref(ni), C.byref(mn), C.byref(mx))
        self.n, self.m, self.delta = n.value, m.value, d.value
        self.num_inlet, self.min_inlet_y, self.max_inlet_y = ni.value, mn.value, mx.value
        return self

    def __del__(self):
        h = getattr(self, "h", None)
        if h:
            lib().or_destroy(h)
            self.h = None

    def step(self, nsteps: int) -> int:
        return lib().or_step(self.h, nsteps)

    @property
    def iter(self) -> int:
        return lib().or_iter(self.h)

    def residuals(self) -> np.ndarray:
        out = np.zeros(4)
        lib().or_residuals(self.h, out.ctypes.data)
        return out

    def force(self, bid: int) -> np.ndarray:
        out = np.zeros(2)
        lib().or_force(self.h, bid, out.ctypes.data)
        return out

    def omegas(self) -> np.ndarray:
        out = np.zeros(2)
        lib().or_omegas(self.h, out.ctypes.data)
        return out

    def cells(self) -> np.ndarray:
        """Copy of the n*m cell array (CELL_DTYPE), storage order j*n+i."""
        ptr = lib().or_cells(self.h)
        buf = (C.c_char * (528 * self.n * self.m)).from_address(ptr)
        return np.frombuffer(buf, dtype=CELL_DTYPE).copy()

    def macro(self):
        nc = self.n * self.m
        rho, u, v = np.zeros(nc), np.zeros(nc), np.zeros(nc)
        lib().or_get_macro(self.h, rho.ctypes.data, u.ctypes.data, v.ctypes.data)
        return rho, u, v

    def write_results(self, path: str, final: bool, row0_bug: bool = True) -> None:
        lib().or_write_results(self.h, path.encode(), int(final), int(row0_bug))


def read_cells_bin(path: str) -> tuple[int, int, np.ndarray]:
    """Read Results/cells.bin written by the reference's capture hook."""
    with open(path, "rb") as fh:
        hdr = np.frombuffer(fh.read(12), dtype="<i4")
        n, m, sz = (int(x) for x in hdr)
        assert sz == 528, sz
        cells = np.frombuffer(fh.read(), dtype=CELL_DTYPE)
    assert cells.shape[0] == n * m
    return n, m, cells


def ref_available() -> bool:
    return os.path.exists(REF_SOLVER)


def run_reference(case_dir: str, workdir: str, ini_values: list) -> str:
    """Run the reference program (oracle/_ref, built from its own sources) in
    ``workdir`` on the mesh files of ``case_dir`` with the 16 ini values given.
    Returns the Results directory path."""
    import shutil
    os.makedirs(os.path.join(workdir, "Mesh"), exist_ok=True)
    os.makedirs(os.path.join(workdir, "Results"), exist_ok=True)
    for fn in ("D2node.dat", "BCconnectors.dat"):
        shutil.copy(os.path.join(case_dir, "Mesh", fn), os.path.join(workdir, "Mesh", fn))
    with open(os.path.join(workdir, "SetUpData.ini"), "w") as fh:
        fh.write("\n".join(str(v) for v in ini_values) + "\n")
    subprocess.run([REF_SOLVER], cwd=workdir, stdout=subprocess.DEVNULL,
                   stderr=subprocess.DEVNULL, check=False, timeout=600)
    return os.path.join(workdir, "Results")
