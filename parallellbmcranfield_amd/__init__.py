"""MI355X-native D2Q9 lattice-Boltzmann engine — Python host interface.

A ctypes binding of the C ABI in ``include/lbm.h`` (library
``parallellbmcranfield_amd/lib/liblbm.so``, HIP kernels for gfx950).  It
mirrors the reference serial solver's interface
(mate-szoke/ParallelLbmCranfield serial/include/Iterate.h:21-27,
serial/include/FilesIO.h:8-23): :func:`Iteration` takes the same 18 arguments
as the reference's ``Iteration`` and writes the same ``Results/`` files;
:class:`Case` and :class:`Solver` expose the set-up and the time loop step by
step.  There is no fallback: without the built library every call raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
# LBM_LIB selects another build of the same library (e.g. a tuning variant).
LIB_PATH = os.environ.get("LBM_LIB") or os.path.join(_PKG, "lib", "liblbm.so")
CLI_PATH = os.path.join(_PKG, "bin", "lbm_solve")

OK, EIO, EPARSE, DIVERGED, EHIP, EINVAL, ENOMEM = range(7)
OUTLET_VELOCITY, OUTLET_OPEN2, OUTLET_OPEN1, OUTLET_PRESSURE = 1, 2, 3, 4
BGKW, TRT, MRT = 1, 2, 3
PREC_F64, PREC_F32, PREC_F16, PREC_F16_SR = 0, 1, 2, 3
INIT_REFERENCE, INIT_REST, INIT_FIELD, INIT_FIELD_NEQ = 0, 1, 2, 3
RESID_REFERENCE, RESID_FULL, RESID_VELOCITY, RESID_NONE = 0, 1, 2, 3
FORMAT_CSV, FORMAT_TECPLOT = 1, 2
(TUNE_XCD_REMAP, TUNE_SPLIT_BOUNDARY, TUNE_NO_GRAPH, TUNE_NO_PAIRS, TUNE_NO_TILE_WALLS,
 TUNE_NO_TWO_STEP, TUNE_NO_THREE_STEP, TUNE_GRAPH) = (1, 2, 4, 8, 16, 32, 64, 128)

# Every function include/lbm.h declares (checked by tests/test_abi.py).
EXPORTS = [
    "lbm_last_error", "lbm_abi_version", "lbm_read_ini", "lbm_options_from_ini",
    "lbm_case_load", "lbm_case_from_arrays", "lbm_case_cavity", "lbm_case_channel",
    "lbm_case_cavity_ex", "lbm_case_pack", "lbm_case_instances",
    "lbm_solver_create_batched", "lbm_solver_create_batched_field", "lbm_batch_status",
    "lbm_case_write", "lbm_case_info_get", "lbm_case_cells", "lbm_case_free",
    "lbm_solver_create", "lbm_solver_create_field", "lbm_solver_destroy", "lbm_solver_step",
    "lbm_solver_step_timed", "lbm_solver_sync", "lbm_solver_iter", "lbm_solver_residuals",
    "lbm_solver_residual_history", "lbm_solver_force_history", "lbm_solver_get_macro",
    "lbm_solver_get_f",
    "lbm_solver_write_results", "lbm_solver_timing", "lbm_solver_device_bytes",
    "lbm_solver_fused", "lbm_solver_two_step_info", "lbm_solver_three_step_info", "lbm_solver_save", "lbm_solver_load", "lbm_device_copy_gbs",
    "lbm_compute_residuals", "lbm_compute_residuals_noalloc", "lbm_device_stream_probe",
    "lbm_run",
    "lbm_run_checkpointed", "lbm_iteration",
    "lbm_net_create", "lbm_net_destroy", "lbm_net_num_params", "lbm_net_get_params",
    "lbm_net_set_params", "lbm_net_predict", "lbm_net_gradients", "lbm_net_set_data",
    "lbm_net_train", "lbm_net_set_weight_decay", "lbm_net_set_div_weight", "lbm_net_save", "lbm_net_load",
]


class LbmError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"lbm status {code}: {msg}")
        self.code = code


class Ini(C.Structure):
    """The 16 positional values of SetUpData.ini (serial/FileIO.c:263-289)."""
    _fields_ = [
        ("uavg", C.c_float), ("vavg", C.c_float), ("rho_ini", C.c_float),
        ("viscosity", C.c_float), ("inlet_profile", C.c_int), ("collision_model", C.c_int),
        ("curved_boundaries", C.c_int), ("outlet_profile", C.c_int), ("iterations", C.c_int),
        ("autosave_every", C.c_int), ("autosave_after", C.c_int), ("postproc_prog", C.c_int),
        ("calculate_drag_lift", C.c_int), ("convergence_crit_veloc", C.c_float),
        ("convergence_crit_rho", C.c_float), ("read_former_data", C.c_int),
    ]


class Options(C.Structure):
    _fields_ = [
        ("collision", C.c_int), ("curved", C.c_int), ("inlet_profile", C.c_int),
        ("outlet_profile", C.c_int), ("drag_lift", C.c_int),
        ("uavg", C.c_float), ("vavg", C.c_float), ("rho_ini", C.c_float),
        ("viscosity", C.c_float),
        ("precision", C.c_int), ("init", C.c_int), ("trt_fp64_parts", C.c_int),
        ("trt_lambda", C.c_double), ("mrt_textbook_meq", C.c_int),
        ("mrt_custom_rates", C.c_int), ("mrt_rates", C.c_double * 9),
        ("resid_kind", C.c_int), ("resid_every", C.c_int), ("use_stop_rule", C.c_int),
        ("crit_v", C.c_float), ("crit_rho", C.c_float), ("max_iterations", C.c_int),
        ("f16_scale", C.c_float), ("block_size", C.c_int), ("tune_flags", C.c_int),
        ("xcd_group", C.c_int), ("outlet_rho", C.c_double), ("force_id", C.c_int),
        ("force_every", C.c_int), ("two_step_rows", C.c_int), ("three_step_rows", C.c_int),
    ]


class CaseInfo(C.Structure):
    _fields_ = [
        ("n", C.c_int), ("m", C.c_int), ("num_nodes", C.c_int), ("num_conn", C.c_int),
        ("num_fluid", C.c_int), ("num_boundary", C.c_int), ("num_inlet", C.c_int),
        ("delta", C.c_double), ("min_inlet_y", C.c_double), ("max_inlet_y", C.c_double),
    ]


class Instance(C.Structure):
    _fields_ = [("viscosity", C.c_float), ("uavg", C.c_float)]


class Timing(C.Structure):
    _fields_ = [("bulk_ms", C.c_double), ("boundary_ms", C.c_double), ("outlet_ms", C.c_double),
                ("residual_ms", C.c_double), ("observe_ms", C.c_double), ("init_ms", C.c_double),
                ("steps", C.c_long)]


_lib = None


def lib() -> C.CDLL:
    """Load liblbm.so (built by ``__graft_entry__.build()``); raises if absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"native library not built: {LIB_PATH} (run __graft_entry__.build())")
    L = C.CDLL(LIB_PATH)
    P, I, D = C.c_void_p, C.c_int, C.c_double
    sig = {
        "lbm_last_error": (C.c_char_p, []),
        "lbm_abi_version": (I, []),
        "lbm_read_ini": (I, [C.c_char_p, C.POINTER(Ini)]),
        "lbm_options_from_ini": (None, [C.POINTER(Ini), C.POINTER(Options)]),
        "lbm_case_load": (I, [C.c_char_p, C.c_char_p, C.POINTER(P)]),
        "lbm_case_from_arrays": (I, [P, I, P, I, C.POINTER(P)]),
        "lbm_case_cavity": (I, [I, I, C.POINTER(P)]),
        "lbm_case_channel": (I, [I, I, D, D, D, C.POINTER(P)]),
        "lbm_case_cavity_ex": (I, [I, I, I, P, I, C.POINTER(P)]),
        "lbm_case_pack": (I, [P, I, C.POINTER(P), P]),
        "lbm_case_instances": (I, [P]),
        "lbm_solver_create_batched": (I, [P, C.POINTER(Options), P, I, C.POINTER(P)]),
        "lbm_batch_status": (I, [P, I, P, P, P]),
        "lbm_solver_create_batched_field": (I, [P, C.POINTER(Options), P, I, P, P, P,
                                                 C.POINTER(P)]),
        "lbm_case_write": (I, [P, C.c_char_p, C.c_char_p]),
        "lbm_case_info_get": (I, [P, C.POINTER(CaseInfo)]),
        "lbm_case_cells": (I, [P] * 11),
        "lbm_case_free": (None, [P]),
        "lbm_solver_create": (I, [P, C.POINTER(Options), C.POINTER(P)]),
        "lbm_solver_create_field": (I, [P, C.POINTER(Options), P, P, P, C.POINTER(P)]),
        "lbm_solver_destroy": (None, [P]),
        "lbm_solver_step": (I, [P, I, C.POINTER(I)]),
        "lbm_solver_step_timed": (I, [P, I, I, C.POINTER(D), C.POINTER(D)]),
        "lbm_solver_sync": (I, [P]),
        "lbm_solver_iter": (I, [P]),
        "lbm_solver_residuals": (I, [P, P]),
        "lbm_solver_residual_history": (I, [P, I, P, P, C.POINTER(I)]),
        "lbm_solver_force_history": (I, [P, I, P, P, C.POINTER(I)]),
        "lbm_solver_get_macro": (I, [P, P, P, P]),
        "lbm_solver_get_f": (I, [P, P, P]),
        "lbm_solver_write_results": (I, [P, C.c_char_p, I, I, I]),
        "lbm_solver_timing": (I, [P, C.POINTER(Timing)]),
        "lbm_solver_device_bytes": (C.c_size_t, [P]),
        "lbm_device_copy_gbs": (I, [C.c_size_t, I, C.POINTER(D)]),
        "lbm_device_stream_probe": (I, [I, I, I, I, C.POINTER(D), C.POINTER(D)]),
        "lbm_solver_fused": (I, [P]),
        "lbm_solver_two_step_info": (I, [P, P, P, P, P]),
        "lbm_solver_three_step_info": (I, [P, P, P, P, P]),
        "lbm_compute_residuals": (I, [P, P, P, P, P, I, I, I, I, P, P]),
        "lbm_compute_residuals_noalloc": (I, [P, P, P, P, P, I, I, I, I, P, P]),
        "lbm_solver_save": (I, [P, C.c_char_p]),
        "lbm_solver_load": (I, [P, C.c_char_p]),
        "lbm_run_checkpointed": (I, [C.c_char_p, C.c_char_p, C.c_char_p, C.c_char_p,
                                     C.POINTER(Options), C.c_char_p, C.c_char_p, I]),
        "lbm_run": (I, [C.c_char_p, C.c_char_p, C.c_char_p, C.c_char_p, C.POINTER(Options), I]),
        "lbm_net_create": (I, [I, I, I, I, I, I, I, C.c_uint, C.POINTER(P)]),
        "lbm_net_destroy": (None, [P]),
        "lbm_net_num_params": (C.c_long, [P]),
        "lbm_net_get_params": (I, [P, P]),
        "lbm_net_set_params": (I, [P, P]),
        "lbm_net_predict": (I, [P, P, I, P]),
        "lbm_net_gradients": (I, [P, P, P, I, C.POINTER(D), P]),
        "lbm_net_set_data": (I, [P, P, P, I]),
        "lbm_net_train": (I, [P, P, I, C.c_float, C.POINTER(D)]),
        "lbm_net_set_weight_decay": (I, [P, C.c_float]),
        "lbm_net_set_div_weight": (I, [P, C.c_float, C.c_float, C.c_float]),
        "lbm_net_save": (I, [P, C.c_char_p]),
        "lbm_net_load": (I, [P, C.c_char_p]),
        "lbm_iteration": (I, [C.c_char_p, C.c_char_p, C.c_float, C.c_float, C.c_float,
                              C.c_float, I, I, I, I, I, I, I, I, I, C.c_float, C.c_float, I]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def _check(rc: int) -> None:
    if rc != OK:
        raise LbmError(rc, lib().lbm_last_error().decode(errors="replace"))


def _ptr(a: np.ndarray | None):
    return None if a is None else a.ctypes.data


def read_ini(path: str) -> Ini:
    """ReadIniData (serial/FileIO.c:263-289)."""
    ini = Ini()
    _check(lib().lbm_read_ini(path.encode(), C.byref(ini)))
    return ini


def options_from_ini(ini: Ini, **overrides) -> Options:
    """Reference behaviour for an ini, then keyword overrides (Options fields)."""
    o = Options()
    lib().lbm_options_from_ini(C.byref(ini), C.byref(o))
    for k, v in overrides.items():
        if k == "mrt_rates":
            o.mrt_rates = (C.c_double * 9)(*v)
            o.mrt_custom_rates = 1
        else:
            setattr(o, k, v)
    return o


def make_options(collision=BGKW, uavg=0.08, vavg=0.0, rho_ini=5.0, viscosity=0.01,
                 inlet_profile=2, curved=2, outlet_profile=2, drag_lift=0, **overrides) -> Options:
    """Options from ini-style values (reference defaults of the shipped cases)."""
    ini = Ini(uavg=uavg, vavg=vavg, rho_ini=rho_ini, viscosity=viscosity,
              inlet_profile=inlet_profile, collision_model=collision, curved_boundaries=curved,
              outlet_profile=outlet_profile, iterations=2**31 - 1, autosave_every=100,
              autosave_after=2**31 - 1, postproc_prog=1, calculate_drag_lift=drag_lift,
              convergence_crit_veloc=-1.0, convergence_crit_rho=-1.0, read_former_data=0)
    return options_from_ini(ini, **overrides)


class Case:
    """A parsed mesh: ReadNodes + ReadBCconn + CompDataNode + CompDataConn + CellIni."""

    def __init__(self, handle):
        self.h = handle
        info = CaseInfo()
        _check(lib().lbm_case_info_get(self.h, C.byref(info)))
        self.info = info
        self.n, self.m = info.n, info.m

    @classmethod
    def load(cls, node_file: str, bc_file: str) -> "Case":
        h = C.c_void_p()
        _check(lib().lbm_case_load(node_file.encode(), bc_file.encode(), C.byref(h)))
        return cls(h)

    @classmethod
    def from_arrays(cls, nodes: np.ndarray, conn: np.ndarray) -> "Case":
        nodes = np.ascontiguousarray(nodes, dtype=np.float32).reshape(-1, 5)
        conn = np.ascontiguousarray(conn, dtype=np.float32).reshape(-1, 7)
        h = C.c_void_p()
        _check(lib().lbm_case_from_arrays(_ptr(nodes), nodes.shape[0], _ptr(conn), conn.shape[0],
                                          C.byref(h)))
        return cls(h)

    @classmethod
    def cavity(cls, N: int, lid_type: int = 2) -> "Case":
        h = C.c_void_p()
        _check(lib().lbm_case_cavity(N, lid_type, C.byref(h)))
        return cls(h)

    @classmethod
    def channel(cls, NX: int, NY: int, cx: float = 0.0, cy: float = 0.0, r: float = 0.0) -> "Case":
        h = C.c_void_p()
        _check(lib().lbm_case_channel(NX, NY, cx, cy, r, C.byref(h)))
        return cls(h)

    @classmethod
    def cavity_ex(cls, NX: int, NY: int, rects=(), lid_type: int = 2) -> "Case":
        """NX x NY lid-driven cavity with solid rectangles (i0, j0, i1, j1),
        inclusive fluid-node indices."""
        r = np.ascontiguousarray(np.asarray(rects, dtype=np.int32).reshape(-1, 4))
        h = C.c_void_p()
        _check(lib().lbm_case_cavity_ex(NX, NY, lid_type, _ptr(r) if r.size else None,
                                        r.shape[0], C.byref(h)))
        return cls(h)

    @classmethod
    def pack(cls, cases) -> "Case":
        """Pack several cases into one lattice for a batched solver; the
        instance rectangles (x0, y0, n, m) are in ``.rects``."""
        arr = (C.c_void_p * len(cases))(*[c.h.value if isinstance(c.h, C.c_void_p) else c.h
                                          for c in cases])
        rects = np.zeros((len(cases), 4), np.int32)
        h = C.c_void_p()
        _check(lib().lbm_case_pack(C.cast(arr, C.c_void_p), len(cases), C.byref(h), _ptr(rects)))
        self = cls(h)
        self.rects = rects
        return self

    @property
    def instances(self) -> int:
        return lib().lbm_case_instances(self.h)

    def write(self, node_file: str, bc_file: str) -> None:
        _check(lib().lbm_case_write(self.h, node_file.encode(), bc_file.encode()))

    def cells(self) -> dict:
        nc = self.n * self.m
        out = {k: np.zeros(nc, np.int32) for k in ("fluid", "corner", "id", "boundary", "boundary_id")}
        out["stream"] = np.zeros((nc, 9), np.int32)
        out["bc"] = np.zeros((nc, 9), np.int32)
        out["q"] = np.zeros((nc, 9))
        out["x"] = np.zeros(nc)
        out["y"] = np.zeros(nc)
        _check(lib().lbm_case_cells(self.h, _ptr(out["fluid"]), _ptr(out["corner"]),
                                    _ptr(out["stream"]), _ptr(out["id"]), _ptr(out["boundary"]),
                                    _ptr(out["boundary_id"]), _ptr(out["bc"]), _ptr(out["q"]),
                                    _ptr(out["x"]), _ptr(out["y"])))
        return out

    def __del__(self):
        h = getattr(self, "h", None)
        if h and _lib is not None:
            _lib.lbm_case_free(h)
            self.h = None


class Solver:
    """Device state of one lattice and its time loop (serial/Iterate.c:168-332)."""

    def __init__(self, case: Case, opts: Options, field=None, instances=None):
        """instances: for a packed case, a sequence of (viscosity, lid speed)."""
        self.case = case  # keep alive: the solver references the case
        self.opts = opts
        h = C.c_void_p()
        if instances is not None:
            arr = (Instance * len(instances))(*[Instance(float(nu), float(u))
                                                for nu, u in instances])
            if field is None:
                _check(lib().lbm_solver_create_batched(case.h, C.byref(opts),
                                                       C.cast(arr, C.c_void_p), len(instances),
                                                       C.byref(h)))
            else:
                rho, u, v = (np.ascontiguousarray(a, dtype=np.float64).ravel() for a in field)
                _check(lib().lbm_solver_create_batched_field(
                    case.h, C.byref(opts), C.cast(arr, C.c_void_p), len(instances), _ptr(rho),
                    _ptr(u), _ptr(v), C.byref(h)))
        elif field is None:
            _check(lib().lbm_solver_create(case.h, C.byref(opts), C.byref(h)))
        else:
            rho, u, v = (np.ascontiguousarray(a, dtype=np.float64).ravel() for a in field)
            _check(lib().lbm_solver_create_field(case.h, C.byref(opts), _ptr(rho), _ptr(u),
                                                 _ptr(v), C.byref(h)))
        self.h = h
        self.n, self.m = case.n, case.m

    def step(self, nsteps: int = 1) -> int:
        done = C.c_int()
        _check(lib().lbm_solver_step(self.h, nsteps, C.byref(done)))
        return done.value

    def step_timed(self, nsteps: int, sample_every: int = 1):
        """Run nsteps; returns (stream ms for all steps, average ms of one
        sampled step-kernel launch)."""
        tot, kern = C.c_double(), C.c_double()
        _check(lib().lbm_solver_step_timed(self.h, nsteps, sample_every, C.byref(tot),
                                           C.byref(kern)))
        return tot.value, kern.value

    def batch_status(self):
        """(conv_iter, conv_seconds, residual) arrays, one entry per instance;
        conv_iter -1 = still running, -2 = diverged."""
        k = lib().lbm_case_instances(self.case.h)
        it = np.zeros(k, np.int32)
        sec = np.zeros(k)
        res = np.zeros(k)
        _check(lib().lbm_batch_status(self.h, k, _ptr(it), _ptr(sec), _ptr(res)))
        return it, sec, res

    def sync(self) -> None:
        _check(lib().lbm_solver_sync(self.h))

    @property
    def iter(self) -> int:
        return lib().lbm_solver_iter(self.h)

    def residuals(self) -> np.ndarray:
        r = np.zeros(4)
        _check(lib().lbm_solver_residuals(self.h, _ptr(r)))
        return r

    def residual_history(self):
        its, vals = [], []
        while True:
            cap = 4096
            i = np.zeros(cap, np.int32)
            v = np.zeros((cap, 4))
            cnt = C.c_int()
            _check(lib().lbm_solver_residual_history(self.h, cap, _ptr(i), _ptr(v), C.byref(cnt)))
            its.append(i[:cnt.value])
            vals.append(v[:cnt.value])
            if cnt.value < cap:
                break
        return np.concatenate(its), np.concatenate(vals)

    def force_history(self):
        """(iterations, forces[k, 2]) sampled since the last call."""
        its, vals = [], []
        while True:
            cap = 4096
            i = np.zeros(cap, np.int32)
            v = np.zeros((cap, 2))
            cnt = C.c_int()
            _check(lib().lbm_solver_force_history(self.h, cap, _ptr(i), _ptr(v), C.byref(cnt)))
            its.append(i[:cnt.value])
            vals.append(v[:cnt.value])
            if cnt.value < cap:
                break
        return np.concatenate(its), np.concatenate(vals)

    def macro(self):
        nc = self.n * self.m
        rho, u, v = np.zeros(nc), np.zeros(nc), np.zeros(nc)
        _check(lib().lbm_solver_get_macro(self.h, _ptr(rho), _ptr(u), _ptr(v)))
        return rho, u, v

    def distributions(self):
        nc = self.n * self.m
        f, mf = np.zeros((nc, 9)), np.zeros((nc, 9))
        _check(lib().lbm_solver_get_f(self.h, _ptr(f), _ptr(mf)))
        return f, mf

    def write_results(self, path: str, fmt: int = FORMAT_CSV, final: bool = False,
                      row0_bug: bool = True) -> None:
        _check(lib().lbm_solver_write_results(self.h, path.encode(), fmt, int(final), int(row0_bug)))

    @property
    def fused(self) -> bool:
        return bool(lib().lbm_solver_fused(self.h) & 1)

    @property
    def paired(self) -> bool:
        """Plain steps run the paired-cell FP16 kernel."""
        return bool(lib().lbm_solver_fused(self.h) & 2)

    def two_step_info(self) -> dict:
        """Two-step launches: nodes advanced two steps per launch (0 = off),
        the kernel's tiles, and the nodes that take two single steps."""
        cells, band = C.c_longlong(0), C.c_longlong(0)
        tiles, rows = C.c_int(0), C.c_int(0)
        _check(lib().lbm_solver_two_step_info(self.h, C.byref(cells), C.byref(tiles),
                                              C.byref(rows), C.byref(band)))
        return dict(cells=cells.value, tiles=tiles.value, rows=rows.value, band=band.value)

    def three_step_info(self) -> dict:
        """Three-step launches: nodes advanced three steps per launch (0 =
        off), the kernel's tiles, and the nodes that take three single steps."""
        cells, band = C.c_longlong(0), C.c_longlong(0)
        tiles, rows = C.c_int(0), C.c_int(0)
        _check(lib().lbm_solver_three_step_info(self.h, C.byref(cells), C.byref(tiles),
                                                C.byref(rows), C.byref(band)))
        return dict(cells=cells.value, tiles=tiles.value, rows=rows.value, band=band.value)

    def save(self, path: str) -> None:
        """Binary checkpoint (lbm_solver_save)."""
        _check(lib().lbm_solver_save(self.h, path.encode()))

    def load(self, path: str) -> None:
        """Restore a checkpoint written by save() for the same case/precision."""
        _check(lib().lbm_solver_load(self.h, path.encode()))

    def device_bytes(self) -> int:
        return lib().lbm_solver_device_bytes(self.h)

    def close(self) -> None:
        h = getattr(self, "h", None)
        if h and _lib is not None:
            _lib.lbm_solver_destroy(h)
        self.h = None

    def __del__(self):
        self.close()


class Net:
    """The warm-start network (lbm_net_*): U-Net of 3x3 convolutions on an
    H x W grid, BF16 on the matrix cores, trained with Adam on the device.
    Inputs are [sample, H, W, in_channels] float32 (channel 0: fluid mask),
    targets and predictions [sample, H, W, 3]."""

    def __init__(self, H: int, W: int, batch: int = 16, base: int = 32, levels: int = 4,
                 seed: int = 0, in_channels: int = 3, dilation: int = 1):
        h = C.c_void_p()
        _check(lib().lbm_net_create(H, W, batch, base, levels, in_channels, dilation, seed,
                                    C.byref(h)))
        self.dilation = dilation
        self.h = h
        self.H, self.W, self.batch, self.base, self.levels = H, W, batch, base, levels
        self.in_channels = in_channels

    @property
    def num_params(self) -> int:
        return lib().lbm_net_num_params(self.h)

    def params(self) -> np.ndarray:
        out = np.zeros(self.num_params, np.float32)
        _check(lib().lbm_net_get_params(self.h, _ptr(out)))
        return out

    def set_params(self, p: np.ndarray) -> None:
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.num_params
        _check(lib().lbm_net_set_params(self.h, _ptr(p)))

    def _arr(self, a, ch=3):
        a = np.ascontiguousarray(a, dtype=np.float32)
        assert a.shape[1:] == (self.H, self.W, ch), a.shape
        return a

    def predict(self, inp: np.ndarray) -> np.ndarray:
        inp = self._arr(inp, self.in_channels)
        out = np.zeros(inp.shape[:3] + (3,), np.float32)
        _check(lib().lbm_net_predict(self.h, _ptr(inp), inp.shape[0], _ptr(out)))
        return out

    def gradients(self, inp: np.ndarray, target: np.ndarray):
        inp, target = self._arr(inp, self.in_channels), self._arr(target)
        loss = C.c_double()
        g = np.zeros(self.num_params, np.float32)
        _check(lib().lbm_net_gradients(self.h, _ptr(inp), _ptr(target), inp.shape[0],
                                       C.byref(loss), _ptr(g)))
        return loss.value, g

    def set_data(self, inp: np.ndarray, target: np.ndarray) -> None:
        inp, target = self._arr(inp, self.in_channels), self._arr(target)
        _check(lib().lbm_net_set_data(self.h, _ptr(inp), _ptr(target), inp.shape[0]))

    def train(self, idx, lr: float) -> float:
        idx = np.ascontiguousarray(idx, dtype=np.int32)
        loss = C.c_double()
        _check(lib().lbm_net_train(self.h, _ptr(idx), idx.size, lr, C.byref(loss)))
        return loss.value

    def set_div_weight(self, weight: float, su: float = 1.0, sv: float = 1.0) -> None:
        """Divergence penalty of the predicted velocity in the loss
        (lbm_net_set_div_weight); 0 turns it off."""
        _check(lib().lbm_net_set_div_weight(self.h, weight, su, sv))

    def set_weight_decay(self, wd: float) -> None:
        """Decoupled weight decay of the Adam step (AdamW); 0 is plain Adam."""
        _check(lib().lbm_net_set_weight_decay(self.h, wd))

    def save(self, path: str) -> None:
        _check(lib().lbm_net_save(self.h, path.encode()))

    def load(self, path: str) -> None:
        _check(lib().lbm_net_load(self.h, path.encode()))

    def close(self) -> None:
        if getattr(self, "h", None) and _lib is not None:
            _lib.lbm_net_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


def run(ini_file: str = "SetUpData.ini", node_file: str = "Mesh/D2node.dat",
        bc_file: str = "Mesh/BCconnectors.dat", out_dir: str = "", opts: Options | None = None,
        verbose: bool = False) -> int:
    """The reference program (serial/main.c + Iteration): returns the lbm status
    (DIVERGED where the reference printed DIVERGENCE! and exited 1)."""
    return lib().lbm_run(ini_file.encode(), node_file.encode(), bc_file.encode(), out_dir.encode(),
                         None if opts is None else C.byref(opts), int(verbose))


def Iteration(NodeDataFile, BCconnectorDataFile, Uavg, Vavg, rho_ini, Viscosity, InletProfile,
              CollisionModel, CurvedBoundaries, OutletProfile, Iterations, AutosaveAfter,
              AutosaveEvery, postproc_prog, CalculateDragLift, ConvergenceCritVeloc,
              ConvergenceCritRho, ReadFormerData) -> int:
    """Same arguments as the reference's Iteration (serial/include/Iterate.h:21-27);
    writes Results/ in the current directory."""
    return lib().lbm_iteration(NodeDataFile.encode(), BCconnectorDataFile.encode(), Uavg, Vavg,
                               rho_ini, Viscosity, InletProfile, CollisionModel, CurvedBoundaries,
                               OutletProfile, Iterations, AutosaveAfter, AutosaveEvery,
                               postproc_prog, CalculateDragLift, ConvergenceCritVeloc,
                               ConvergenceCritRho, ReadFormerData)


def device_stream_probe(n: int, m: int, elem_bytes: int = 4, reps: int = 20):
    """(GB/s, average us) of the step's loads and stores without arithmetic on
    an n x m lattice's layout (lbm_device_stream_probe)."""
    g, us = C.c_double(), C.c_double()
    _check(lib().lbm_device_stream_probe(n, m, elem_bytes, reps, C.byref(g), C.byref(us)))
    return g.value, us.value


def compute_residuals(f: int, metaf: int, n: int, m: int, out: int, kind: int = RESID_FULL,
                      boundary_id: int | None = None, drag: int | None = None,
                      lift: int | None = None, drag_lift: int = 0, stream: int | None = None):
    """ComputeResiduals (serial/ComputeResiduals.c:6-51) over device arrays
    (lbm_compute_residuals): f, metaf, boundary_id, drag, lift and out are
    device addresses (ints); out receives the four Residuals on the stream."""
    _check(lib().lbm_compute_residuals(f, metaf, boundary_id, drag, lift, n, m, drag_lift, kind,
                                       out, stream))


def device_copy_gbs(nbytes: int = 1 << 30, reps: int = 10) -> float:
    """Best-of-reps bandwidth (GB/s, read + write) of a plain device copy."""
    g = C.c_double()
    _check(lib().lbm_device_copy_gbs(nbytes, reps, C.byref(g)))
    return g.value


def strouhal(lift: np.ndarray, sample_steps: int, diameter: float, velocity: float,
             discard: float = 0.5) -> float:
    """Strouhal number f D / U from a lift series sampled every `sample_steps`
    lattice steps: the dominant frequency of the (mean-removed, Hann-windowed)
    last (1 - discard) of the series, refined by parabolic interpolation of
    the spectral peak."""
    x = np.asarray(lift, dtype=np.float64)
    x = x[int(len(x) * discard):]
    x = x - x.mean()
    n = len(x)
    if n < 16:
        raise ValueError("lift series too short")
    spec = np.abs(np.fft.rfft(x * np.hanning(n)))
    k = int(np.argmax(spec[1:]) + 1)
    delta = 0.0
    if 1 <= k < len(spec) - 1:
        a, b, c = np.log(spec[k - 1] + 1e-300), np.log(spec[k] + 1e-300), np.log(spec[k + 1] + 1e-300)
        den = a - 2 * b + c
        delta = 0.5 * (a - c) / den if den != 0 else 0.0
    f = (k + delta) / (n * sample_steps)  # cycles per lattice step
    return f * diameter / velocity


def gpu_available() -> bool:
    """True when a HIP device is visible (probed without torch)."""
    try:
        hip = C.CDLL("libamdhip64.so")
    except OSError:
        return False
    n = C.c_int(0)
    return hip.hipGetDeviceCount(C.byref(n)) == 0 and n.value > 0
