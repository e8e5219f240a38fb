"""aither_amd: MI355X-native hot path of the aither structured-grid CFD solver.

The compute path is the in-tree native library ``libaither_amd.so`` (hand-
written HIP kernels for gfx950 behind the C ABI of ``include/aither_amd.h``).
This module is a thin ctypes binding used by the tests and the benchmark; it
has no compute path of its own and raises if the native library is missing.

Typical use mirrors the reference's main loop (src/main.cpp:231-302)::

    s = Solver("case.inp")
    for nn in range(s.info.iterations):
        s.begin_step(nn)
        for mm in range(s.info.nonlinear_iterations):
            norms = s.iterate(nn, mm)
"""
import ctypes
import os

import numpy as np

__all__ = ["Solver", "Server", "SlotInfo", "Info", "Norms", "Timing", "AitherError", "load_library",
           "eval_face_flux", "eval_sst_cells", "inspect_info", "inspect_field", "LIB_PATH"]

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libaither_amd.so")
MAX_EQ = 8

ATH_OK = 0
ERRORS = {
    -1: "ATH_ERR_ARG", -2: "ATH_ERR_CONFIG", -3: "ATH_ERR_IO", -4: "ATH_ERR_NOMEM",
    -5: "ATH_ERR_NONPHYSICAL", -6: "ATH_ERR_HIP", -7: "ATH_ERR_DEVICE",
}


class AitherError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class Info(ctypes.Structure):
    _fields_ = [
        ("num_blocks", ctypes.c_int), ("num_equations", ctypes.c_int),
        ("num_ghosts", ctypes.c_int), ("iterations", ctypes.c_int),
        ("nonlinear_iterations", ctypes.c_int), ("implicit", ctypes.c_int),
        ("total_cells", ctypes.c_double), ("r_ref", ctypes.c_double),
        ("t_ref", ctypes.c_double), ("l_ref", ctypes.c_double), ("a_ref", ctypes.c_double),
        ("gas_constant", ctypes.c_double), ("gamma", ctypes.c_double),
        ("block_matrix", ctypes.c_int),
    ]


class Norms(ctypes.Structure):
    _fields_ = [
        ("num_equations", ctypes.c_int),
        ("l2_sq", ctypes.c_double * MAX_EQ), ("l2", ctypes.c_double * MAX_EQ),
        ("normalized", ctypes.c_double * MAX_EQ), ("linf", ctypes.c_double),
        ("linf_block", ctypes.c_int), ("linf_i", ctypes.c_int), ("linf_j", ctypes.c_int),
        ("linf_k", ctypes.c_int), ("linf_eqn", ctypes.c_int),
        ("matrix_resid", ctypes.c_double), ("cfl", ctypes.c_double),
    ]

    def as_dict(self):
        n = self.num_equations
        return {
            "l2_sq": list(self.l2_sq[:n]), "l2": list(self.l2[:n]),
            "normalized": list(self.normalized[:n]), "linf": self.linf,
            "linf_loc": (self.linf_block, self.linf_i, self.linf_j, self.linf_k, self.linf_eqn),
            "matrix_resid": self.matrix_resid, "cfl": self.cfl,
        }


class SlotInfo(ctypes.Structure):
    _fields_ = [
        ("state", ctypes.c_int), ("error", ctypes.c_int), ("iterations", ctypes.c_int),
        ("num_blocks", ctypes.c_int), ("cells", ctypes.c_double), ("device_bytes", ctypes.c_double),
        ("drop", ctypes.c_double), ("last", Norms),
    ]


class Timing(ctypes.Structure):
    _fields_ = [
        ("run_ms", ctypes.c_double), ("flux_ms", ctypes.c_double),
        ("flux_launches", ctypes.c_int), ("sweep_ms", ctypes.c_double),
        ("sweep_launches", ctypes.c_int), ("sweep_passes", ctypes.c_int),
    ]


_lib = None

_P = ctypes.c_void_p
_D = ctypes.POINTER(ctypes.c_double)
_SIGS = {
    "ath_open": [ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(_P)],
    "ath_close": [_P],
    "ath_get_info": [_P, ctypes.POINTER(Info)],
    "ath_block_dims": [_P, ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                       ctypes.POINTER(ctypes.c_int)],
    "ath_set_implicit_precision": [_P, ctypes.c_int],
    "ath_set_matrix_residual": [_P, ctypes.c_int],
    "ath_set_sweep_ordering": [_P, ctypes.c_int],
    "ath_get_sweep_ordering": [_P, ctypes.POINTER(ctypes.c_int)],
    "ath_autotune": [_P, ctypes.c_int, ctypes.POINTER(ctypes.c_int), _D],
    "ath_get_state": [_P, ctypes.c_int, _D],
    "ath_set_state": [_P, ctypes.c_int, _D],
    "ath_begin_step": [_P, ctypes.c_int],
    "ath_iterate": [_P, ctypes.c_int, ctypes.c_int, ctypes.POINTER(Norms)],
    "ath_run": [_P, ctypes.c_int, ctypes.c_int, ctypes.POINTER(Norms)],
    "ath_assign_ghost_cells": [_P],
    "ath_calc_residual": [_P],
    "ath_relax": [_P, _D],
    "ath_update": [_P, ctypes.c_int, ctypes.POINTER(Norms)],
    "ath_get_residual": [_P, ctypes.c_int, _D],
    "ath_get_dt": [_P, ctypes.c_int, _D],
    "ath_get_spectral_radius": [_P, ctypes.c_int, _D],
    "ath_get_diagonal": [_P, ctypes.c_int, _D],
    "ath_get_update": [_P, ctypes.c_int, _D],
    "ath_get_block_diagonal": [_P, ctypes.c_int, ctypes.c_int, _D],
    "ath_get_wall_distance": [_P, ctypes.c_int, _D],
    "ath_enable_timing": [_P, ctypes.c_int],
    "ath_get_timing": [_P, ctypes.POINTER(Timing)],
    "ath_inspect_info": [ctypes.c_char_p, ctypes.POINTER(Info)],
    "ath_inspect_field": [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, _D, ctypes.c_long,
                          ctypes.POINTER(ctypes.c_long)],
    "ath_eval_face_flux": [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double, _D, _D, _D,
                           ctypes.c_int, _D],
    "ath_eval_sst_cells": [ctypes.c_int, _D, ctypes.c_int, _D],
    "ath_write_restart": [_P, ctypes.c_char_p, ctypes.c_int],
    "ath_read_restart": [_P, ctypes.c_char_p, ctypes.POINTER(ctypes.c_int)],
    "ath_write_output": [_P, ctypes.c_char_p, ctypes.c_int],
    "ath_open_logs": [_P, ctypes.c_char_p, ctypes.c_int],
    "ath_close_logs": [_P],
    "ath_server_open": [ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.POINTER(_P)],
    "ath_server_close": [_P],
    "ath_server_memory": [_P, _D, _D],
    "ath_estimate_case_bytes": [ctypes.c_char_p, ctypes.c_int, _D],
    "ath_admit": [_P, ctypes.c_char_p, ctypes.c_double, ctypes.c_int, ctypes.POINTER(ctypes.c_int)],
    "ath_step_batch": [_P, ctypes.c_int, ctypes.POINTER(ctypes.c_int)],
    "ath_get_slot_info": [_P, ctypes.c_int, ctypes.POINTER(SlotInfo)],
    "ath_slot_block_dims": [_P, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                            ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)],
    "ath_get_slot_state": [_P, ctypes.c_int, ctypes.c_int, _D],
    "ath_set_slot_state": [_P, ctypes.c_int, ctypes.c_int, _D],
    "ath_retire": [_P, ctypes.c_int],
}


def load_library(path=LIB_PATH):
    """Load the native library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise AitherError(-6, f"native library {path} is missing; run `python -m aither_amd.build`")
    lib = ctypes.CDLL(path)
    for name, args in _SIGS.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = None if name in ("ath_close", "ath_server_close") else ctypes.c_int
    lib.ath_last_error.restype = ctypes.c_char_p
    lib.ath_last_error.argtypes = []
    _lib = lib
    return lib


def _check(rc):
    if rc != ATH_OK:
        raise AitherError(rc, _lib.ath_last_error().decode(errors="replace"))


def _ptr(a):
    return a.ctypes.data_as(_D)


def inspect_info(inp_path):
    lib = load_library()
    info = Info()
    _check(lib.ath_inspect_info(os.fsencode(inp_path), ctypes.byref(info)))
    return info


def inspect_field(inp_path, block, which):
    """Host geometry / initial state as built by the loader (no device)."""
    lib = load_library()
    need = ctypes.c_long(0)
    _check(lib.ath_inspect_field(os.fsencode(inp_path), block, which, None, 0, ctypes.byref(need)))
    out = np.empty(need.value, dtype=np.float64)
    _check(lib.ath_inspect_field(os.fsencode(inp_path), block, which, _ptr(out), out.size,
                                 ctypes.byref(need)))
    return out


def eval_face_flux(flux, gas_r, gas_n, WL, WR, n, device=0):
    """Device face flux (0 Roe, 1 AUSMPW+) for arrays of states and normals."""
    lib = load_library()
    WL = np.ascontiguousarray(WL, dtype=np.float64)
    WR = np.ascontiguousarray(WR, dtype=np.float64)
    n = np.ascontiguousarray(n, dtype=np.float64)
    count = WL.shape[0]
    F = np.empty((count, 5), dtype=np.float64)
    _check(lib.ath_eval_face_flux(device, flux, gas_r, gas_n, _ptr(WL), _ptr(WR), _ptr(n), count, _ptr(F)))
    return F


def eval_sst_cells(cells, device=0):
    """Device SST k-omega 2003 cell model: ``cells`` is (n, 27) (primitive
    state with k and omega, velocity gradient, k and omega gradients, laminar
    viscosity, wall distance, volume, scaling, 1/scaling); returns (n, 7):
    F1, F2, eddy viscosity, k and omega sources, source Jacobian (0,0), (1,1)."""
    lib = load_library()
    cells = np.ascontiguousarray(cells, dtype=np.float64)
    if cells.ndim != 2 or cells.shape[1] != 27:
        raise ValueError("cells must be (n, 27)")
    out = np.empty((cells.shape[0], 7), dtype=np.float64)
    _check(lib.ath_eval_sst_cells(device, _ptr(cells), cells.shape[0], _ptr(out)))
    return out


class Solver:
    """One case resident on one GPU (an ``ath_ctx``)."""

    FP64 = 0
    FP32 = 1

    def __init__(self, inp_path, device=0):
        self._lib = load_library()
        h = _P()
        _check(self._lib.ath_open(os.fsencode(inp_path), device, ctypes.byref(h)))
        self._h = h
        self.info = Info()
        _check(self._lib.ath_get_info(self._h, ctypes.byref(self.info)))
        self.dims = []
        for b in range(self.info.num_blocks):
            ni, nj, nk = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
            _check(self._lib.ath_block_dims(self._h, b, ctypes.byref(ni), ctypes.byref(nj), ctypes.byref(nk)))
            self.dims.append((ni.value, nj.value, nk.value))

    def close(self):
        if getattr(self, "_h", None):
            self._lib.ath_close(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def ncells(self, b):
        ni, nj, nk = self.dims[b]
        return ni * nj * nk

    # --- configuration
    def set_implicit_precision(self, p):
        _check(self._lib.ath_set_implicit_precision(self._h, p))

    TILED = 0
    HYPERPLANE = 1
    PENCIL = 2
    ROWS = 3

    def set_sweep_ordering(self, ordering):
        _check(self._lib.ath_set_sweep_ordering(self._h, ordering))

    def sweep_ordering(self):
        """the order the sweeps actually run in (block solvers and the
        approximateRoe Jacobian map some requests to another order)"""
        v = ctypes.c_int()
        _check(self._lib.ath_get_sweep_ordering(self._h, ctypes.byref(v)))
        return v.value

    def autotune(self, iterations=2):
        """time every sweep order the case admits and keep the fastest;
        returns (order, mean ms per sweep pass)"""
        order, ms = ctypes.c_int(), ctypes.c_double()
        _check(self._lib.ath_autotune(self._h, int(iterations), ctypes.byref(order), ctypes.byref(ms)))
        return order.value, ms.value

    def set_matrix_residual(self, on):
        _check(self._lib.ath_set_matrix_residual(self._h, 1 if on else 0))

    def enable_timing(self, on=True):
        _check(self._lib.ath_enable_timing(self._h, 1 if on else 0))

    def timing(self):
        t = Timing()
        _check(self._lib.ath_get_timing(self._h, ctypes.byref(t)))
        return t

    # --- state
    def get_state(self, b=0):
        out = np.empty(self.ncells(b) * 5, dtype=np.float64)
        _check(self._lib.ath_get_state(self._h, b, _ptr(out)))
        return out.reshape(-1, 5)

    def set_state(self, W, b=0):
        W = np.ascontiguousarray(W, dtype=np.float64).reshape(-1)
        _check(self._lib.ath_set_state(self._h, b, _ptr(W)))

    # --- files
    def write_restart(self, path, iteration):
        """the reference's binary restart file (WriteRestart)"""
        _check(self._lib.ath_write_restart(self._h, os.fsencode(path), int(iteration)))

    def read_restart(self, path):
        """state and residual normalisation from a restart file; returns the
        file's iteration"""
        it = ctypes.c_int()
        _check(self._lib.ath_read_restart(self._h, os.fsencode(path), ctypes.byref(it)))
        return it.value

    def write_output(self, directory, iteration):
        """cell-centre grid, function file and meta file (WriteOutput)"""
        _check(self._lib.ath_write_output(self._h, os.fsencode(directory), int(iteration)))

    def open_logs(self, directory, append=False):
        """<simName>.resid and <simName>.tme, a line per iteration"""
        _check(self._lib.ath_open_logs(self._h, os.fsencode(directory), 1 if append else 0))

    def close_logs(self):
        _check(self._lib.ath_close_logs(self._h))

    # --- hot path
    def begin_step(self, nn):
        _check(self._lib.ath_begin_step(self._h, nn))

    def iterate(self, nn, mm=0):
        n = Norms()
        _check(self._lib.ath_iterate(self._h, nn, mm, ctypes.byref(n)))
        return n

    def run(self, nn0, nsteps):
        n = Norms()
        _check(self._lib.ath_run(self._h, nn0, nsteps, ctypes.byref(n)))
        return n

    # --- granular
    def assign_ghost_cells(self):
        _check(self._lib.ath_assign_ghost_cells(self._h))

    def calc_residual(self):
        _check(self._lib.ath_calc_residual(self._h))

    def relax(self):
        v = ctypes.c_double()
        _check(self._lib.ath_relax(self._h, ctypes.byref(v)))
        return v.value

    def update(self, mm=0):
        n = Norms()
        _check(self._lib.ath_update(self._h, mm, ctypes.byref(n)))
        return n

    def _field(self, fn, b, width):
        out = np.empty(self.ncells(b) * width, dtype=np.float64)
        _check(fn(self._h, b, _ptr(out)))
        return out.reshape(-1, width) if width > 1 else out

    def residual(self, b=0):
        return self._field(self._lib.ath_get_residual, b, 5)

    def dt(self, b=0):
        return self._field(self._lib.ath_get_dt, b, 1)

    def spectral_radius(self, b=0):
        return self._field(self._lib.ath_get_spectral_radius, b, 1)

    def diagonal(self, b=0):
        return self._field(self._lib.ath_get_diagonal, b, 1)

    def wall_distance(self, b=0):
        """Distance from each cell centre to the nearest viscous-wall face
        centre (procBlock::CalcWallDistance), i fastest."""
        return self._field(self._lib.ath_get_wall_distance, b, 1)

    def update_vector(self, b=0):
        return self._field(self._lib.ath_get_update, b, 5)

    def block_diagonal(self, b=0, inverse=False):
        """Block main diagonal after AddDiagonalTerms (or its inverse), one
        5x5 row-major matrix per cell (blusgs / bdplur cases only)."""
        out = np.empty(self.ncells(b) * 25, dtype=np.float64)
        _check(self._lib.ath_get_block_diagonal(self._h, b, int(inverse), _ptr(out)))
        return out.reshape(-1, 5, 5)


class Server:
    """Many steady cases resident on one GPU, advanced together (an
    ``ath_server``): ``admit`` uploads a case if memory and a slot allow
    (else returns None), ``step`` runs every running case for some
    iterations in shared launches, ``info``/``state`` read a case and
    ``retire`` frees it."""

    RUNNING, CONVERGED, FAILED, LIMIT = 0, 1, 2, 3

    def __init__(self, device=0, hbm_budget=0.0, max_slots=64, precision=0):
        self._lib = load_library()
        h = _P()
        _check(self._lib.ath_server_open(device, float(hbm_budget), int(max_slots), int(precision),
                                         ctypes.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            self._lib.ath_server_close(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def memory(self):
        used, free = ctypes.c_double(), ctypes.c_double()
        _check(self._lib.ath_server_memory(self._h, ctypes.byref(used), ctypes.byref(free)))
        return used.value, free.value

    @staticmethod
    def estimate(inp_path, precision=0):
        """device bytes admission reserves for the case"""
        lib = load_library()
        v = ctypes.c_double()
        _check(lib.ath_estimate_case_bytes(os.fsencode(inp_path), int(precision), ctypes.byref(v)))
        return v.value

    def admit(self, inp_path, residual_drop=1e-8, max_iterations=10000):
        """slot id, or None when the case does not fit now (ATH_ERR_NOMEM)"""
        slot = ctypes.c_int(-1)
        rc = self._lib.ath_admit(self._h, os.fsencode(inp_path), float(residual_drop), int(max_iterations),
                                 ctypes.byref(slot))
        if rc == -4:
            return None
        _check(rc)
        return slot.value

    def step(self, iterations):
        running = ctypes.c_int(0)
        _check(self._lib.ath_step_batch(self._h, int(iterations), ctypes.byref(running)))
        return running.value

    def info(self, slot):
        o = SlotInfo()
        _check(self._lib.ath_get_slot_info(self._h, slot, ctypes.byref(o)))
        return o

    def dims(self, slot, block=0):
        ni, nj, nk = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        _check(self._lib.ath_slot_block_dims(self._h, slot, block, ctypes.byref(ni), ctypes.byref(nj),
                                             ctypes.byref(nk)))
        return ni.value, nj.value, nk.value

    def state(self, slot, block=0):
        ni, nj, nk = self.dims(slot, block)
        out = np.empty(ni * nj * nk * 5, dtype=np.float64)
        _check(self._lib.ath_get_slot_state(self._h, slot, block, _ptr(out)))
        return out.reshape(-1, 5)

    def set_state(self, slot, W, block=0):
        W = np.ascontiguousarray(W, dtype=np.float64).reshape(-1)
        _check(self._lib.ath_set_slot_state(self._h, slot, block, _ptr(W)))

    def retire(self, slot):
        _check(self._lib.ath_retire(self._h, slot))
