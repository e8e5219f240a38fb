"""ctypes wrapper of the oracle (test infrastructure only).

Imported only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
leg, always as the checker.  The product library never loads it.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
LIB_OMP = os.path.join(HERE, "liboracle_omp.so")  # same code, parallel loops (OpenMP)

_libs = {}
_V = ctypes.c_void_p
_D = ctypes.POINTER(ctypes.c_double)


class Gas(ctypes.Structure):
    _fields_ = [("R", ctypes.c_double), ("n", ctypes.c_double)]


def lib(parallel=False):
    """the serial oracle, or (parallel=True) its OpenMP build, which gives
    the same results bit for bit"""
    path = LIB_OMP if parallel else LIB
    if path not in _libs:
        if not os.path.exists(path):
            subprocess.run(["make", "-s", "-C", HERE, os.path.basename(path)], check=True)
        L = ctypes.CDLL(path)
        L.orc_num_threads.restype = ctypes.c_int
        L.orc_num_threads.argtypes = []
        L.orc_load.restype = _V
        L.orc_load.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int]
        L.orc_free.argtypes = [_V]
        for f in ("orc_num_blocks", "orc_num_eq", "orc_num_ghosts", "orc_iterations",
                  "orc_nonlinear_iterations"):
            getattr(L, f).argtypes = [_V]
            getattr(L, f).restype = ctypes.c_int
        L.orc_block_dims.argtypes = [_V, ctypes.c_int] + [ctypes.POINTER(ctypes.c_int)] * 3
        L.orc_refs.argtypes = [_V, _D]
        L.orc_gas_of.argtypes = [_V, ctypes.POINTER(Gas)]
        L.orc_set_matrix_residual.argtypes = [_V, ctypes.c_int]
        for f in ("orc_get_state", "orc_set_state", "orc_get_state_padded", "orc_get_residual",
                  "orc_get_dt", "orc_get_specrad", "orc_get_diag", "orc_get_update", "orc_set_update",
                  "orc_get_volume", "orc_get_center"):
            getattr(L, f).argtypes = [_V, ctypes.c_int, _D]
        for f in ("orc_get_face_area", "orc_get_cell_width"):
            getattr(L, f).argtypes = [_V, ctypes.c_int, ctypes.c_int, _D]
        L.orc_begin_step.argtypes = [_V, ctypes.c_int]
        L.orc_iterate.argtypes = [_V, ctypes.c_int, _D, _D, ctypes.POINTER(ctypes.c_int), _D]
        L.orc_normalize.argtypes = [_V, ctypes.c_int, ctypes.c_int, _D, _D]
        L.orc_cfl.argtypes = [_V]
        L.orc_cfl.restype = ctypes.c_double
        for f in ("orc_bc", "orc_residual", "orc_time_step", "orc_invert_diagonal", "orc_init_update",
                  "orc_reset_diagonal"):
            getattr(L, f).argtypes = [_V]
        L.orc_relax.argtypes = [_V]
        L.orc_relax.restype = ctypes.c_double
        L.orc_update.argtypes = [_V, ctypes.c_int, _D, _D, ctypes.POINTER(ctypes.c_int)]
        G = ctypes.POINTER(Gas)
        L.orc_gamma.argtypes = [G]
        L.orc_gamma.restype = ctypes.c_double
        L.orc_sos.argtypes = [G, _D]
        L.orc_sos.restype = ctypes.c_double
        L.orc_enthalpy.argtypes = [G, _D]
        L.orc_enthalpy.restype = ctypes.c_double
        L.orc_prim_to_cons.argtypes = [G, _D, _D]
        L.orc_cons_to_prim.argtypes = [G, _D, _D]
        for f in ("orc_roe_flux", "orc_ausm_flux"):
            getattr(L, f).argtypes = [G, _D, _D, _D, _D]
        L.orc_phys_flux.argtypes = [G, _D, _D, _D]
        L.orc_muscl.argtypes = [_D, _D, _D, ctypes.c_double, ctypes.c_int, ctypes.c_double,
                                ctypes.c_double, ctypes.c_double, _D]
        L.orc_lagrange_coeff.argtypes = [_D, ctypes.c_int, ctypes.c_int, ctypes.c_int, _D]
        L.orc_weno_beta.argtypes = [ctypes.c_int] + [ctypes.c_double] * 6
        L.orc_weno_beta.restype = ctypes.c_double
        L.orc_weno.argtypes = [_D] * 6 + [ctypes.c_int, _D]
        L.orc_update_aux.argtypes = [_V]
        L.orc_set_state_padded.argtypes = [_V, ctypes.c_int, _D]
        L.orc_face_grads.argtypes = [_V] + [ctypes.c_int] * 5 + [_D]
        L.orc_viscosity.argtypes = [_V, ctypes.c_double]
        L.orc_viscosity.restype = ctypes.c_double
        L.orc_conductivity.argtypes = [_V, ctypes.c_double]
        L.orc_conductivity.restype = ctypes.c_double
        L.orc_pyramid_volume.argtypes = [_D] * 5
        L.orc_pyramid_volume.restype = ctypes.c_double
        L.orc_hyperplane_order.argtypes = [ctypes.c_int] * 3 + [ctypes.POINTER(ctypes.c_int)]
        L.orc_inv_cell_specrad.argtypes = [G, _D, _D, _D]
        L.orc_inv_cell_specrad.restype = ctypes.c_double
        L.orc_scalar_offdiag.argtypes = [G, _D, _D, _D, ctypes.c_int, _D]
        L.orc_roe_offdiag.argtypes = [G, _D, _D, _D, _D, ctypes.c_int, _D]
        L.orc_rusanov_jacobian.argtypes = [G, _D, _D, ctypes.c_int, _D]
        L.orc_block_inverse.argtypes = [_D]
        L.orc_block_inverse.restype = ctypes.c_int
        for f in ("orc_get_block_diag", "orc_get_block_inverse"):
            getattr(L, f).argtypes = [_V, ctypes.c_int, _D]
        L.orc_viscous_flux.argtypes = [_V, _D, _D, _D, _D, ctypes.c_double, ctypes.c_int, _D]
        L.orc_tsl_jacobian.argtypes = [_V, _D, ctypes.c_double, _D, ctypes.c_double, ctypes.c_int, _D, _D]
        L.orc_get_cell_grad.argtypes = [_V, ctypes.c_int, _D]
        L.orc_is_block.argtypes = [_V]
        L.orc_is_block.restype = ctypes.c_int
        L.orc_sst_cell.argtypes = [_D, _D]
        _libs[path] = L
    return _libs[path]


def P(a):
    return a.ctypes.data_as(_D)


def arr(x):
    return np.ascontiguousarray(x, dtype=np.float64)


def sst_cells(cells):
    """the oracle's SST 2003 cell model over rows of 27 inputs -> rows of 7"""
    L = lib()
    cells = arr(np.atleast_2d(cells))
    out = np.empty((cells.shape[0], 7))
    for r in range(cells.shape[0]):
        L.orc_sst_cell(P(cells[r]), P(out[r]))
    return out


class OracleCase:
    """One case in the oracle, with the same call sequence as the reference."""

    def __init__(self, inp_path, parallel=False):
        L = lib(parallel)
        self.threads = L.orc_num_threads()
        err = ctypes.create_string_buffer(1024)
        h = L.orc_load(os.fsencode(inp_path), err, 1024)
        if not h:
            raise RuntimeError("oracle load failed: " + err.value.decode())
        self.h = _V(h)
        self.L = L
        self.nb = L.orc_num_blocks(self.h)
        self.dims = []
        for b in range(self.nb):
            ni, nj, nk = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
            L.orc_block_dims(self.h, b, ctypes.byref(ni), ctypes.byref(nj), ctypes.byref(nk))
            self.dims.append((ni.value, nj.value, nk.value))
        self.iterations = L.orc_iterations(self.h)
        self.nonlinear = L.orc_nonlinear_iterations(self.h)
        g = Gas()
        L.orc_gas_of(self.h, ctypes.byref(g))
        self.gas = g

    def __del__(self):
        try:
            self.L.orc_free(self.h)
        except Exception:
            pass

    def ncells(self, b=0):
        ni, nj, nk = self.dims[b]
        return ni * nj * nk

    def set_matrix_residual(self, on):
        self.L.orc_set_matrix_residual(self.h, 1 if on else 0)

    def begin_step(self, nn):
        self.L.orc_begin_step(self.h, nn)

    def iterate(self, nn, mm=0):
        l2 = np.zeros(5)
        linf = np.zeros(1)
        loc = (ctypes.c_int * 5)()
        mr = np.zeros(1)
        self.L.orc_iterate(self.h, mm, P(l2), P(linf), loc, P(mr))
        l2r = np.sqrt(l2)
        norm = np.zeros(5)
        self.L.orc_normalize(self.h, nn, mm, P(arr(l2r)), P(norm))
        return dict(l2_sq=l2, l2=l2r, normalized=norm, linf=linf[0], linf_loc=tuple(loc), matrix_resid=mr[0])

    def _get(self, fn, b, width):
        out = np.empty(self.ncells(b) * width)
        fn(self.h, b, P(out))
        return out.reshape(-1, width) if width > 1 else out

    def state(self, b=0):
        return self._get(self.L.orc_get_state, b, 5)

    def set_state_padded(self, W, b=0):
        """all cells, ghost layers included: (nk+2g, nj+2g, ni+2g, 5) flattened"""
        W = arr(W)
        self.L.orc_set_state_padded(self.h, b, P(W))

    def set_state(self, W, b=0):
        self.L.orc_set_state(self.h, b, P(arr(W).reshape(-1)))

    def residual(self, b=0):
        return self._get(self.L.orc_get_residual, b, 5)

    def dt(self, b=0):
        return self._get(self.L.orc_get_dt, b, 1)

    def spectral_radius(self, b=0):
        return self._get(self.L.orc_get_specrad, b, 1)

    def diagonal(self, b=0):
        return self._get(self.L.orc_get_diag, b, 1)

    def block_diagonal(self, b=0):
        return self._get(self.L.orc_get_block_diag, b, 25).reshape(-1, 5, 5)

    def block_inverse(self, b=0):
        return self._get(self.L.orc_get_block_inverse, b, 25).reshape(-1, 5, 5)

    def cell_gradients(self, b=0):
        ni, nj, nk = self.dims[b]
        g = self.L.orc_num_ghosts(self.h)
        out = np.empty((ni + 2 * g) * (nj + 2 * g) * (nk + 2 * g) * 9)
        self.L.orc_get_cell_grad(self.h, b, P(out))
        return out.reshape(nk + 2 * g, nj + 2 * g, ni + 2 * g, 3, 3)

    def is_block(self):
        return bool(self.L.orc_is_block(self.h))

    def update_vector(self, b=0):
        return self._get(self.L.orc_get_update, b, 5)

    def volume(self, b=0):
        return self._get(self.L.orc_get_volume, b, 1)

    def centre(self, b=0):
        return self._get(self.L.orc_get_center, b, 3)

    def cell_width(self, d, b=0):
        ni, nj, nk = self.dims[b]
        g = self.L.orc_num_ghosts(self.h)
        out = np.empty((ni + 2 * g) * (nj + 2 * g) * (nk + 2 * g))
        self.L.orc_get_cell_width(self.h, b, d, P(out))
        return out

    def face_area(self, d, b=0):
        ni, nj, nk = self.dims[b]
        g = self.L.orc_num_ghosts(self.h)
        n = (ni + 2 * g + (d == 0)) * (nj + 2 * g + (d == 1)) * (nk + 2 * g + (d == 2))
        out = np.empty(4 * n)
        self.L.orc_get_face_area(self.h, b, d, P(out))
        return out

    # granular
    def bc(self):
        self.L.orc_bc(self.h)

    def calc_residual(self):
        self.L.orc_residual(self.h)

    def time_step(self):
        self.L.orc_time_step(self.h)

    def invert_diagonal(self):
        self.L.orc_invert_diagonal(self.h)

    def init_update(self):
        self.L.orc_init_update(self.h)

    def relax(self):
        return self.L.orc_relax(self.h)

    def update(self, mm=0):
        l2 = np.zeros(5)
        linf = np.zeros(1)
        loc = (ctypes.c_int * 5)()
        self.L.orc_update(self.h, mm, P(l2), P(linf), loc)
        return dict(l2_sq=l2, linf=linf[0], linf_loc=tuple(loc))

    def reset_diagonal(self):
        self.L.orc_reset_diagonal(self.h)


def run_reference_loop(inp_path, iterations=None):
    """Run the oracle's main loop; returns (case, list of per-iteration norms)."""
    c = OracleCase(inp_path)
    its = iterations if iterations is not None else c.iterations
    hist = []
    for nn in range(its):
        c.begin_step(nn)
        for mm in range(c.nonlinear):
            hist.append(c.iterate(nn, mm))
    return c, hist
