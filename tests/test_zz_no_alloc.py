"""ath_eval_face_flux_static (no device allocation) against
ath_eval_face_flux, through the C ABI: the same fluxes bit for bit on two
calls in a row, the second with fewer faces, and a refusal beyond the
array's capacity."""
import ctypes

import numpy as np
import pytest

import aither_amd

pytestmark = pytest.mark.gpu

_D = ctypes.POINTER(ctypes.c_double)


def states(rng, count):
    W = np.empty((count, 5))
    W[:, 0] = rng.uniform(0.5, 2.0, count)
    W[:, 1:4] = rng.uniform(-0.8, 0.8, (count, 3))
    W[:, 4] = rng.uniform(0.3, 1.5, count)
    return W


def call(fn, flux, WL, WR, n):
    count = WL.shape[0]
    F = np.full((count, 5), np.nan)
    p = lambda a: a.ctypes.data_as(_D)
    rc = fn(0, flux, 1.0 / 1.4, 2.5, p(WL), p(WR), p(n), count, p(F))
    return rc, F


def test_static_equals_allocating_entry_point():
    lib = aither_amd.load_library()
    fn = lib.ath_eval_face_flux_static
    fn.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double, _D, _D, _D, ctypes.c_int, _D]
    fn.restype = ctypes.c_int
    rng = np.random.default_rng(5)
    for flux in (0, 1):
        for count in (5000, 1237):
            WL, WR = states(rng, count), states(rng, count)
            n = rng.normal(size=(count, 3))
            n /= np.linalg.norm(n, axis=1, keepdims=True)
            print("flux", flux, "faces", count, flush=True)
            rc, Fs = call(fn, flux, WL, WR, n)
            assert rc == 0, aither_amd._lib.ath_last_error()
            Fa = aither_amd.eval_face_flux(flux, 1.0 / 1.4, 2.5, WL, WR, n)
            assert np.array_equal(Fs, Fa)
    # beyond the array's capacity: refused before any device work
    big = 65537
    z = np.ones((big, 5))
    rc, _ = call(fn, 0, z, z, np.ones((big, 3)))
    assert rc == -1
    assert b"at most 65536" in aither_amd._lib.ath_last_error()
