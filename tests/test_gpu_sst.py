"""The SST k-omega 2003 cell model (turbKWSst, src/turbulence.cpp:573-760) on
the GPU against the FP64 oracle, through the C ABI.  Both evaluate the same
operations in the same order; they differ only in tanh and pow, whose device
and host libraries may round differently in the last place, so the bar is a
relative 1e-13 per output, and exact where the model saturates (F1 = F2 = 1
at the wall) or has no transcendental input (the eddy viscosity without
strain, the source Jacobian of the inner coefficient set)."""
import numpy as np
import pytest

import aither_amd
import pyoracle

pytestmark = pytest.mark.gpu


def _cells(n, seed):
    rng = np.random.default_rng(seed)
    c = np.zeros((n, 27))
    c[:, 0] = rng.uniform(0.2, 3.0, n)                 # rho
    c[:, 1:4] = rng.normal(size=(n, 3)) * 50           # velocity
    c[:, 4] = rng.uniform(0.5, 2.0, n) * 1e5           # p
    c[:, 5] = 10 ** rng.uniform(-6, 1, n)              # k
    c[:, 6] = 10 ** rng.uniform(-1, 6, n)              # omega
    scale = 10 ** rng.uniform(-2, 5, n)[:, None]
    c[:, 7:16] = rng.normal(size=(n, 9)) * scale       # velocity gradient
    c[::7, 7:16] = 0.0                                 # no strain
    c[1::7, 7:16] = np.eye(3).ravel() * rng.normal(size=(len(c[1::7]), 1)) * 100  # dilatation
    c[:, 16:19] = rng.normal(size=(n, 3)) * 10 ** rng.uniform(-3, 1, n)[:, None]
    c[:, 19:22] = rng.normal(size=(n, 3)) * 10 ** rng.uniform(0, 5, n)[:, None]
    c[:, 22] = 10 ** rng.uniform(-7, -3, n)            # mu
    c[:, 23] = 10 ** rng.uniform(-8, 2, n)             # wall distance
    c[:, 24] = 10 ** rng.uniform(-9, -3, n)            # volume
    sc = 10 ** rng.uniform(-1, 1, n)
    c[:, 25], c[:, 26] = sc, 1.0 / sc
    return c


@pytest.mark.parametrize("n,seed", [(1, 0), (333, 1), (20000, 2)])
def test_sst_cells_match_oracle(n, seed):
    c = _cells(n, seed)
    got = aither_amd.eval_sst_cells(c)
    ref = pyoracle.sst_cells(c)
    assert np.all(np.isfinite(got))
    den = np.maximum(np.abs(ref), 1e-300)
    worst = float(np.max(np.abs(got - ref) / den))
    assert worst < 1e-13, worst
    near = c[:, 23] < 1e-7
    # the outer blend is exact where F1 saturates; the Jacobian (0,0) has no
    # transcendental input at all
    assert np.array_equal(got[:, 5], ref[:, 5])
    sat = ref[:, 0] == 1.0
    assert np.array_equal(got[sat, 6], ref[sat, 6])
    assert near.sum() == 0 or np.all(got[near & sat, 0] == 1.0)


def test_sst_cells_two_million_cells_sampled():
    # a 2,097,152-cell launch (8,192 workgroups) checked on a seeded sample
    c = _cells(1 << 21, 5)
    got = aither_amd.eval_sst_cells(c)
    assert np.all(np.isfinite(got))
    rows = np.random.default_rng(6).choice(len(c), 4000, replace=False)
    ref = pyoracle.sst_cells(c[rows])
    den = np.maximum(np.abs(ref), 1e-300)
    assert float(np.max(np.abs(got[rows] - ref) / den)) < 1e-13


def test_sst_cells_bad_count_is_refused():
    with pytest.raises(aither_amd.AitherError):
        aither_amd.eval_sst_cells(np.zeros((0, 27)))
