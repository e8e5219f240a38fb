"""approximateRoe off-diagonals (RoeOffDiagonal, src/fluxJacobian.cpp:240-290)
on the GPU against the FP64 oracle, through the C ABI: per-step residuals
within 1e-12 and matrix residuals within 1e-10 (the bars of the Rusanov
operators), with scalar LU-SGS (hyperplane order) and DP-LUR, on one block
and across a rotated interblock connection.  The FP32 operator is refused.
"""
import numpy as np
import pytest

import aither_amd
from aither_amd import Solver, cases
import pyoracle

pytestmark = pytest.mark.gpu


def rel(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def _roe(path, solver="lusgs", sweeps=1):
    text = open(path).read()
    text = text.replace("matrixSolver: lusgs", f"matrixSolver: {solver}\ninviscidFluxJacobian: approximateRoe")
    text = text.replace("matrixSweeps: 1", f"matrixSweeps: {sweeps}")
    out = path.replace(".inp", f"_roe{solver}{sweeps}.inp")
    open(out, "w").write(text)
    return out


@pytest.mark.parametrize("solver,sweeps", [("lusgs", 1), ("lusgs", 2), ("dplur", 3)])
def test_roe_offdiagonal_step_parity(tmp_path, solver, sweeps):
    p = cases.bump_case(str(tmp_path), seed=11, n=10, iterations=6, cfl=5.0)[0]
    path = _roe(p, solver, sweeps)
    print("approximateRoe", solver, sweeps, flush=True)
    s = Solver(path)
    s.set_matrix_residual(1)
    o = pyoracle.OracleCase(path)
    o.set_matrix_residual(1)
    for nn in range(4):
        s.begin_step(nn)
        o.begin_step(nn)
        ns = s.iterate(nn, 0)
        no = o.iterate(nn, 0)
        assert rel(s.residual(), o.residual()) <= 1e-12, nn
        assert abs(ns.matrix_resid - no["matrix_resid"]) <= 1e-10 * no["matrix_resid"], nn
    assert rel(s.get_state(), o.state()) <= 1e-12
    with pytest.raises(aither_amd.AitherError):
        s.set_implicit_precision(1)
    s.close()


def test_roe_offdiagonal_runs_in_hyperplane_order(tmp_path):
    """approximateRoe runs in the hyperplane order whatever order is set;
    the order actually used is read back from the library."""
    p = cases.bump_case(str(tmp_path), seed=11, n=10, iterations=3, cfl=5.0)[0]
    path = _roe(p)
    out = []
    for order in ("PENCIL", "TILED", "HYPERPLANE"):
        print("approximateRoe ordering", order, flush=True)
        s = Solver(path)
        s.set_sweep_ordering(getattr(Solver, order))
        assert s.sweep_ordering() == Solver.HYPERPLANE
        for nn in range(3):
            s.begin_step(nn)
            s.iterate(nn, 0)
        out.append(s.get_state())
        s.close()
    assert np.array_equal(out[0], out[2]) and np.array_equal(out[1], out[2])


def test_roe_offdiagonal_across_connections(tmp_path):
    import conn_cases as cc
    r0, r1 = cc.orientation_set()[211]
    path, _ = cc.write_pair(str(tmp_path), "cut", r0, r1, timeint="implicitEuler")
    text = open(path).read().replace("matrixSolver: lusgs", "matrixSolver: lusgs\ninviscidFluxJacobian: approximateRoe")
    open(path, "w").write(text)
    print("approximateRoe connections", flush=True)
    s = Solver(path)
    o = pyoracle.OracleCase(path)
    for nn in range(4):
        s.begin_step(nn)
        o.begin_step(nn)
        s.iterate(nn, 0)
        o.iterate(nn, 0)
        Rs = np.concatenate([s.residual(b) for b in range(o.nb)])
        Ro = np.concatenate([o.residual(b) for b in range(o.nb)])
        assert rel(Rs, Ro) <= 1e-12, (nn, rel(Rs, Ro))
    Ws = np.concatenate([s.get_state(b) for b in range(o.nb)])
    Wo = np.concatenate([o.state(b) for b in range(o.nb)])
    assert rel(Ws, Wo) <= 1e-12
    s.close()
