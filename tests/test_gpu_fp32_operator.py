"""The FP32 implicit operator (D^-1, off-diagonal products and du in FP32;
residual, spectral radius and time step in FP64) against the FP64 oracle at
convergence: steady bump channels run to a 1e-11 residual drop on the GPU
with each operator precision and in the oracle.  The converged state of
either GPU run is within 1e-10 of the oracle's; the FP32 operator's extra
iterations are printed.  Cases: scalar LU-SGS (inviscid), BLU-SGS with 5x5
blocks (inviscid, and laminar with no-slip walls)."""
import sys

import numpy as np
import pytest

from aither_amd import Solver, cases
from aither_amd.serve import drop_of
import pyoracle

pytestmark = pytest.mark.gpu

DROP = 1e-11


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def gpu_converge(path, precision, chunk=10, cap=20000):
    s = Solver(path)
    s.set_implicit_precision(precision)
    it = 0
    while it < cap:
        n = s.run(it, chunk)
        it += chunk
        if drop_of(list(n.l2[:5]), list(n.normalized[:5])) <= DROP:
            break
    W = np.concatenate([s.get_state(b).reshape(-1) for b in range(s.info.num_blocks)])
    s.close()
    return W, it


def oracle_converge(path, chunk=10, cap=20000):
    o = pyoracle.OracleCase(path, parallel=True)
    it = 0
    while it < cap:
        o.begin_step(it)
        r = o.iterate(it, 0)
        it += 1
        if it % chunk == 0 and drop_of(list(r["l2"]), list(r["normalized"])) <= DROP:
            break
    return np.concatenate([o.state(b).reshape(-1) for b in range(o.nb)]), it


@pytest.mark.parametrize("kind,n,viscous,solver", [("scalar", 10, False, "lusgs"),
                                                   ("block", 10, False, "blusgs"),
                                                   ("block-viscous", 8, True, "blusgs")])
def test_fp32_operator_converged_error(tmp_path, kind, n, viscous, solver):
    path, meta = cases.bump_case(str(tmp_path), seed=5, n=n, iterations=1, cfl=20.0, viscous=viscous,
                                 limiter="minmod")
    text = open(path).read().replace("matrixSolver: lusgs", f"matrixSolver: {solver}")
    open(path, "w").write(text)
    say(f"{kind}: {n}^3 bump, Mach {meta['mach']:.3f}, h/l {meta['h_over_l']:.3f}, to a {DROP:g} drop")
    Wo, io = oracle_converge(path)
    W64, i64 = gpu_converge(path, Solver.FP64)
    W32, i32 = gpu_converge(path, Solver.FP32)
    e64 = np.linalg.norm(W64 - Wo) / np.linalg.norm(Wo)
    e32 = np.linalg.norm(W32 - Wo) / np.linalg.norm(Wo)
    say(f"  iterations: oracle {io}, GPU FP64 operator {i64}, FP32 operator {i32} ({i32 - i64:+d}); "
        f"converged state error vs oracle: FP64 {e64:.2e}, FP32 {e32:.2e}")
    assert e64 <= 1e-10 and e32 <= 1e-10
