"""Block-implicit solvers (matrixSolver blusgs / bdplur) on the GPU against
the FP64 oracle, through the C ABI.

Bars: the block main diagonal and its Gauss-Jordan inverse equal the
oracle's bit for bit where the reconstruction calls no std::pow (constant,
MUSCL), else within 1e-12 relative; per-step residuals within 1e-12 (the
bar of the scalar path) and the state after the run within 1e-12.
"""
import os
import sys

import numpy as np
import pytest

import aither_amd
from aither_amd import Solver, cases
import pyoracle

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def _variant(path, tag, **repl):
    text = open(path).read()
    for a, b in repl.items():
        assert a in text, a
        text = text.replace(a, b)
    out = path.replace(".inp", f"_{tag}.inp")
    with open(out, "w") as f:
        f.write(text)
    return out


def _bump(tmp, recon="thirdOrder", solver="blusgs", sweeps=1, timeint="implicitEuler", cfl=20.0, its=6):
    p = cases.bump_case(str(tmp), seed=11, n=10, iterations=its, cfl=cfl, recon=recon)[0]
    return _variant(p, f"{solver}{sweeps}{timeint}",
                    **{"matrixSolver: lusgs": f"matrixSolver: {solver}",
                       "matrixSweeps: 1": f"matrixSweeps: {sweeps}",
                       "timeIntegration: implicitEuler": f"timeIntegration: {timeint}"})


@pytest.mark.parametrize("recon,flux", [("constant", "roe"), ("thirdOrder", "roe"), ("thirdOrder", "ausm"),
                                        ("weno", "roe")])
def test_block_diagonal_and_inverse_match_oracle(tmp_path, recon, flux):
    path = _bump(tmp_path, recon=recon)
    if flux == "ausm":
        path = _variant(path, "ausm", **{"limiter: vanAlbada": "limiter: vanAlbada\ninviscidFlux: ausm"})
    say("block diagonal", recon, flux)
    s = Solver(path)
    assert s.info.block_matrix == 1
    o = pyoracle.OracleCase(path)
    assert o.is_block()
    s.begin_step(0)
    o.begin_step(0)
    s.assign_ghost_cells()
    s.calc_residual()
    o.bc()
    o.calc_residual()
    o.time_step()
    o.invert_diagonal()
    A, Ai = s.block_diagonal(), s.block_diagonal(inverse=True)
    Ao, Aio = o.block_diagonal(), o.block_inverse()
    assert rel(s.residual(), o.residual()) <= 1e-12
    if recon == "weno":
        assert rel(A, Ao) <= 1e-12 and rel(Ai, Aio) <= 1e-12
    else:
        assert np.array_equal(A, Ao)
        assert np.array_equal(Ai, Aio)
    with pytest.raises(aither_amd.AitherError):
        s.diagonal()
    s.close()


@pytest.mark.parametrize("solver,sweeps,timeint", [("blusgs", 1, "implicitEuler"), ("blusgs", 2, "implicitEuler"),
                                                   ("bdplur", 3, "implicitEuler"), ("blusgs", 1, "bdf2"),
                                                   ("bdplur", 2, "crankNicholson")])
def test_block_solvers_step_parity(tmp_path, solver, sweeps, timeint):
    path = _bump(tmp_path, solver=solver, sweeps=sweeps, timeint=timeint)
    if timeint != "implicitEuler":
        path = _variant(path, "nl", **{"matrixRelaxation: 1.0": "matrixRelaxation: 1.0\nnonlinearIterations: 2"})
    say("block solver", solver, sweeps, timeint)
    s = Solver(path)
    s.set_matrix_residual(1)
    o = pyoracle.OracleCase(path)
    o.set_matrix_residual(1)
    for nn in range(5):
        s.begin_step(nn)
        o.begin_step(nn)
        for mm in range(o.nonlinear):
            ns = s.iterate(nn, mm)
            no = o.iterate(nn, mm)
            assert rel(s.residual(), o.residual()) <= 1e-12, (nn, mm)
            assert abs(ns.matrix_resid - no["matrix_resid"]) <= 1e-10 * no["matrix_resid"], (nn, mm)
    assert rel(s.get_state(), o.state()) <= 1e-12
    s.close()


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("sweeps", [1, 2])
def test_block_sweep_orderings_agree(tmp_path, precision, sweeps):
    """every sweep order of the block solvers gives the same bits, with the
    FP64 and with the FP32 operator, in first and in full sweeps (the pencil
    order runs the row-parallel sweep with the FP64 operator, the tiled one
    with the FP32 operator)"""
    path = _bump(tmp_path, sweeps=sweeps)
    out = []
    for order in ("ROWS", "PENCIL", "TILED", "HYPERPLANE"):
        say("block ordering", order, "precision", precision, "sweeps", sweeps)
        s = Solver(path)
        s.set_implicit_precision(precision)
        s.set_sweep_ordering(getattr(Solver, order))
        pencil = Solver.TILED if precision == 1 else Solver.ROWS
        assert s.sweep_ordering() == (pencil if order == "PENCIL" else getattr(Solver, order))
        for nn in range(3):
            s.begin_step(nn)
            s.iterate(nn, 0)
        out.append(s.get_state())
        s.close()
    for q in range(3):
        assert np.array_equal(out[q], out[3]), q


def test_block_graph_run_equals_iterate(tmp_path):
    path = _bump(tmp_path, solver="blusgs", sweeps=2)
    say("block graph replay")
    a = Solver(path)
    for nn in range(4):
        a.begin_step(nn)
        a.iterate(nn, 0)
    b = Solver(path)
    b.run(0, 4)
    assert np.array_equal(a.get_state(), b.get_state())
    a.close()
    b.close()


@pytest.mark.parametrize("solver,sweeps", [("blusgs", 1), ("bdplur", 2)])
def test_block_solvers_across_connections(tmp_path, solver, sweeps):
    import conn_cases as cc
    r0, r1 = cc.orientation_set()[211]
    path, _ = cc.write_pair(str(tmp_path), "cut", r0, r1, timeint="implicitEuler", solver=solver, sweeps=sweeps)
    say("block connections", solver, sweeps)
    s = Solver(path)
    o = pyoracle.OracleCase(path)
    for nn in range(4):
        s.begin_step(nn)
        o.begin_step(nn)
        ns = s.iterate(nn, 0)
        no = o.iterate(nn, 0)
        Rs = np.concatenate([s.residual(b) for b in range(o.nb)])
        Ro = np.concatenate([o.residual(b) for b in range(o.nb)])
        assert rel(Rs, Ro) <= 1e-12, (nn, rel(Rs, Ro))
    Ws = np.concatenate([s.get_state(b) for b in range(o.nb)])
    Wo = np.concatenate([o.state(b) for b in range(o.nb)])
    assert rel(Ws, Wo) <= 1e-12
    s.close()


def _common_state_parity(s, o, iterations, tol=1e-12):
    """per-step residual and matrix residual from a common state (the oracle
    restarts each step from the GPU state): Sutherland's T^1.5 is correctly
    rounded on the device and not always in glibc"""
    for nn in range(iterations):
        for b in range(o.nb):
            o.set_state(s.get_state(b), b)
        s.begin_step(nn)
        o.begin_step(nn)
        for mm in range(o.nonlinear):
            ns = s.iterate(nn, mm)
            no = o.iterate(nn, mm)
            Rs = np.concatenate([s.residual(b) for b in range(o.nb)])
            Ro = np.concatenate([o.residual(b) for b in range(o.nb)])
            assert rel(Rs, Ro) <= tol, (nn, mm, rel(Rs, Ro))
            if no["matrix_resid"] > 0:
                assert abs(ns.matrix_resid - no["matrix_resid"]) <= 1e-10 * no["matrix_resid"], (nn, mm)


@pytest.mark.parametrize("visc_recon,recon", [("central", "thirdOrder"), ("centralFourth", "thirdOrder"),
                                              ("central", "weno")])
def test_viscous_block_diagonal_matches_oracle(tmp_path, visc_recon, recon):
    """The wall cells' inviscid sums are formed before the viscous ghost
    cells, every other cell's after them (three ghost layers deep with
    WENO); both against the oracle."""
    p = cases.bump_case(str(tmp_path), seed=4, n=8, iterations=3, viscous=True, visc_recon=visc_recon,
                        recon=recon)[0]
    path = _variant(p, "blk", **{"matrixSolver: lusgs": "matrixSolver: blusgs"})
    say("viscous block diagonal", visc_recon, recon)
    s = Solver(path)
    o = pyoracle.OracleCase(path)
    # from a non-uniform state: at the uniform initial state the limiter
    # hides the wall ghost states from the inviscid face states
    for nn in range(2):
        s.begin_step(nn)
        s.iterate(nn, 0)
    o.set_state(s.get_state())
    s.begin_step(2)
    o.begin_step(2)
    s.assign_ghost_cells()
    s.calc_residual()
    o.bc()
    o.calc_residual()
    o.time_step()
    o.invert_diagonal()
    assert rel(s.residual(), o.residual()) <= 1e-12
    assert rel(s.block_diagonal(), o.block_diagonal()) <= 1e-12
    assert rel(s.block_diagonal(inverse=True), o.block_inverse()) <= 1e-12
    s.close()


@pytest.mark.parametrize("solver,sweeps,timeint", [("blusgs", 1, "implicitEuler"), ("bdplur", 2, "implicitEuler"),
                                                   ("blusgs", 2, "bdf2")])
def test_viscous_block_solvers_step_parity(tmp_path, solver, sweeps, timeint):
    p = cases.bump_case(str(tmp_path), seed=4, n=8, iterations=4, viscous=True)[0]
    repl = {"matrixSolver: lusgs": f"matrixSolver: {solver}", "matrixSweeps: 1": f"matrixSweeps: {sweeps}",
            "timeIntegration: implicitEuler": f"timeIntegration: {timeint}"}
    if timeint == "bdf2":
        repl["matrixRelaxation: 1.0"] = "matrixRelaxation: 1.0\nnonlinearIterations: 2"
    path = _variant(p, f"{solver}{sweeps}{timeint}", **repl)
    say("viscous block solver", solver, sweeps, timeint)
    s = Solver(path)
    o = pyoracle.OracleCase(path)
    _common_state_parity(s, o, 4)
    s.close()


def test_viscous_block_across_connections(tmp_path):
    """Viscous block off-diagonals across a rotated connection read the
    partner's cell velocity gradients through the ghost cells."""
    import conn_cases as cc
    r0, r1 = cc.orientation_set()[350]
    path, _ = cc.write_pair(str(tmp_path), "cut", r0, r1, timeint="implicitEuler", eq="navierStokes", scale=1e-3,
                            solver="blusgs", sweeps=2)
    say("viscous block connections")
    s = Solver(path)
    o = pyoracle.OracleCase(path)
    _common_state_parity(s, o, 4)
    s.close()


@pytest.mark.parametrize("refine,solver", [(0.125, "blusgs"), (0.25, "blusgs"), (0.125, "bdplur")])
def test_multiblock_channel_parity(tmp_path, refine, solver):
    """The benchmark's multiblock laminar channel (eight blocks joined by
    interblock connections, spanwise periodic, no-slip wall, AUSMPW+, BDF2
    with block solver subiterations) at reduced resolution: per-step
    residuals from a common state, then the free trajectory."""
    path, meta = cases.channel_case(str(tmp_path), refine=refine, steps=2, subits=3, solver=solver,
                                    sweeps=1 if solver == "blusgs" else 2)
    say("multiblock channel", refine, solver, meta["cells"], "cells")
    s = Solver(path)
    o = pyoracle.OracleCase(path)
    _common_state_parity(s, o, 2)
    s.close()
    s = Solver(path)
    o = pyoracle.OracleCase(path)
    for nn in range(2):
        s.begin_step(nn)
        o.begin_step(nn)
        for mm in range(o.nonlinear):
            s.iterate(nn, mm)
            o.iterate(nn, mm)
    Ws = np.concatenate([s.get_state(b) for b in range(o.nb)])
    Wo = np.concatenate([o.state(b) for b in range(o.nb)])
    assert rel(Ws, Wo) <= 1e-10
    s.close()


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("case", ["viscous_connection", "channel"])
def test_row_sweep_equals_tiled(tmp_path, case, precision):
    """The row-parallel block sweep (ATH_SWEEP_ROWS) gives the tiled sweep's
    bits: viscous blocks joined by a rotated connection with two sweeps per
    relax (full sweeps, partial tiles), and the multiblock channel (eight
    blocks, BDF2 subiterations), with the FP64 and the FP32 operator.  The
    tiled sweep is compared with the oracle above."""
    if case == "viscous_connection":
        import conn_cases as cc
        r0, r1 = cc.orientation_set()[350]
        path, _ = cc.write_pair(str(tmp_path), "cut", r0, r1, timeint="implicitEuler", eq="navierStokes",
                                scale=1e-3, solver="blusgs", sweeps=2)
        steps, subits = 4, 1
    else:
        path, meta = cases.channel_case(str(tmp_path), refine=0.25, steps=2, subits=3)
        steps, subits = 2, 3
    out = []
    for order in ("TILED", "ROWS"):
        say("row sweep", case, precision, order)
        s = Solver(path)
        s.set_implicit_precision(precision)
        s.set_sweep_ordering(getattr(Solver, order))
        assert s.sweep_ordering() == getattr(Solver, order)
        norms = []
        for nn in range(steps):
            s.begin_step(nn)
            for mm in range(subits):
                norms.append(s.iterate(nn, mm).l2_sq[:5])
        out.append((np.concatenate([s.get_state(b).reshape(-1) for b in range(s.info.num_blocks)]), norms))
        s.close()
    assert np.array_equal(out[0][0], out[1][0])
    assert out[0][1] == out[1][1]


def test_split_calls_equal_iterate_on_the_channel(tmp_path):
    """The C ABI's separate calls (ghost cells, residual, relax, update) give
    the bits of ath_iterate on the multiblock channel, whose viscous block
    diagonal also forms the one-sweep relax's right-hand side for the
    row-parallel sweep."""
    path, meta = cases.channel_case(str(tmp_path), refine=0.25, steps=2, subits=3)
    say("split calls", "channel")
    out = []
    for split in (False, True):
        s = Solver(path)
        assert s.sweep_ordering() == Solver.ROWS or s.sweep_ordering() == Solver.PENCIL
        norms = []
        for nn in range(2):
            s.begin_step(nn)
            for mm in range(3):
                if split:
                    s.assign_ghost_cells()
                    s.calc_residual()
                    s.relax()
                    norms.append(s.update(mm).l2_sq[:5])
                else:
                    norms.append(s.iterate(nn, mm).l2_sq[:5])
        out.append((np.concatenate([s.get_state(b).reshape(-1) for b in range(s.info.num_blocks)]), norms))
        s.close()
    assert np.array_equal(out[0][0], out[1][0])
    assert out[0][1] == out[1][1]


@pytest.mark.parametrize("solver,sweeps,viscous", [("blusgs", 1, True), ("bdplur", 2, False)])
def test_crank_nicolson_step_is_twice_implicit_euler_on_gpu(tmp_path, solver, sweeps, viscous):
    """The relation that pins Crank-Nicolson in the oracle, on the GPU: one
    Crank-Nicolson relax is twice the implicit-Euler relax at half the CFL,
    bit for bit (linear block operators, theta in -R/theta and in the time
    term of the diagonal)."""
    out = {}
    for tag, timeint, cfl in (("cn", "crankNicholson", 10.0), ("ie", "implicitEuler", 5.0)):
        d = tmp_path / tag
        d.mkdir()
        base = cases.bump_case(str(d), seed=7, n=8, iterations=1, cfl=cfl, viscous=viscous)[0]
        path = _variant(base, "v", **{"timeIntegration: implicitEuler": f"timeIntegration: {timeint}",
                                      "matrixSolver: lusgs": f"matrixSolver: {solver}",
                                      "matrixSweeps: 1": f"matrixSweeps: {sweeps}"})
        say("crank-nicolson relation", solver, sweeps, viscous, tag)
        s = Solver(path)
        s.begin_step(0)
        s.assign_ghost_cells()
        s.calc_residual()
        s.relax()
        out[tag] = np.concatenate([s.update_vector(b).reshape(-1) for b in range(s.info.num_blocks)])
        s.close()
    assert np.abs(out["ie"]).max() > 0
    assert np.array_equal(out["cn"], 2.0 * out["ie"])


def test_sweep_order_changed_between_residual_and_relax(tmp_path):
    """A residual computed with one sweep order and a relax run with another
    (the row-parallel order's relax reads a right-hand side that its
    residual stage forms) give the same update as either order alone."""
    path, meta = cases.channel_case(str(tmp_path), refine=0.125, steps=1, subits=1)
    say("order change between residual and relax")
    ups = []
    for first, second in (("TILED", "TILED"), ("ROWS", "ROWS"), ("TILED", "ROWS"), ("ROWS", "TILED")):
        s = Solver(path)
        s.begin_step(0)
        s.set_sweep_ordering(getattr(Solver, first))
        s.assign_ghost_cells()
        s.calc_residual()
        s.set_sweep_ordering(getattr(Solver, second))
        s.relax()
        ups.append(np.concatenate([s.update_vector(b).reshape(-1) for b in range(s.info.num_blocks)]))
        s.close()
    assert np.abs(ups[0]).max() > 0
    for u in ups[1:]:
        assert np.array_equal(u, ups[0])
