"""Parity of the HIP path (through the C ABI) with the FP64 oracle.

Bars (FP64 residual path, FP64 implicit operator unless stated):
  * face fluxes: relative difference <= 1e-13 per face;
  * per-step residual field: ||R_gpu - R_ref||_2 / ||R_ref||_2 <= 1e-12 while
    ||R_ref|| >= 1e-4 ||R_ref,0||, else ||R_gpu - R_ref||_2 <= 1e-12 ||R_ref,0||;
  * solution after the run: ||W_gpu - W_ref||_2 / ||W_ref||_2 <= 1e-10;
  * the reference-held known answers (1 %, testCases/regressionTests.py).
Both sides evaluate the reference's operation sequence without fused
multiply-add, and the GPU state is bit-identical to the oracle's on every
case that does not call std::pow.  std::pow appears in the stagnation-inlet
state (non-integer exponent), the AUSMPW+ pressure weight (cube) and the WENO
smoothness integrals (cubes) and Sutherland's law (T^1.5): the device
evaluates it correctly rounded,
while glibc misrounds about 0.09 % of near-tie inputs, so those cases agree
to the last bit only until the first such input.  In subsonicCylinder
(stagnation inlet, CFL 1e5, residual not decaying) a one-ulp difference in a
ghost state is amplified by the residual's cancellation along the trajectory
(1e-10 relative after 100 iterations while the states stay within 1e-13); in
multiblockCylinder (AUSMPW+, CFL ramp to 100) the per-step residual drifts to
1.1e-12 by step 84.  For the reference cases that call std::pow the per-step
residual bar is checked from a common state (the oracle restarts each step
from the GPU state) and the free trajectory is checked on the state.  The norm reductions
are summed in a different order on the GPU.
"""
import json
import os
import sys

import numpy as np
import pytest

import aither_amd
from aither_amd import Solver, cases
import pyoracle

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
with open(os.path.join(GOLDEN, "known_answers.json")) as f:
    KNOWN = {k: v for k, v in json.load(f).items() if not k.startswith("_")}


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def make_cases(tmp):
    out = {}
    for name, ka in KNOWN.items():
        out[name] = os.path.join(GOLDEN, ka["inp"])
    out["vortex12"] = cases.vortex_case(os.path.join(tmp, "v12"), n=12, iterations=30)
    out["bump10"] = cases.bump_case(os.path.join(tmp, "b10"), seed=11, n=10, iterations=30)[0]
    return out


@pytest.fixture(scope="module")
def case_paths(tmp_path_factory):
    return make_cases(str(tmp_path_factory.mktemp("cases")))


@pytest.mark.parametrize("flux", [0, 1])
def test_face_flux_matches_oracle(flux):
    rng = np.random.default_rng(10 + flux)
    n = 4096
    gas = pyoracle.Gas(R=1 / 1.4, n=2.5)
    WL = np.empty((n, 5))
    WR = np.empty((n, 5))
    for W in (WL, WR):
        W[:, 0] = rng.uniform(0.3, 3.0, n)
        W[:, 4] = rng.uniform(0.3, 3.0, n) / 1.4
        W[:, 1:4] = rng.normal(size=(n, 3)) * rng.uniform(0.01, 2.0, (n, 1))
    nv = rng.normal(size=(n, 3))
    nv /= np.linalg.norm(nv, axis=1, keepdims=True)
    say("device flux", flux, "faces", n)
    F = aither_amd.eval_face_flux(flux, gas.R, gas.n, WL, WR, nv)
    fn = pyoracle.lib().orc_roe_flux if flux == 0 else pyoracle.lib().orc_ausm_flux
    Fo = np.empty(5)
    worst = 0.0
    for q in range(n):
        fn(pyoracle.ctypes.byref(gas), pyoracle.P(WL[q].copy()), pyoracle.P(WR[q].copy()),
           pyoracle.P(nv[q].copy()), pyoracle.P(Fo))
        scale = np.abs(Fo).max()
        if scale == 0.0:
            assert np.abs(F[q]).max() == 0.0
            continue
        worst = max(worst, np.abs(F[q] - Fo).max() / scale)
    assert worst <= 1e-13, worst


@pytest.mark.parametrize("name", ["subsonicCylinder", "supersonicWedge", "vortex12", "bump10"])
def test_residual_dt_diagonal_parity(case_paths, name):
    inp = case_paths[name]
    say("residual parity", name)
    s = Solver(inp)
    o = pyoracle.OracleCase(inp)
    s.begin_step(0)
    o.begin_step(0)
    s.assign_ghost_cells()
    s.calc_residual()
    o.bc()
    o.calc_residual()
    o.time_step()
    if s.info.implicit:
        o.invert_diagonal()
    for b in range(o.nb):
        assert rel(s.residual(b), o.residual(b)) <= 1e-12
        assert rel(s.spectral_radius(b), o.spectral_radius(b)) <= 1e-14
        assert rel(s.dt(b), o.dt(b)) <= 1e-14
        if s.info.implicit:
            assert rel(s.diagonal(b), o.diagonal(b)) <= 1e-14
    s.close()


# cases whose free trajectory is not bit-identical (std::pow); their per-step
# residual is compared from a common state
RESYNC = {"subsonicCylinder", "multiblockCylinder", "shockTube", "viscousFlatPlate", "couette"}


def run_both(inp, iterations, precision=0, check_fields=True, step_tol=1e-12, resync=False):
    s = Solver(inp)
    s.set_implicit_precision(precision)
    o = pyoracle.OracleCase(inp)
    r0 = None
    hist_s, hist_o = [], []
    for nn in range(iterations):
        if resync:
            for b in range(o.nb):
                o.set_state(s.get_state(b), b)
        s.begin_step(nn)
        o.begin_step(nn)
        for mm in range(o.nonlinear):
            ns = s.iterate(nn, mm)
            no = o.iterate(nn, mm)
            hist_s.append(ns.as_dict())
            hist_o.append(no)
            if check_fields and precision == 0:
                Rs = np.concatenate([s.residual(b) for b in range(o.nb)])
                Ro = np.concatenate([o.residual(b) for b in range(o.nb)])
                nro = np.linalg.norm(Ro)
                if r0 is None:
                    r0 = nro
                if nro >= 1e-4 * r0:
                    assert rel(Rs, Ro) <= step_tol, (nn, mm, rel(Rs, Ro))
                else:
                    assert np.linalg.norm(Rs - Ro) <= step_tol * r0
    return s, o, hist_s, hist_o


@pytest.mark.parametrize("name", sorted(KNOWN))
def test_known_answers_on_gpu(case_paths, name):
    ka = KNOWN[name]
    say("known answer", name, ka["iterations"], "iterations")
    if name in RESYNC:
        s, o, hs, ho = run_both(case_paths[name], ka["iterations"], resync=True)
        s.close()
    s, o, hs, ho = run_both(case_paths[name], ka["iterations"], check_fields=name not in RESYNC)
    got = hs[-1]["normalized"]
    for e, want in enumerate(ka["residuals"]):
        if e not in ka["ignore"]:
            assert abs(got[e] - want) <= 0.01 * abs(want), (e, got[e], want)
    for b in range(o.nb):
        assert rel(s.get_state(b), o.state(b)) <= 1e-10
    # norms agree with the oracle's through the whole history
    rtol = 1e-6 if name in RESYNC else 1e-10
    for a, c in zip(hs, ho):
        # the ignored equation of the 2D cases is rounding noise (~1e-14)
        atol = rtol * max(c["l2"]) if name in RESYNC else 0.0
        assert np.allclose(a["l2"], c["l2"], rtol=rtol, atol=atol)
        assert abs(a["matrix_resid"] - c["matrix_resid"]) <= rtol * max(c["matrix_resid"], 1e-300)
    s.close()


@pytest.mark.parametrize("name", ["vortex12", "bump10"])
def test_generated_cases_state_parity(case_paths, name):
    say("state parity", name)
    s, o, hs, ho = run_both(case_paths[name], 30)
    assert rel(s.get_state(), o.state()) <= 1e-10
    assert hs[-1]["linf_loc"] == tuple(ho[-1]["linf_loc"])
    s.close()


def test_benchmark_configuration_is_bit_identical(tmp_path):
    """The benchmark workload's numerics (vortex, characteristic BCs, Roe,
    van Albada MUSCL, LU-SGS) on a 20^3 grid: after 20 iterations every state
    value equals the oracle's bit for bit."""
    inp = cases.vortex_case(str(tmp_path), n=20, iterations=20)
    say("bit-identical vortex 20^3")
    s, o, hs, ho = run_both(inp, 20)
    assert np.array_equal(s.get_state(), o.state())
    assert np.array_equal(s.residual(), o.residual())
    s.close()


def test_graph_run_equals_iterate(tmp_path):
    inp = cases.vortex_case(str(tmp_path), n=16, iterations=6)
    say("graph replay vs direct launches")
    a = Solver(inp)
    for nn in range(6):
        a.begin_step(nn)
        na = a.iterate(nn, 0)
    b = Solver(inp)
    nb = b.run(0, 6)
    assert np.array_equal(a.get_state(), b.get_state())
    assert list(na.normalized[:5]) == list(nb.normalized[:5])
    a.close()
    b.close()


def test_dplur_and_multisweep_lusgs(tmp_path):
    base = cases.vortex_case(str(tmp_path), n=10, iterations=5)
    for solver, sweeps in (("dplur", 3), ("lusgs", 2)):
        text = open(base).read().replace("matrixSolver: lusgs", f"matrixSolver: {solver}")
        text = text.replace("matrixSweeps: 1", f"matrixSweeps: {sweeps}")
        path = os.path.join(str(tmp_path), f"{solver}.inp")
        open(path, "w").write(text)
        say("solver", solver, sweeps)
        s, o, hs, ho = run_both(path, 5)
        assert rel(s.get_state(), o.state()) <= 1e-10
        s.close()


def test_explicit_and_constant_reconstruction(tmp_path):
    base = cases.vortex_case(str(tmp_path), n=10, iterations=5, cfl=0.5)
    text = open(base).read().replace("timeIntegration: implicitEuler", "timeIntegration: explicitEuler")
    text = text.replace("faceReconstruction: thirdOrder", "faceReconstruction: constant")
    path = os.path.join(str(tmp_path), "expl.inp")
    open(path, "w").write(text)
    say("explicit Euler, constant reconstruction")
    s, o, hs, ho = run_both(path, 5)
    assert rel(s.get_state(), o.state()) <= 1e-12
    s.close()


@pytest.mark.parametrize("recon", ["weno", "wenoZ"])
@pytest.mark.parametrize("grid", ["vortex", "bump"])
def test_weno_reconstruction(tmp_path, recon, grid):
    """WENO5 / WENO-Z with three ghost layers, on the uniform vortex grid and
    the bump grid (non-uniform widths), Roe and AUSMPW+."""
    if grid == "vortex":
        path = cases.vortex_case(str(tmp_path), n=10, iterations=6, recon=recon)
    else:
        path = cases.bump_case(str(tmp_path), seed=4, n=10, iterations=6, recon=recon)[0]
    for flux in ("roe", "ausm"):
        text = open(path).read()
        if flux == "ausm":
            text = text.replace("boundaryConditions:", "inviscidFlux: ausm\nboundaryConditions:", 1)
        p = os.path.join(str(tmp_path), f"{flux}.inp")
        open(p, "w").write(text)
        say("WENO parity", recon, grid, flux)
        s, o, hs, ho = run_both(p, 6)
        assert s.info.num_ghosts == 3
        assert rel(s.get_state(), o.state()) <= 1e-10
        s.close()


def test_ausm_minmod(tmp_path):
    path = cases.vortex_case(str(tmp_path), n=10, iterations=5, flux="ausm", limiter="minmod")
    say("AUSMPW+ with minmod")
    s, o, hs, ho = run_both(path, 5)
    assert rel(s.get_state(), o.state()) <= 1e-10
    s.close()


def test_fp32_implicit_operator(case_paths):
    """FP32 LU-SGS operator with the FP64 residual: the known answer of the
    implicit case still holds at the reference's 1 % tolerance."""
    ka = KNOWN["subsonicCylinder"]
    say("fp32 operator", "subsonicCylinder")
    s, o, hs, ho = run_both(case_paths["subsonicCylinder"], ka["iterations"], precision=1, check_fields=False)
    got = hs[-1]["normalized"]
    for e, want in enumerate(ka["residuals"]):
        if e not in ka["ignore"]:
            assert abs(got[e] - want) <= 0.01 * abs(want), (e, got[e], want)
    s.close()


def test_full_size_vortex_first_iterations(tmp_path):
    """The 64^3 benchmark case: two iterations against the oracle, and free-
    stream preservation (uniform state => R = 0) at full size."""
    inp = cases.vortex_case(str(tmp_path), n=64, iterations=2)
    say("64^3 vortex, 2 iterations")
    s, o, hs, ho = run_both(inp, 2)
    assert rel(s.get_state(), o.state()) <= 1e-12
    W = s.get_state()
    W[:] = W[0] * 0 + np.array([1.0, 0.5, 0.0, 0.0, 1 / 1.4])
    s.set_state(W)
    s.begin_step(0)
    s.assign_ghost_cells()
    s.calc_residual()
    R = s.residual()
    s.close()
    # ghost states from the characteristic BC at a different freestream make
    # only the boundary layer of cells nonzero; interior cells are exactly
    # free-stream preserving on the uniform grid
    Rb = R.reshape(64, 64, 64, 5)[2:-2, 2:-2, 2:-2]
    assert np.abs(Rb).max() <= 1e-15


@pytest.mark.parametrize("dims,sweeps,precision", [((20, 20, 20), 1, 0), ((13, 11, 7), 1, 0),
                                                   ((13, 11, 7), 3, 0), ((17, 9, 10), 2, 1),
                                                   ((9, 17, 3), 1, 1)])
@pytest.mark.parametrize("ordering", ["PENCIL", "TILED"])
def test_sweep_orderings_equal_hyperplane_order(tmp_path, dims, sweeps, precision, ordering):
    """The pencil-wavefront and tiled dataflow sweeps and the one-launch-per-
    hyperplane sweep are topological orders of the same Gauss-Seidel
    dependencies: identical bits."""
    ni, nj, nk = dims
    base = cases.vortex_case(str(tmp_path), n=max(dims), iterations=4, ni=ni, nj=nj, nk=nk)
    text = open(base).read().replace("matrixSweeps: 1", f"matrixSweeps: {sweeps}")
    open(base, "w").write(text)
    say(ordering, "vs hyperplane", dims, sweeps, precision)
    out = []
    for order in (getattr(Solver, ordering), Solver.HYPERPLANE):
        s = Solver(base)
        s.set_implicit_precision(precision)
        s.set_sweep_ordering(order)
        for nn in range(4):
            s.begin_step(nn)
            s.iterate(nn, 0)
        out.append((s.get_state(), s.update_vector()))
        s.close()
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1])


# two-block cases cut from one block, each half re-indexed by a proper
# rotation: cut surfaces on every face type and several relative orientations
CONN_PAIRS = [0, 97, 211, 350, 463, 575]


@pytest.mark.parametrize("q", CONN_PAIRS)
def test_connections_every_sweep_ordering(tmp_path, q):
    """Interblock connections (ghost state gather, lagged du terms before
    every pass, matrix residual) against the oracle for the pencil, tiled and
    hyperplane LU-SGS orders: per-step residual and final state."""
    import conn_cases as cc
    r0, r1 = cc.orientation_set()[q]
    path, _ = cc.write_pair(str(tmp_path), "cut", r0, r1, timeint="implicitEuler")
    for order in ("PENCIL", "TILED", "HYPERPLANE"):
        say("connections", q, order)
        s = Solver(path)
        s.set_sweep_ordering(getattr(Solver, order))
        o = pyoracle.OracleCase(path)
        for nn in range(4):
            s.begin_step(nn)
            o.begin_step(nn)
            ns = s.iterate(nn, 0)
            no = o.iterate(nn, 0)
            Rs = np.concatenate([s.residual(b) for b in range(o.nb)])
            Ro = np.concatenate([o.residual(b) for b in range(o.nb)])
            assert rel(Rs, Ro) <= 1e-12, (order, nn, rel(Rs, Ro))
            assert abs(ns.matrix_resid - no["matrix_resid"]) <= 1e-10 * no["matrix_resid"]
        Ws = np.concatenate([s.get_state(b) for b in range(o.nb)])
        Wo = np.concatenate([o.state(b) for b in range(o.nb)])
        assert rel(Ws, Wo) <= 1e-12, (order, rel(Ws, Wo))
        s.close()


@pytest.mark.parametrize("solver,sweeps", [("lusgs", 3), ("dplur", 3)])
def test_connections_multisweep_and_dplur(tmp_path, solver, sweeps):
    import conn_cases as cc
    r0, r1 = cc.orientation_set()[350]
    path, _ = cc.write_pair(str(tmp_path), "cut", r0, r1, timeint="implicitEuler", solver=solver,
                            sweeps=sweeps)
    say("connections", solver, sweeps)
    s, o, hs, ho = run_both(path, 4)
    Ws = np.concatenate([s.get_state(b) for b in range(o.nb)])
    Wo = np.concatenate([o.state(b) for b in range(o.nb)])
    assert rel(Ws, Wo) <= 1e-12
    s.close()


def test_connections_fp32_operator(tmp_path):
    """FP32 implicit operator across connections: close to the FP64 oracle
    after four iterations (the operator's rounding only)."""
    import conn_cases as cc
    r0, r1 = cc.orientation_set()[463]
    path, _ = cc.write_pair(str(tmp_path), "cut", r0, r1, timeint="implicitEuler")
    say("connections fp32")
    s, o, hs, ho = run_both(path, 4, precision=1, check_fields=False)
    Ws = np.concatenate([s.get_state(b) for b in range(o.nb)])
    Wo = np.concatenate([o.state(b) for b in range(o.nb)])
    assert rel(Ws, Wo) <= 1e-6
    s.close()


def _parity_from_common_state(s, o, iterations, tol=1e-12):
    """per-step residual and matrix residual from a common state (the oracle
    restarts each step from the GPU state)"""
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
                assert abs(ns.matrix_resid - no["matrix_resid"]) <= 1e-10 * no["matrix_resid"]


@pytest.mark.parametrize("visc_recon", ["central", "centralFourth"])
def test_viscous_channel_every_sweep_ordering(tmp_path, visc_recon):
    """Laminar channel: adiabatic wall on the bump, isothermal wall moving
    along x on top, characteristic inflow/outflow, slip side walls; the
    viscous residual (edge and wall ghost cells, face gradients, Sutherland
    transport, viscous spectral radii) and the viscous implicit terms in the
    pencil, tiled and hyperplane LU-SGS orders."""
    path = cases.bump_case(str(tmp_path), seed=4, n=10, iterations=6, viscous=True, visc_recon=visc_recon)[0]
    for order in ("PENCIL", "TILED", "HYPERPLANE"):
        say("viscous channel", visc_recon, order)
        s = Solver(path)
        s.set_sweep_ordering(getattr(Solver, order))
        o = pyoracle.OracleCase(path)
        _parity_from_common_state(s, o, 6)
        s.close()
    # the free trajectory
    s, o, hs, ho = run_both(path, 6, check_fields=False)
    assert rel(s.get_state(), o.state()) <= 1e-10
    s.close()


def test_viscous_explicit_dplur_fp32(tmp_path):
    base = cases.bump_case(str(tmp_path), seed=5, n=8, iterations=4, viscous=True)[0]
    text = open(base).read()
    variants = {
        "explicit": text.replace("timeIntegration: implicitEuler", "timeIntegration: explicitEuler")
                        .replace("cflStart: 5.0", "cflStart: 0.5").replace("cflMax: 5.0", "cflMax: 0.5"),
        "dplur": text.replace("matrixSolver: lusgs", "matrixSolver: dplur").replace("matrixSweeps: 1", "matrixSweeps: 3"),
        "lusgs3": text.replace("matrixSweeps: 1", "matrixSweeps: 3"),
    }
    for name, t in variants.items():
        p = os.path.join(str(tmp_path), name + ".inp")
        open(p, "w").write(t)
        say("viscous", name)
        s = Solver(p)
        o = pyoracle.OracleCase(p)
        _parity_from_common_state(s, o, 4)
        s.close()
    say("viscous fp32 operator")
    s, o, hs, ho = run_both(base, 4, precision=1, check_fields=False)
    assert rel(s.get_state(), o.state()) <= 1e-6
    s.close()


def test_viscous_connections(tmp_path):
    """Two blocks joined in a rotated orientation with the viscous terms on:
    the connection's ghost viscosity and centre distance enter the implicit
    terms across the cut."""
    import conn_cases as cc
    r0, r1 = cc.orientation_set()[350]
    path, _ = cc.write_pair(str(tmp_path), "cut", r0, r1, timeint="implicitEuler", eq="navierStokes", scale=1e-3)
    for order in ("PENCIL", "HYPERPLANE"):
        say("viscous connections", order)
        s = Solver(path)
        s.set_sweep_ordering(getattr(Solver, order))
        o = pyoracle.OracleCase(path)
        _parity_from_common_state(s, o, 4)
        s.close()


@pytest.mark.parametrize("eq", ["euler", "navierStokes"])
def test_periodic_every_sweep_ordering(tmp_path, eq):
    """A block periodic in i (ghost states and lagged du across the periodic
    connection, mirrored ghost geometry) against the oracle."""
    from test_oracle_connections import _periodic_pair
    path = _periodic_pair(str(tmp_path), eq=eq)
    for order in ("PENCIL", "TILED", "HYPERPLANE"):
        say("periodic", eq, order)
        s = Solver(path)
        s.set_sweep_ordering(getattr(Solver, order))
        o = pyoracle.OracleCase(path)
        _parity_from_common_state(s, o, 4)
        s.close()
