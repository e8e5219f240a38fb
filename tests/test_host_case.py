"""Host side of the drop-in boundary (no GPU needed): the C ABI library loads
and exports every entry point of include/aither_amd.h, and the host loader
(input file, Plot3D grid, FP64 geometry, initial state) reproduces the
oracle's geometry bit for bit."""
import ctypes
import os
import re

import numpy as np
import pytest

import aither_amd
from aither_amd import cases, inspect_field, inspect_info
import pyoracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    text = open(os.path.join(ROOT, "include", "aither_amd.h")).read()
    return sorted(set(re.findall(r"\b(ath_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = aither_amd.load_library()
    names = header_functions()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n


def _cases(tmp):
    out = [os.path.join(ROOT, "tests", "golden", "subsonicCylinder", "subsonicCylinder.inp"),
           os.path.join(ROOT, "tests", "golden", "supersonicWedge", "supersonicWedge.inp"),
           os.path.join(ROOT, "tests", "golden", "shockTube", "shockTube.inp"),
           os.path.join(ROOT, "tests", "golden", "multiblockCylinder", "multiblockCylinder.inp"),
           os.path.join(ROOT, "tests", "golden", "viscousFlatPlate", "viscousFlatPlate.inp"),
           os.path.join(ROOT, "tests", "golden", "couette", "couette.inp"),
           cases.vortex_case(os.path.join(tmp, "v"), n=7, iterations=1),
           cases.bump_case(os.path.join(tmp, "b"), seed=3, n=9)[0]]
    # two-block cases with interblock connections in several orientations
    import conn_cases as cc
    pairs = cc.orientation_set()
    for q in (0, 97, 211, 350, 463, 575):
        out.append(cc.write_pair(os.path.join(tmp, f"c{q}"), "cut", *pairs[q], recon="weno")[0])
    return out


def test_host_geometry_matches_oracle(tmp_path):
    for inp in _cases(str(tmp_path)):
        orc = pyoracle.OracleCase(inp)
        info = inspect_info(inp)
        assert info.num_blocks == orc.nb
        for b in range(orc.nb):
            assert np.array_equal(inspect_field(inp, b, 0), orc.volume(b)), inp
            for d in range(3):
                assert np.array_equal(inspect_field(inp, b, 1 + d), orc.cell_width(d, b)), (inp, d)
                assert np.array_equal(inspect_field(inp, b, 4 + d), orc.face_area(d, b)), (inp, d)
            assert np.array_equal(inspect_field(inp, b, 7), orc.state(b).reshape(-1)), inp


def test_reference_quantities(tmp_path):
    inp = cases.vortex_case(str(tmp_path), n=4, iterations=1)
    info = inspect_info(inp)
    R = 8.3144598 / 28.97e-3
    assert abs(info.a_ref - np.sqrt(1.4 * R * 288.15)) < 1e-12 * info.a_ref
    assert abs(info.gamma - 1.4) < 1e-15
    assert abs(info.gas_constant - 1 / 1.4) < 1e-15
    assert info.num_ghosts == 2 and info.implicit == 1 and info.num_equations == 5


def test_errors_are_reported_not_fatal(tmp_path):
    with pytest.raises(aither_amd.AitherError) as e:
        inspect_info(str(tmp_path / "missing.inp"))
    assert e.value.code == -3
    inp = cases.vortex_case(str(tmp_path), n=4, iterations=1)
    text = open(inp).read().replace("equationSet: euler", "equationSet: rans")
    bad = str(tmp_path / "bad.inp")
    open(bad, "w").write(text)
    with pytest.raises(aither_amd.AitherError) as e:
        inspect_info(bad)
    assert e.value.code == -2 and "rans" in str(e.value)


@pytest.mark.parametrize("solver,timeint,block", [("blusgs", "implicitEuler", 1), ("bdplur", "bdf2", 1),
                                                  ("blusgs", "explicitEuler", 0), ("lusgs", "implicitEuler", 0)])
def test_block_matrix_selection(tmp_path, solver, timeint, block):
    """input::IsBlockMatrix (src/input.cpp:713-720): blusgs and bdplur give a
    block matrix only with an implicit time integration, like the oracle"""
    import pyoracle
    inp = cases.vortex_case(str(tmp_path), n=4, iterations=1, solver=solver)
    text = open(inp).read().replace("timeIntegration: implicitEuler", f"timeIntegration: {timeint}")
    open(inp, "w").write(text)
    assert inspect_info(inp).block_matrix == block
    assert pyoracle.OracleCase(inp).is_block() == bool(block)


def test_block_matrix_with_viscous_terms(tmp_path):
    """navierStokes with blusgs selects the block matrix in both readers"""
    import pyoracle
    inp = cases.bump_case(str(tmp_path), seed=1, n=4, iterations=1, viscous=True)[0]
    text = open(inp).read().replace("matrixSolver: lusgs", "matrixSolver: blusgs")
    open(inp, "w").write(text)
    assert inspect_info(inp).block_matrix == 1
    assert pyoracle.OracleCase(inp).is_block()


def test_estimate_case_bytes(tmp_path):
    """The device memory admission reserves for a case grows with the cell
    count, with the FP32 operator's buffers and with the block operator."""
    from aither_amd import Server
    small = cases.bump_case(str(tmp_path / "s"), seed=1, n=8)[0]
    big = cases.bump_case(str(tmp_path / "b"), seed=1, n=16)[0]
    e_small, e_big = Server.estimate(small, 0), Server.estimate(big, 0)
    assert e_big > 4 * (e_small - 1.1 * (1 << 20))
    assert Server.estimate(small, 1) > e_small
    blk = small.replace(".inp", "_blk.inp")
    open(blk, "w").write(open(small).read().replace("matrixSolver: lusgs", "matrixSolver: blusgs"))
    assert Server.estimate(blk, 0) > e_small + 8 ** 3 * 25 * 8
    with pytest.raises(aither_amd.AitherError):
        Server.estimate(str(tmp_path / "missing.inp"), 0)
