"""Interblock connections in the oracle, pinned without the reference.

A block cut in two, with each half re-indexed by any proper rotation of the
index cube, must give the uncut block's residual to rounding: the ghost cells
of a connection (states, volumes, centroids, face areas and centres, hence
cell widths) continue the partner block as if the grid were not cut
(src/utility.cpp:180-256, src/procBlock.cpp:3167-4011).  All 576 rotation
pairs are checked; they cover all six cut surfaces on each side, the eight
relative orientations of the patch axes (src/boundaryConditions.cpp:654-727)
and the lower/upper, lower/lower and upper/upper pairings.  The implicit
coupling across connections (lagged du swapped before every pass) is pinned
by the multiblockCylinder and shockTube known answers.
"""
import os

import numpy as np
import pytest

import pyoracle
import conn_cases as cc

N = (6, 5, 8)


def residual_by_uncut_cell(o, cmap):
    Rb = [o.residual(b) for b in range(o.nb)]
    out = np.zeros((N[2], N[1], N[0], 5))
    for k in range(N[2]):
        for j in range(N[1]):
            for i in range(N[0]):
                b, (a0, a1, a2) = cmap((i, j, k))
                ni, nj, _ = o.dims[b]
                out[k, j, i] = Rb[b][a0 + ni * (a1 + nj * a2)]
    return out


def state_by_uncut_cell(o, cmap):
    Wb = [o.state(b) for b in range(o.nb)]
    out = np.zeros((N[2], N[1], N[0], 5))
    for k in range(N[2]):
        for j in range(N[1]):
            for i in range(N[0]):
                b, (a0, a1, a2) = cmap((i, j, k))
                ni, nj, _ = o.dims[b]
                out[k, j, i] = Wb[b][a0 + ni * (a1 + nj * a2)]
    return out


@pytest.mark.parametrize("recon,tol", [("thirdOrder", 1e-14), ("weno", 1e-12), ("wenoZ", 1e-12)])
def test_cut_block_residual_equals_uncut(tmp_path, recon, tol):
    p0, _ = cc.write_pair(str(tmp_path / "u"), "uncut", recon=recon, n=N)
    o = pyoracle.OracleCase(p0)
    o.bc()
    o.calc_residual()
    R0 = o.residual(0).reshape(N[2], N[1], N[0], 5)
    scale = np.abs(R0).max()
    pairs = cc.orientation_set()
    assert len(pairs) == 576
    surfs = set()
    for q, (r0, r1) in enumerate(pairs):
        p, cmap = cc.write_pair(str(tmp_path / f"c{q}"), "cut", r0, r1, recon=recon, n=N)
        oc = pyoracle.OracleCase(p)
        oc.bc()
        oc.calc_residual()
        err = np.abs(residual_by_uncut_cell(oc, cmap) - R0).max() / scale
        assert err <= tol, (q, r0, r1, err)
        surfs.add((cc.cut_surface(r0, False), cc.cut_surface(r1, True)))
    assert len(surfs) == 36


def test_cut_block_explicit_steps_equal_uncut(tmp_path):
    """Five explicit Euler steps: the state of every cell follows the uncut
    block's to rounding."""
    p0, _ = cc.write_pair(str(tmp_path / "u"), "uncut", n=N)
    o = pyoracle.OracleCase(p0)
    for nn in range(5):
        o.begin_step(nn)
        o.iterate(nn, 0)
    W0 = o.state(0).reshape(N[2], N[1], N[0], 5)
    pairs = cc.orientation_set()
    for q in range(0, len(pairs), 23):
        r0, r1 = pairs[q]
        p, cmap = cc.write_pair(str(tmp_path / f"c{q}"), "cut", r0, r1, n=N)
        oc = pyoracle.OracleCase(p)
        for nn in range(5):
            oc.begin_step(nn)
            oc.iterate(nn, 0)
        W = state_by_uncut_cell(oc, cmap)
        assert np.abs(W - W0).max() <= 1e-13 * np.abs(W0).max(), q


def test_unmatched_connection_is_an_error(tmp_path):
    r = cc.rotations()[0]
    p, _ = cc.write_pair(str(tmp_path), "cut", r, r, n=N)
    text = open(p).read()
    # point block 0's connection at block 1's k-upper face, which is not
    # an interblock patch
    bad = text.replace("4     4  5001", "4     4  6001")
    assert bad != text
    open(p, "w").write(bad)
    with pytest.raises(RuntimeError, match="interblock"):
        pyoracle.OracleCase(p)
