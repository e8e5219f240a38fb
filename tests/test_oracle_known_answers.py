"""The oracle against the reference's own regression known answers.

The reference's tests run the full solver on its own grids and compare the
normalised L2 residual of the last iteration with stored values at 1 %
relative (testCases/regressionTests.py:93-115).  These two cases cover the
Euler path the oracle restates: implicit Euler + LU-SGS with stagnation
inlet / pressure outlet / slip walls and unlimited third-order MUSCL
(subsonicCylinder), explicit Euler with van Albada MUSCL and supersonic
inflow/outflow (supersonicWedge), two blocks joined by interblock
connections with AUSMPW+ and a CFL ramp (multiblockCylinder), and a two-block
Sod shock tube with WENO5, BDF2 dual time stepping (five nonlinear iterations
per step, dual-time CFL) and per-block initial states (shockTube), laminar
Navier-Stokes with an adiabatic no-slip wall (viscousFlatPlate), and laminar
Couette flow between isothermal walls, one moving, with a translational
periodic connection (couette).
"""
import json
import os

import numpy as np
import pytest

from pyoracle import run_reference_loop

with open(os.path.join(os.path.dirname(__file__), "golden", "known_answers.json")) as f:
    KNOWN = {k: v for k, v in json.load(f).items() if not k.startswith("_")}


@pytest.mark.parametrize("case", sorted(KNOWN))
def test_known_answer(case, golden):
    ka = KNOWN[case]
    c, hist = run_reference_loop(os.path.join(golden, ka["inp"]), ka["iterations"])
    got = hist[-1]["normalized"]
    for e, want in enumerate(ka["residuals"]):
        if e in ka["ignore"]:
            continue
        assert abs(got[e] - want) <= 0.01 * abs(want), (case, e, got[e], want)
    # the printed values carry 4 significant digits; the restatement matches
    # every checked one exactly at that precision (one case is documented as
    # held to the reference's tolerance only)
    if not ka.get("exact", True):
        return
    printed = [float(f"{v:.4e}") for v in got]
    for e, want in enumerate(ka["residuals"]):
        if e not in ka["ignore"]:
            assert printed[e] == want, (case, e, printed[e], want)
