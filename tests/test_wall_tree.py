"""The heat-flux walls' distance search (aither_amd/csrc/wall_tree.hpp, the
k-d tree standing for kdtree::NearestNeighbor, src/kdtree.cpp:123-224)
returns the full scan's minimum squared distance bit for bit: point sets
from 0 to 5000 points (uniform, a curved wall's face centres, a lattice with
repeated points and equal distances), queried on, near and far from the
set.  Host C++ only, built with the library's floating-point flags."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "aither_amd", "csrc")


def test_wall_tree_matches_full_scan(tmp_path):
    cxx = shutil.which("g++")
    if cxx is None:
        pytest.skip("no g++")
    exe = str(tmp_path / "wall_tree_check")
    subprocess.run([cxx, "-std=c++17", "-O3", "-ffp-contract=off", "-I", CSRC,
                    os.path.join(HERE, "wall_tree_check.cpp"), "-o", exe], check=True)
    out = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.startswith("ok"), out.stdout
