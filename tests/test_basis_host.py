"""Host build of the kernels' shape-function header (jump_amd/csrc/mpm_basis.hpp):
the branch-free quadratic B-spline the kernels use equals the piecewise
definition of Bsplines.jl:18-83 bit for bit at every knot, its neighbouring
doubles and a dense grid, for all four boundary types."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_branch_free_quadratic_bspline_is_bitwise_piecewise(tmp_path):
    exe = str(tmp_path / "basis_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-fno-fast-math",
                    os.path.join(HERE, "host", "basis_check.cpp"), "-o", exe], check=True)
    out = subprocess.run([exe], capture_output=True, text=True)
    assert out.returncode == 0, out.stdout + out.stderr
    assert int(out.stdout.strip()) > 2_000_000
