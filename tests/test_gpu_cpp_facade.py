"""Runs the C++ facade test program (the reference's ie_solver_test.cpp
cases through include/ie_solver_gpu.hpp) on the GPU, at the reference's
default ID tolerance and at 1e-10."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "ie_solver_gpu_test.bin")


@pytest.mark.parametrize("tol", ["1e-6", "1e-10"])
def test_reference_cases_through_cpp_facade(tol):
    assert os.path.exists(BIN), "build first: python ie-solver_amd/build.py"
    env = dict(os.environ, IE_TEST_TOL=tol)
    p = subprocess.run([BIN], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       env=env, timeout=300)
    print(p.stdout, flush=True)
    assert p.returncode == 0, p.stdout
