"""Runs the reference's own test cases against the C++ drop-in classes
(moleculardynamics_amd/host/moldyn_gpu.hpp) on the GPU: tests/cpp/test_mirror.cpp."""
import os
import subprocess

import pytest

from conftest import GOLDEN, ROOT

pytestmark = pytest.mark.gpu

EXE = os.path.join(ROOT, "tests", "cpp", "_build", "test_mirror")


def test_reference_cases_through_cpp_mirror(gpu_ok):
    if not os.path.exists(EXE):
        pytest.fail(f"{EXE} is not built (python -m moleculardynamics_amd.build)")
    r = subprocess.run([EXE, GOLDEN], capture_output=True, text=True, timeout=240)
    print(r.stdout)
    print(r.stderr)
    assert r.returncode == 0, r.stdout[-4000:]
    assert "0 failed" in r.stdout
