"""Runs the C++ drop-in test binary (tests/cpp/test_dropin.cpp, built by
__graft_entry__.build()) on the GPU: the reference's unit-test scenarios
through this repository's pdlp::PdlpSolver headers."""
import subprocess
from pathlib import Path

import pytest

BIN = Path(__file__).resolve().parent / "cpp" / "test_dropin"


@pytest.mark.gpu
def test_cpp_dropin_suite():
    assert BIN.exists(), "build first: python -m optimisation_solver_amd.build"
    p = subprocess.run([str(BIN)], capture_output=True, text=True, timeout=110)
    print(p.stdout)
    assert p.returncode == 0, p.stdout + p.stderr
