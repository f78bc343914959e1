"""Runs the C++ test binaries (built by __graft_entry__.build()) on the GPU:
tests/cpp/test_dropin.cpp, the reference's unit-test scenarios through this
repository's pdlp::PdlpSolver headers, and tests/cpp/test_abi_paths.cpp, the
C ABI's batch, slicing, tuning and profiling paths driven from C++."""
import subprocess
from pathlib import Path

import pytest

BIN = Path(__file__).resolve().parent / "cpp" / "test_dropin"
ABI_BIN = Path(__file__).resolve().parent / "cpp" / "test_abi_paths"
ROOT = Path(__file__).resolve().parent.parent


@pytest.mark.gpu
def test_cpp_dropin_suite():
    assert BIN.exists(), "build first: python -m optimisation_solver_amd.build"
    p = subprocess.run([str(BIN)], capture_output=True, text=True, timeout=110)
    print(p.stdout)
    assert p.returncode == 0, p.stdout + p.stderr


@pytest.mark.gpu
def test_cpp_abi_paths():
    assert ABI_BIN.exists(), "build first: python -m optimisation_solver_amd.build"
    p = subprocess.run([str(ABI_BIN)], capture_output=True, text=True, timeout=110, cwd=ROOT)
    print(p.stdout)
    assert p.returncode == 0, p.stdout + p.stderr
    assert "0 failure(s)" in p.stdout
