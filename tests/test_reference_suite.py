"""The reference's own PDLP unit tests (pdlp_engine/tests/test_pdlp.cpp, 25
tests: sparse-matrix assembly, SpMV, Ruiz scaling, statuses, closed-form
optima, rays, limits, options) compiled unchanged against this repository's
drop-in headers and linked to the GPU engine (`make -C oracle dropin`, built
where the reference sources exist; the binary travels with the tree).  Every
solve inside them runs on the GPU through pdlp::PdlpSolver::solve."""
import subprocess
from pathlib import Path

import pytest

BIN = Path(__file__).resolve().parent.parent / "oracle" / "_ref" / "test_pdlp_on_engine"


@pytest.mark.gpu
@pytest.mark.skipif(not BIN.exists(), reason="oracle/_ref/test_pdlp_on_engine not built (needs the reference sources)")
def test_reference_unit_tests_pass_on_the_gpu_engine():
    r = subprocess.run([str(BIN)], capture_output=True, text=True, timeout=240)
    print(r.stdout[-3000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0
