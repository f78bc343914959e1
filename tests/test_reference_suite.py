"""The reference's own tests run against this repository's PDLP.

`make -C oracle dropin` and `make -C oracle pipeline` (built where the
reference sources exist; the binaries travel with the tree in oracle/_ref/)
compile, unchanged and from where they lie:
  - pdlp_engine/tests/test_pdlp.cpp: 25 PDLP unit tests (assembly, SpMV,
    Ruiz, row kinds, closed-form optima, rays, limits, options);
  - tests/adapter/test_pdlp_adapter.cpp: the model <-> CompiledLp adapter
    with its objective and dual-sign checks;
  - tests/pipeline/test_end_to_end.cpp and test_postsolve_integration.cpp:
    the full pipeline (classifier, presolve, dispatcher, orchestrator,
    postsolve; the reference's own MILP and QP engines for those models),
against include/pdlp/*.h, with every PDLP solve served by libpdlp_amd.so on
the GPU through pdlp::PdlpSolver::solve."""
import subprocess
from pathlib import Path

import pytest

REF = Path(__file__).resolve().parent.parent / "oracle" / "_ref"
CASES = ["lp_engines", "qp_signs", "all_eliminated", "bound_only", "integer_models", "unavailable_duals",
         "failure_statuses", "limit_no_solution", "limit_feasible_solution", "reduced_api"]

pytestmark = pytest.mark.gpu


def _run(name, *args):
    exe = REF / name
    if not exe.exists():
        pytest.skip(f"oracle/_ref/{name} not built (needs the reference sources)")
    r = subprocess.run([str(exe), *args], capture_output=True, text=True, timeout=240)
    print(r.stdout[-3000:], r.stderr[-2000:], flush=True)
    return r


def test_reference_pdlp_unit_tests():
    assert _run("test_pdlp_on_engine").returncode == 0


def test_reference_adapter_tests():
    assert _run("test_pdlp_adapter_on_engine").returncode == 0


def test_reference_end_to_end_pipeline():
    assert _run("test_end_to_end_on_engine").returncode == 0


@pytest.mark.parametrize("case", CASES)
def test_reference_postsolve_integration(case):
    r = _run("test_postsolve_integration_on_engine", case)
    assert r.returncode == 0 and f"[PASSED] {case}" in r.stdout
