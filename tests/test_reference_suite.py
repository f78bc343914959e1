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


def test_reference_validation_tool_answers_match_the_oracle(tmp_path):
    """The reference's validation tool (tools/validate_against_reference.cpp,
    unchanged) generates its six instance flavours (mixed, degenerate,
    badly scaled, staircase, infeasible, unbounded), solves them at 1e-8
    through pdlp::PdlpSolver (here: the GPU engine) and writes .lp.txt and
    .answer.txt files.  Each written instance is read back and solved by the
    oracle: same status, objective within 1e-6 relative, the answer's x
    primal-feasible to the oracle's evaluator."""
    import oracle_lib as ol
    from optimisation_solver_amd import default_options
    from optimisation_solver_amd.lpio import read_answer, read_lp_txt

    r = _run("validate_on_engine", str(tmp_path), "6")
    assert r.returncode == 0
    o = default_options(primal_tolerance=1e-8, dual_tolerance=1e-8, gap_tolerance=1e-8, thread_count=1,
                        iteration_limit=300000, time_limit_seconds=60.0)
    for i in range(6):
        lp = read_lp_txt(str(tmp_path / f"instance_{i}.lp.txt"))
        ans = read_answer(str(tmp_path / f"instance_{i}.answer.txt"))
        cpu = ol.solve(lp, o)
        print(f"instance {i}: gpu {ans.status} {ans.primal_objective:.10g}; cpu {cpu.status_name} "
              f"{cpu.primal_objective:.10g}", flush=True)
        assert ans.status == cpu.status_name
        if ans.status == "optimal":
            assert abs(ans.primal_objective - cpu.primal_objective) / (1 + abs(cpu.primal_objective)) <= 1e-6
            import numpy as np
            m = ol.evaluate(lp, o, ans.primal, np.zeros(lp.m))
            assert m["primal_residual"] <= 1e-8
