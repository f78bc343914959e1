"""Known-answer LPs from the reference repository's benchmark set (MPS files
and recorded optima, tests/golden/mps/known_answers.json): the MPS reader
builds the LPs, the CPU oracle (the reference PDLP restated) and the GPU
engine both solve them at the production tolerance 1e-8 (the orchestrator's
default) and must reproduce the recorded optimum to 1e-6 relative, the
recorded status for the infeasible and unbounded instances, and agree with
each other."""
import json
from pathlib import Path

import pytest

import oracle_lib as ol
from optimisation_solver_amd import Engine, default_options
from optimisation_solver_amd.mps import read_mps

GOLD = Path(__file__).resolve().parent / "golden" / "mps"
KNOWN = json.loads((GOLD / "known_answers.json").read_text())


def _opts():
    return default_options(primal_tolerance=1e-8, dual_tolerance=1e-8, gap_tolerance=1e-8, thread_count=1,
                           iteration_limit=200000)


def test_reader_shapes_and_semantics():
    m = read_mps(str(GOLD / "afiro.mps"))
    assert [m.lp.m, m.lp.n, m.lp.nnz] == KNOWN["shapes"]["afiro.mps"]
    r = read_mps(str(GOLD / "ranged_lp.mps")).lp
    assert (r.rlo[0], r.rhi[0]) == (6.0, 10.0)            # L row 10 with range 4
    f = read_mps(str(GOLD / "free_fixed.mps"))
    lo, hi = dict(zip(f.col_names, f.lp.lo)), dict(zip(f.col_names, f.lp.hi))
    assert lo["X"] == -float("inf") and hi["X"] == float("inf")
    assert lo["Y"] == hi["Y"] == 3.0 and lo["Z"] == -float("inf") and hi["Z"] == 5.0
    mo = read_mps(str(GOLD / "max_offset_lp.mps"))
    assert mo.maximize and mo.lp.offset == -100.0 and list(mo.lp.c) == [-2.0, -3.0]


@pytest.mark.parametrize("name", sorted(KNOWN["optimal"]))
def test_oracle_reaches_known_optimum(name):
    model = read_mps(str(GOLD / name))
    r = ol.solve(model.lp, _opts())
    val = model.objective(r.primal_objective)
    want = KNOWN["optimal"][name]
    assert r.status_name == "optimal"
    assert abs(val - want) <= 1e-6 * (1 + abs(want)), (name, val, want)


@pytest.mark.parametrize("name,status", [(n, "infeasible") for n in KNOWN["infeasible"]] +
                         [(n, "unbounded") for n in KNOWN["unbounded"]])
def test_oracle_statuses(name, status):
    assert ol.solve(read_mps(str(GOLD / name)).lp, _opts()).status_name == status


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(KNOWN["optimal"]) + KNOWN["infeasible"] + KNOWN["unbounded"])
def test_gpu_engine_known_answers(name):
    model = read_mps(str(GOLD / name))
    cpu = ol.solve(model.lp, _opts())
    with Engine() as e:
        e.load(model.lp)
        r = e.solve(_opts())
    print(f"{name}: gpu {r.status} {r.iterations} its objective {model.objective(r.primal_objective):.10g}; "
          f"cpu {cpu.status_name} {cpu.iterations} its", flush=True)
    assert r.status == cpu.status_name
    if name in KNOWN["optimal"]:
        want = KNOWN["optimal"][name]
        val = model.objective(r.primal_objective)
        assert abs(val - want) <= 1e-6 * (1 + abs(want))
        m = ol.evaluate(model.lp, _opts(), r.primal, r.row_dual)
        assert max(m["primal_residual"], m["dual_residual"], m["relative_gap"]) <= 1e-8
    else:
        assert r.status in ("infeasible", "unbounded")


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(KNOWN["optimal"]) + KNOWN["infeasible"] + KNOWN["unbounded"])
def test_gpu_engine_known_answers_mixed_precision(name):
    """FP32 matrix values with FP64 iterates at 1e-6: the recorded status, and
    the recorded optimum to 1e-5 relative with the oracle's KKT bar met."""
    model = read_mps(str(GOLD / name))
    o = default_options(primal_tolerance=1e-6, dual_tolerance=1e-6, gap_tolerance=1e-6, thread_count=1,
                        iteration_limit=200000, precision=2)
    with Engine() as e:
        e.load(model.lp, 2)
        r = e.solve(o)
    print(f"{name}: gpu mixed {r.status} {r.iterations} its objective {model.objective(r.primal_objective):.10g}",
          flush=True)
    if name in KNOWN["optimal"]:
        want = KNOWN["optimal"][name]
        assert r.status == "optimal"
        assert abs(model.objective(r.primal_objective) - want) <= 1e-5 * (1 + abs(want))
        m = ol.evaluate(model.lp, o, r.primal, r.row_dual)
        assert max(m["primal_residual"], m["dual_residual"], m["relative_gap"]) <= 1e-6
    else:
        assert r.status == ("infeasible" if name in KNOWN["infeasible"] else "unbounded")
