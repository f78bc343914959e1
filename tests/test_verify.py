"""The independent verifier (optimisation_solver_amd/verify.py) on points the
oracle produces and on planted mistakes (made here, on the CPU)."""
import numpy as np

import oracle_lib as ol
import problems
from optimisation_solver_amd import box_lp, default_options, zipf_lp
from optimisation_solver_amd.verify import check_solution


def _opts(tol=1e-8):
    return default_options(primal_tolerance=tol, dual_tolerance=tol, gap_tolerance=tol, thread_count=1)


def test_oracle_optimum_is_verified():
    for lp in (box_lp(300, 750, 5, seed=5), problems.closed_form()["equality_row"][0], zipf_lp(400, 1000, seed=8)):
        r = ol.solve(lp, _opts())
        assert r.status_name == "optimal"
        rep = check_solution(lp, r.x, r.y)
        assert rep["verdict"] == "optimal_verified", rep
        assert abs(rep["objective_recomputed"] - r.primal_objective) <= 1e-9 * (1 + abs(r.primal_objective))


def test_planted_mistakes_are_caught():
    lp = box_lp(300, 750, 5, seed=6)
    r = ol.solve(lp, _opts())
    x, y = r.x.copy(), r.y.copy()
    bad = x.copy()
    bad[np.argmax(np.abs(x))] += 1.0                       # moves a row off its bounds
    assert check_solution(lp, bad, y)["verdict"] == "infeasible_point"
    assert check_solution(lp, x, -y)["verdict"] == "feasible"   # wrong dual sign: no optimality proof
    assert check_solution(lp, x)["verdict"] == "feasible"       # no duals: feasibility only
    nan = x.copy()
    nan[0] = np.nan
    assert check_solution(lp, nan, y)["verdict"] == "nonfinite"
    assert check_solution(lp, x[:-1], y)["verdict"] == "malformed"
