"""Batched solves: many LPs of uneven sizes in one device run.  Each instance
must match its own single solve against the oracle (same status, 1e-6
relative objective, oracle-scored KKT within tolerance); invalid instances in
a batch are reported without disturbing the others."""
import numpy as np
import pytest

import oracle_lib as ol
import problems
from optimisation_solver_amd import Engine, box_lp, default_options

pytestmark = pytest.mark.gpu


def _opts(tol, **kw):
    o = default_options(primal_tolerance=tol, dual_tolerance=tol, gap_tolerance=tol, thread_count=1)
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def _mixed_batch():
    cf = problems.closed_form()
    lps = [box_lp(300 + 37 * i, int(2.5 * (300 + 37 * i)), 5, seed=900 + i) for i in range(6)]
    lps += [cf["equality_row"][0], cf["degenerate"][0], problems.infeasible_rows(), problems.unbounded(),
            problems.random_feasible(200, 300, 5)]
    return lps


def test_batch_matches_individual_oracle_solves():
    lps = _mixed_batch()
    o = _opts(1e-6, iteration_limit=50000, termination_check_frequency=50, polishing_iterations=500)
    print(f"batch of {len(lps)} instances", flush=True)
    with Engine() as e:
        e.load_batch(lps)
        res = e.solve_batch(o)
    for lp, r in zip(lps, res):
        cpu = ol.solve(lp, o)
        print(f"  {lp.name or lp.m}: gpu {r.status} {r.iterations} its obj {r.primal_objective:.10g}; "
              f"cpu {cpu.status_name} {cpu.iterations} its obj {cpu.primal_objective:.10g}", flush=True)
        assert r.status == cpu.status_name
        if r.status == "optimal":
            assert abs(r.primal_objective - cpu.primal_objective) / (1 + abs(cpu.primal_objective)) <= 1e-6
            m = ol.evaluate(lp, o, r.primal, r.row_dual)
            assert m["primal_residual"] <= 1e-6 and m["dual_residual"] <= 1e-6 and m["relative_gap"] <= 1e-6


def test_batch_equals_single_engine_solve():
    lps = [box_lp(500, 1250, 5, seed=41), box_lp(800, 2000, 5, seed=42)]
    o = _opts(1e-6)
    with Engine() as e:
        e.load_batch(lps)
        res = e.solve_batch(o)
    for lp, r in zip(lps, res):
        with Engine() as e1:
            e1.load(lp)
            s = e1.solve(o)
        # same kernels, same block partition per instance: identical trajectory
        assert (s.status, s.iterations, s.step_trials) == (r.status, r.iterations, r.step_trials)
        assert np.array_equal(s.primal, r.primal) and np.array_equal(s.row_dual, r.row_dual)


def test_batch_with_invalid_member_and_warm_starts():
    good = box_lp(400, 1000, 5, seed=7)
    bad = problems.lp(1, 1, [(0, 0, 1.0)], [1], [0], [1], [5], [1])
    o = _opts(1e-6)
    with Engine() as e:
        e.load_batch([good, bad, good])
        res = e.solve_batch(o)
        assert res[1].status == "invalid_problem" and res[1].status_message == "Invalid row bounds"
        assert res[0].status == res[2].status == "optimal"
        warm = e.solve_batch(o, warm_x=[res[0].primal, None, None], warm_y=[res[0].row_dual, None, None])
    assert warm[0].status == "optimal" and warm[0].iterations <= res[0].iterations
    assert warm[2].iterations == res[2].iterations


def test_batch_fp32_mixed_precision_reaches_fp64_tolerance():
    """FP32 SpMV/iterates with FP64 checks and an FP64 hand-off: every
    instance still reaches the FP64 tolerance as scored by the oracle."""
    lps = [box_lp(400 + 150 * i, int(2.5 * (400 + 150 * i)), 5, seed=3100 + i) for i in range(5)]
    o = _opts(1e-6, precision=1)
    with Engine() as e:
        e.load_batch(lps, precision=1)
        res = e.solve_batch(o)
    for lp, r in zip(lps, res):
        cpu = ol.solve(lp, _opts(1e-6))
        print(f"  fp32 batch: {r.status} its={r.iterations} fp32_its={r.diagnostics['fp32_iterations']} "
              f"obj={r.primal_objective:.10g} cpu obj={cpu.primal_objective:.10g}", flush=True)
        assert r.status == "optimal"
        assert r.diagnostics["fp32_iterations"] > 0
        assert abs(r.primal_objective - cpu.primal_objective) / (1 + abs(cpu.primal_objective)) <= 1e-6
        m = ol.evaluate(lp, o, r.primal, r.row_dual)
        assert max(m["primal_residual"], m["dual_residual"], m["relative_gap"]) <= 1e-6


@pytest.mark.parametrize("handoff", [1.0, 1e9])
def test_batch_mixed_fp32_values_fp64_iterates(handoff):
    """FP32 matrix values with FP64 iterates (PDLPX_MIXED): every instance
    reaches the FP64 tolerance as scored by the oracle on the FP64 matrix.
    handoff=1e9 switches every instance to FP64 values in place at its first
    check, which exercises the in-place hand-off (no restart)."""
    from optimisation_solver_amd import zipf_lp
    lps = [box_lp(400 + 150 * i, int(2.5 * (400 + 150 * i)), 5, seed=3200 + i) for i in range(4)]
    lps.append(zipf_lp(3000, 7500, False, seed=3300))
    o = _opts(1e-6, precision=2, fp32_handoff_factor=handoff)
    with Engine() as e:
        e.load_batch(lps, precision=2)
        res = e.solve_batch(o)
    for lp, r in zip(lps, res):
        cpu = ol.solve(lp, _opts(1e-6))
        print(f"  mixed batch (handoff {handoff:g}) {lp.name}: {r.status} its={r.iterations} "
              f"fp32_value_its={r.diagnostics['fp32_iterations']} obj={r.primal_objective:.10g} "
              f"cpu {cpu.status_name} its={cpu.iterations} obj={cpu.primal_objective:.10g}", flush=True)
        assert r.status == cpu.status_name == "optimal"
        assert r.diagnostics["fp32_iterations"] > 0
        if handoff > 1:
            assert r.diagnostics["fp32_iterations"] < r.iterations
        assert abs(r.primal_objective - cpu.primal_objective) / (1 + abs(cpu.primal_objective)) <= 1e-6
        m = ol.evaluate(lp, o, r.primal, r.row_dual)
        assert max(m["primal_residual"], m["dual_residual"], m["relative_gap"]) <= 1e-6


@pytest.mark.parametrize("precision", [0, 2])
def test_compacted_batch_equals_single_solves(precision):
    """Instances of very different sizes finish at different times, so the
    trial kernels' plan is cut down to the running instances (and widened
    again for the final FP64 polish in mixed precision).  Every member must
    still equal its single solve bitwise."""
    sizes = [300, 2400, 450, 3000, 600, 1800, 350, 900]
    lps = [box_lp(m, int(2.5 * m), 5, seed=700 + i) for i, m in enumerate(sizes)]
    o = _opts(1e-6, precision=precision)
    print(f"batch of {len(lps)}, precision {precision}", flush=True)
    with Engine() as e:
        e.load_batch(lps, precision)
        res = e.solve_batch(o)
    for lp, r in zip(lps, res):
        with Engine() as e1:
            e1.load(lp, precision)
            s = e1.solve(o)
        print(f"  m={lp.m}: {r.status} {r.iterations} its (single {s.iterations})", flush=True)
        assert (s.status, s.iterations, s.step_trials) == (r.status, r.iterations, r.step_trials)
        assert np.array_equal(s.primal, r.primal) and np.array_equal(s.row_dual, r.row_dual)


@pytest.mark.parametrize("precision", [0, 2])
def test_batch_with_structural_corner_cases(precision):
    """Instances with an empty row and column, a fixed variable, no rows at
    all and a free row, batched with ordinary LPs: every member equals its
    single solve bitwise and reaches the optimum fixed by inspection."""
    st = problems.structural()
    lps = [box_lp(400, 1000, 5, seed=81)] + [v[0] for v in st.values()] + [box_lp(900, 2250, 5, seed=82)]
    o = _opts(1e-8, precision=precision, iteration_limit=100000, termination_check_frequency=50,
              polishing_iterations=2000)
    print(f"batch of {len(lps)} with corner cases, precision {precision}", flush=True)
    with Engine() as e:
        e.load_batch(lps, precision)
        res = e.solve_batch(o)
    for lp, r in zip(lps, res):
        with Engine() as e1:
            e1.load(lp, precision)
            s = e1.solve(o)
        print(f"  m={lp.m} n={lp.n}: {r.status} {r.iterations} its (single {s.iterations})", flush=True)
        assert (s.status, s.iterations, s.step_trials) == (r.status, r.iterations, r.step_trials)
        assert np.array_equal(s.primal, r.primal) and np.array_equal(s.row_dual, r.row_dual)
    for (lp, obj), r in zip(st.values(), res[1:-1]):
        assert r.status == "optimal" and abs(r.primal_objective - obj) <= 1e-6 * (1 + abs(obj))


def test_update_vectors_equals_fresh_load():
    """pdlpx_update_vectors: new costs and row bounds on a loaded matrix solve
    exactly like a fresh load of the modified LP (bitwise), in a batch member
    too; an invalid update is rejected and leaves the instance unchanged."""
    from optimisation_solver_amd.solver import EngineError
    base = box_lp(600, 1500, 5, seed=61)
    g = np.random.default_rng(5)
    c2 = base.c * (1 + 0.05 * g.standard_normal(base.n))
    rlo2 = np.where(np.isfinite(base.rlo), base.rlo * 1.02, base.rlo)
    rhi2 = np.where(np.isfinite(base.rhi), base.rhi * 1.02, base.rhi)
    rhi2 = np.maximum(rhi2, rlo2)
    moved = base.with_costs_and_rows(c2, rlo2, rhi2)
    other = box_lp(300, 750, 5, seed=62)
    o = _opts(1e-6)
    with Engine() as e:
        e.load(moved)
        fresh = e.solve(o)
    with Engine() as e:
        e.load_batch([other, base])
        before = e.solve_batch(o)
        e.update_vectors(1, c2, rlo2, rhi2)
        after = e.solve_batch(o)
        bad = c2.copy()
        bad[3] = np.nan
        with pytest.raises(EngineError):
            e.update_vectors(1, bad, None, None)
        with pytest.raises(EngineError):
            e.update_vectors(1, None, rhi2 + 1.0, rlo2)    # crossed row bounds
        with pytest.raises(EngineError):
            e.update_vectors(2, c2, None, None)             # no such instance
        again = e.solve_batch(o)
    print(f"fresh {fresh.iterations} its, updated {after[1].iterations} its", flush=True)
    assert (after[1].status, after[1].iterations) == (fresh.status, fresh.iterations)
    assert np.array_equal(after[1].primal, fresh.primal) and np.array_equal(after[1].row_dual, fresh.row_dual)
    for a, b in ((before[0], after[0]), (after[0], again[0]), (after[1], again[1])):
        assert a.iterations == b.iterations and np.array_equal(a.primal, b.primal)
    assert not np.array_equal(before[1].primal, after[1].primal)
