"""Writes tests/golden/highs_objectives.json: optimal objectives of seeded
instances from an independent LP solver, HiGHS (scipy.optimize.linprog,
method="highs"), the cross-check the reference's own tooling uses
(pdlp_engine/tools/reference_check.py:82).  The instances are the
repository's seeded generators, so the fixture pins the oracle and the
engine to an optimum neither of them computed.

Run from the repository root: python tests/golden/make_highs.py
"""
from __future__ import annotations

import json
import math
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "tests"))

import problems  # noqa: E402
from optimisation_solver_amd import box_lp, zipf_lp  # noqa: E402

OUT = Path(__file__).resolve().parent / "highs_objectives.json"


def instances():
    """name -> Lp; every generator call is seeded."""
    return {
        "c0_box_2000x5000": box_lp(2000, 5000, 5),
        "box_1500x3750_s9": box_lp(1500, 3750, 5, seed=9),
        "zipf_3000x7000_s5": zipf_lp(3000, 7000, seed=5),
        "zipf_6000x15000_s12": zipf_lp(6000, 15000, seed=12),
        "random_feasible_300x420_s21": problems.random_feasible(300, 420, 21),
        "random_feasible_260x360_s33": problems.random_feasible(260, 360, 33),
    }


def status_instances():
    """name -> Lp whose status (infeasible / unbounded) HiGHS classifies."""
    return {
        "infeasible_rows": problems.infeasible_rows(),
        "infeasible_equalities": problems.infeasible_equalities(),
        "unbounded": problems.unbounded(),
    }


def highs(lp) -> dict:
    """min c x + offset s.t. rlo <= A x <= rhi, lo <= x <= hi, as linprog's
    A_ub x <= b_ub (finite upper and negated finite lower row bounds of
    inequality rows) and A_eq x = b_eq (rows with rlo == rhi)."""
    import scipy.sparse as sp
    from scipy.optimize import linprog

    a = sp.csr_matrix((lp.rv, lp.ci, lp.rp), shape=(lp.m, lp.n))
    eq = np.isfinite(lp.rlo) & (lp.rlo == lp.rhi)
    up = ~eq & np.isfinite(lp.rhi)
    lo = ~eq & np.isfinite(lp.rlo)
    a_ub = sp.vstack([a[up], -a[lo]]).tocsr()
    b_ub = np.concatenate([lp.rhi[up], -lp.rlo[lo]])
    bounds = [(None if math.isinf(l) else l, None if math.isinf(h) else h) for l, h in zip(lp.lo, lp.hi)]
    r = linprog(lp.c, A_ub=a_ub if a_ub.shape[0] else None, b_ub=b_ub if a_ub.shape[0] else None,
                A_eq=a[eq] if eq.any() else None, b_eq=lp.rlo[eq] if eq.any() else None,
                bounds=bounds, method="highs")
    return {"status": int(r.status), "message": str(r.message),
            "objective": float(r.fun) + float(lp.offset) if r.status == 0 else None,
            "m": int(lp.m), "n": int(lp.n), "nnz": int(lp.rp[-1]), "sha256": lp.sha256()}


def main() -> None:
    import scipy

    out = {"solver": f"HiGHS via scipy {scipy.__version__} linprog(method='highs'), default tolerances",
           "instances": {}}
    for name, lp in instances().items():
        out["instances"][name] = highs(lp)
        print(name, out["instances"][name]["objective"], out["instances"][name]["message"], flush=True)
    # linprog status: 2 = infeasible, 3 = unbounded
    out["statuses"] = {name: highs(lp) for name, lp in status_instances().items()}
    for name, h in out["statuses"].items():
        print(name, h["status"], h["message"], flush=True)
    OUT.write_text(json.dumps(out, indent=1, sort_keys=True) + "\n")


if __name__ == "__main__":
    main()
