"""Independent verification of a returned LP point against the original
problem: nothing here uses the solver's own residuals or status.

The frozen tolerances and scalings are those of the reference's verifier
(benchmarks/lib/verify.py:13-33, benchmarks/README.md):

    row i         |violation_i| <= 1e-6 + 1e-8 * max(1, |l_i|, |u_i|, sum_j |a_ij x_j|)
    variable j    |violation_j| <= 1e-6 + 1e-8 * max(1, |lb_j|, |ub_j|, |x_j|)
    optimality    gap_norm = |p - d| / (1 + |p| + |d|) <= 1e-6

with p = c.x + offset recomputed here.  The dual objective d is recomputed
from the row duals in the PDLP sign convention (stationarity c + A^T y = r,
y the negative of the shadow price), after projecting r and y onto the cones
the bounds allow (an infinite lower bound needs r <= 0, an infinite upper
bound r >= 0; likewise y for the rows): d = offset + sum_j inf_{x in [l, u]}
r_j x_j - sum_i sup_{a in [l_r, u_r]} y_i a.  The projection drift, over
1 + ||c||, is the normalised dual residual and must also be within 1e-6 for
"optimal_verified".  The verdict is reported beside, never merged with, the
solver's status:

    optimal_verified  feasible, gap_norm and dual residual within tolerance
    feasible          feasible, optimality not shown
    infeasible_point  a row or bound is violated beyond its tolerance
    nonfinite         NaN or infinity in the point
    malformed         wrong length
"""
from __future__ import annotations

import math

import numpy as np

from .lp import Lp

FEAS_ATOL = 1e-6
FEAS_RTOL = 1e-8
OPT_NORM_TOL = 1e-6
INF = math.inf


def _row_products(lp: Lp, x: np.ndarray):
    rows = np.repeat(np.arange(lp.m), np.diff(lp.rp))
    prod = lp.rv * x[lp.ci]
    act = np.bincount(rows, weights=prod, minlength=lp.m)
    mag = np.bincount(rows, weights=np.abs(prod), minlength=lp.m)
    return act, mag


def _finite_abs(v: np.ndarray) -> np.ndarray:
    return np.where(np.isfinite(v), np.abs(v), 0.0)


def check_solution(lp: Lp, x, y=None) -> dict:
    x = np.asarray(x, dtype=np.float64)
    rep = {"verdict": None, "tolerances": {"feasibility_absolute": FEAS_ATOL, "feasibility_relative": FEAS_RTOL,
                                           "optimality_normalised": OPT_NORM_TOL}}
    if x.shape != (lp.n,):
        rep["verdict"] = "malformed"
        return rep
    if not np.all(np.isfinite(x)):
        rep["verdict"] = "nonfinite"
        return rep
    act, mag = _row_products(lp, x)
    below = np.where(np.isfinite(lp.rlo), lp.rlo - act, 0.0)
    above = np.where(np.isfinite(lp.rhi), act - lp.rhi, 0.0)
    rviol = np.maximum(0.0, np.maximum(below, above))
    rscale = np.maximum.reduce([np.ones(lp.m), _finite_abs(lp.rlo), _finite_abs(lp.rhi), mag])
    rtol = FEAS_ATOL + FEAS_RTOL * rscale
    vb = np.where(np.isfinite(lp.lo), lp.lo - x, 0.0)
    va = np.where(np.isfinite(lp.hi), x - lp.hi, 0.0)
    cviol = np.maximum(0.0, np.maximum(vb, va))
    cscale = np.maximum.reduce([np.ones(lp.n), _finite_abs(lp.lo), _finite_abs(lp.hi), np.abs(x)])
    ctol = FEAS_ATOL + FEAS_RTOL * cscale
    p = float(lp.c @ x + lp.offset)
    rep["objective_recomputed"] = p
    rep["primal"] = {"max_row_violation_abs": float(rviol.max(initial=0.0)),
                     "max_row_violation_scaled": float((rviol / rtol).max(initial=0.0)),
                     "rows_violating": int(np.count_nonzero(rviol > rtol)),
                     "max_bound_violation_abs": float(cviol.max(initial=0.0)),
                     "variables_violating": int(np.count_nonzero(cviol > ctol))}
    feasible = rep["primal"]["rows_violating"] == 0 and rep["primal"]["variables_violating"] == 0
    rep["dual"] = {"available": y is not None}
    proven = False
    if y is not None:
        y = np.asarray(y, dtype=np.float64)
        cols = np.repeat(np.arange(lp.n), np.diff(lp.cp))
        r = lp.c + np.bincount(cols, weights=lp.cv * y[lp.ri], minlength=lp.n)
        # project onto the cones the bounds allow; the drift is dual infeasibility
        rp = np.where(np.isfinite(lp.lo), r, np.minimum(r, 0.0))
        rp = np.where(np.isfinite(lp.hi), rp, np.maximum(rp, 0.0))
        yp = np.where(np.isfinite(lp.rhi), y, np.minimum(y, 0.0))
        yp = np.where(np.isfinite(lp.rlo), yp, np.maximum(yp, 0.0))
        drift = math.sqrt(float(np.sum((r - rp) ** 2) + np.sum((y - yp) ** 2)))
        with np.errstate(invalid="ignore"):
            inf_terms = np.where(rp > 0, rp * lp.lo, np.where(rp < 0, rp * lp.hi, 0.0))
            sup_terms = np.where(yp > 0, yp * lp.rhi, np.where(yp < 0, yp * lp.rlo, 0.0))
        d = lp.offset + float(np.sum(inf_terms)) - float(np.sum(sup_terms))
        gn = abs(p - d) / (1 + abs(p) + abs(d)) if math.isfinite(d) else INF
        dres = drift / (1.0 + float(np.linalg.norm(lp.c)))
        rep["dual"].update({"dual_objective": d, "gap_abs": abs(p - d) if math.isfinite(d) else INF,
                            "gap_normalised": gn, "sign_drift_abs": drift, "dual_residual_normalised": dres})
        proven = gn <= OPT_NORM_TOL and dres <= OPT_NORM_TOL
    if not feasible:
        rep["verdict"] = "infeasible_point"
    else:
        rep["verdict"] = "optimal_verified" if proven else "feasible"
    return rep
