"""Plain-text instance and answer files, the interchange format of the
reference's validation tooling (pdlp_engine/tools/validate_against_reference.cpp:
153-189 writes them, tools/reference_check.py:20-42 reads them).

instance (.lp.txt):
    rows cols nnz
    objective_offset
    c[0..cols)            (one line; "inf" / "-inf" for infinities)
    var_lower[0..cols)
    var_upper[0..cols)
    row_lower[0..rows)
    row_upper[0..rows)
    row col value         (nnz lines, triplets)
answer (.answer.txt):
    status primal_objective primal_residual dual_residual relative_gap seconds
    x[0..cols)
Numbers are written with 17 significant digits, so a round trip is exact.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from .lp import Lp


def _fmt(v: float) -> str:
    if v == math.inf:
        return "inf"
    if v == -math.inf:
        return "-inf"
    return f"{v:.17g}"


def write_lp_txt(lp: Lp, path: str, triplets: tuple | None = None) -> None:
    """Writes `lp`; `triplets` (rows, cols, vals) keeps a caller's entry
    order, otherwise the CSR entries are written row by row."""
    if triplets is None:
        rows = np.repeat(np.arange(lp.m), np.diff(lp.rp))
        triplets = (rows, lp.ci, lp.rv)
    r, c, v = triplets
    with open(path, "w") as f:
        f.write(f"{lp.m} {lp.n} {len(v)}\n{lp.offset:.17g}\n")
        for vec in (lp.c, lp.lo, lp.hi, lp.rlo, lp.rhi):
            f.write(" ".join(_fmt(x) for x in vec) + " \n")
        for a, b, x in zip(r, c, v):
            f.write(f"{int(a)} {int(b)} {x:.17g}\n")


def read_lp_txt(path: str, name: str = "") -> Lp:
    with open(path) as f:
        tokens = f.read().split()
    pos = 0

    def take(k):
        nonlocal pos
        out = tokens[pos:pos + k]
        pos += k
        return out
    m, n, nnz = (int(t) for t in take(3))
    offset = float(take(1)[0])
    vec = lambda k: np.array([float(t) for t in take(k)])   # float() parses "inf"/"-inf"
    c, lo, hi, rlo, rhi = vec(n), vec(n), vec(n), vec(m), vec(m)
    trip = take(3 * nnz)
    rows = np.array(trip[0::3], dtype=np.int64)
    cols = np.array(trip[1::3], dtype=np.int64)
    vals = np.array([float(t) for t in trip[2::3]])
    return Lp.from_triplets(m, n, rows, cols, vals, c, lo, hi, rlo, rhi, offset, name=name or path)


@dataclass
class Answer:
    status: str
    primal_objective: float
    primal_residual: float
    dual_residual: float
    relative_gap: float
    seconds: float
    primal: np.ndarray


def write_answer(path: str, status: str, primal_objective: float, primal_residual: float, dual_residual: float,
                 relative_gap: float, seconds: float, primal) -> None:
    with open(path, "w") as f:
        f.write(f"{status} {primal_objective:.17g} {primal_residual:.17g} {dual_residual:.17g} "
                f"{relative_gap:.17g} {seconds:.17g}\n")
        f.write(" ".join(f"{x:.17g}" for x in primal) + " \n")


def read_answer(path: str) -> Answer:
    with open(path) as f:
        head = f.readline().split()
        x = np.array([float(t) for t in f.read().split()])
    return Answer(head[0], *(float(t) for t in head[1:6]), primal=x)
