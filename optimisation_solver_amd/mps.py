"""MPS reader for LPs (fixed or free format, whitespace separated fields).

Builds the engine's ``Lp`` (the CompiledLp layout of
pdlp_engine/include/pdlp/compiled_lp.h:17-63) from the standard MPS
sections; it is the entry point of the reference's MPS -> CompiledLp path
(src/mps/mps_reader.cpp, then adapter::toCompiledLp in
src/adapter/pdlp_adapter.cpp) restricted to LPs:

  NAME, OBJSENSE (MIN/MAX, on its own line or after the keyword),
  ROWS (N = objective, the first N row; E, L, G),
  COLUMNS (integer MARKER lines are rejected: LP only),
  RHS (a value on the objective row is the negated objective constant),
  RANGES (E: [b, b + |R|] if R > 0 else [b - |R|, b]; L: [b - |R|, b];
          G: [b, b + |R|]),
  BOUNDS (UP, LO, FX, FR, MI, PL; default [0, +inf)), ENDATA.

A maximisation is returned as the equivalent minimisation (costs and
constant negated) with ``maximize=True`` so callers can flip the objective
back.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from .lp import Lp

INF = math.inf


@dataclass
class MpsModel:
    lp: Lp
    maximize: bool
    row_names: list[str]
    col_names: list[str]

    def objective(self, minimised_value: float) -> float:
        """The model's objective value from the minimisation's value."""
        return -minimised_value if self.maximize else minimised_value


class MpsError(ValueError):
    pass


def read_mps(path_or_text: str, is_text: bool = False) -> MpsModel:
    text = path_or_text if is_text else open(path_or_text).read()
    section = None
    maximize = False
    obj_row = None
    row_kind: dict[str, str] = {}
    row_order: list[str] = []
    cols: dict[str, int] = {}
    col_order: list[str] = []
    entries: list[tuple[str, int, float]] = []
    cost: dict[int, float] = {}
    rhs: dict[str, float] = {}
    rng: dict[str, float] = {}
    offset = 0.0
    bounds: dict[int, list[float]] = {}

    def col_index(name: str) -> int:
        if name not in cols:
            cols[name] = len(col_order)
            col_order.append(name)
        return cols[name]

    for raw in text.splitlines():
        line = raw.rstrip()
        if not line or line.lstrip().startswith("*"):
            continue
        head = line.split()
        if not line[0].isspace():   # section header
            key = head[0].upper()
            if key == "NAME":
                section = "NAME"
            elif key == "OBJSENSE":
                section = "OBJSENSE"
                if len(head) > 1:
                    maximize = head[1].upper() in ("MAX", "MAXIMIZE")
            elif key in ("ROWS", "COLUMNS", "RHS", "RANGES", "BOUNDS"):
                section = key
            elif key == "ENDATA":
                break
            else:
                raise MpsError(f"unsupported MPS section {key}")
            continue
        f = head
        if section == "OBJSENSE":
            maximize = f[0].upper() in ("MAX", "MAXIMIZE")
        elif section == "ROWS":
            kind, name = f[0].upper(), f[1]
            if kind == "N":
                if obj_row is None:
                    obj_row = name
                continue   # further free rows are ignored
            if kind not in ("E", "L", "G"):
                raise MpsError(f"bad row type {kind}")
            row_kind[name] = kind
            row_order.append(name)
        elif section == "COLUMNS":
            if len(f) >= 3 and f[1].upper() == "'MARKER'":
                raise MpsError("integer columns are not supported (LP only)")
            j = col_index(f[0])
            pairs = f[1:]
            for q in range(0, len(pairs) - 1, 2):
                r, v = pairs[q], float(pairs[q + 1])
                if r == obj_row:
                    cost[j] = cost.get(j, 0.0) + v
                elif r in row_kind:
                    entries.append((r, j, v))
                else:
                    raise MpsError(f"unknown row {r}")
        elif section in ("RHS", "RANGES"):
            pairs = f[1:] if len(f) % 2 == 1 else f   # the set name may be omitted
            for q in range(0, len(pairs) - 1, 2):
                r, v = pairs[q], float(pairs[q + 1])
                if section == "RHS":
                    if r == obj_row:
                        offset = -v
                    else:
                        rhs[r] = v
                else:
                    rng[r] = v
        elif section == "BOUNDS":
            kind = f[0].upper()
            name = f[2] if len(f) >= 3 and f[2] in cols else f[1]
            val = float(f[-1]) if kind not in ("FR", "MI", "PL", "BV") else None
            j = col_index(name)
            b = bounds.setdefault(j, [0.0, INF])
            if kind == "UP":
                b[1] = val
            elif kind == "LO":
                b[0] = val
            elif kind == "FX":
                b[0] = b[1] = val
            elif kind == "FR":
                b[0], b[1] = -INF, INF
            elif kind == "MI":
                b[0] = -INF
            elif kind == "PL":
                b[1] = INF
            else:
                raise MpsError(f"unsupported bound type {kind} (LP only)")
    m, n = len(row_order), len(col_order)
    ridx = {r: i for i, r in enumerate(row_order)}
    rlo, rhi = np.empty(m), np.empty(m)
    for i, r in enumerate(row_order):
        b = rhs.get(r, 0.0)
        k = row_kind[r]
        lo, hi = (b, b) if k == "E" else ((-INF, b) if k == "L" else (b, INF))
        if r in rng:
            R = rng[r]
            if k == "E":
                lo, hi = (b, b + abs(R)) if R > 0 else (b - abs(R), b)
            elif k == "L":
                lo = b - abs(R)
            else:
                hi = b + abs(R)
        rlo[i], rhi[i] = lo, hi
    c = np.zeros(n)
    for j, v in cost.items():
        c[j] = v
    lo, hi = np.zeros(n), np.full(n, INF)
    for j, (a, b) in bounds.items():
        lo[j], hi[j] = a, b
    sign = -1.0 if maximize else 1.0
    rows = [ridx[r] for r, _, _ in entries]
    colj = [j for _, j, _ in entries]
    vals = [v for _, _, v in entries]
    lp = Lp.from_triplets(m, n, rows, colj, vals, sign * c, lo, hi, rlo, rhi, sign * offset)
    return MpsModel(lp, maximize, row_order, col_order)
