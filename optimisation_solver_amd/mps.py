"""MPS reader for LPs (fixed or free format, whitespace separated fields).

Builds the engine's ``Lp`` (the CompiledLp layout of
pdlp_engine/include/pdlp/compiled_lp.h:17-63) from the standard MPS
sections; it is the entry point of the reference's MPS -> CompiledLp path
(src/mps/mps_reader.cpp, then adapter::toCompiledLp in
src/adapter/pdlp_adapter.cpp) restricted to LPs, with the reference
reader's conventions (src/mps/mps_reader.cpp:195-460, 609-632):

  NAME, OBJSENSE (MIN/MAX, on its own line or after the keyword),
  ROWS (the first N row is the objective; later N rows are free rows whose
        coefficients are discarded, with a warning; E, L, G; a duplicate row
        name is an error),
  COLUMNS (integer MARKER lines are rejected: LP only; a coefficient on an
           undeclared row is ignored with a warning),
  RHS and RANGES (a record with an odd field count names its vector and only
        the first named vector is used, later ones are ignored with a
        warning; an even count omits the name and always applies;
        an RHS value on the objective row is the negated objective
        constant; a RANGES entry on it is ignored;
        RANGES: E: [b, b + |R|] if R > 0 else [b - |R|, b]; L: [b - |R|, b];
        G: [b, b + |R|]),
  BOUNDS (UP, LO, FX, FR, MI, PL; default [0, +inf); the bound-set name may
          be omitted; a negative UP on a column whose lower bound is still
          the implicit 0 opens that bound to -inf, with a warning;
          BV/LI/UI are rejected: LP only), ENDATA.

A maximisation is returned as the equivalent minimisation (costs and
constant negated) with ``maximize=True`` so callers can flip the objective
back.  ``warnings`` lists what was ignored or adjusted.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from .lp import Lp

INF = math.inf


@dataclass
class MpsModel:
    lp: Lp
    maximize: bool
    row_names: list[str]
    col_names: list[str]
    name: str = ""
    warnings: list[str] = field(default_factory=list)

    def objective(self, minimised_value: float) -> float:
        """The model's objective value from the minimisation's value."""
        return -minimised_value if self.maximize else minimised_value


class MpsError(ValueError):
    pass


def read_mps(path_or_text: str, is_text: bool = False) -> MpsModel:
    text = path_or_text if is_text else open(path_or_text).read()
    section = None
    name = ""
    maximize = False
    obj_row = None
    free_rows: set[str] = set()
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
    lower_set: set[int] = set()            # lower bound given by LO, FX, FR or MI
    vector = {"RHS": None, "RANGES": None}   # the first named vector of each
    warnings: list[str] = []

    def col_index(cname: str) -> int:
        if cname not in cols:
            cols[cname] = len(col_order)
            col_order.append(cname)
        return cols[cname]

    for raw in text.splitlines():
        line = raw.rstrip()
        if not line or line.lstrip().startswith("*"):
            continue
        head = line.split()
        if not line[0].isspace():   # section header
            key = head[0].upper()
            if key == "NAME":
                section = "NAME"
                name = head[1] if len(head) > 1 else ""
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
            kind, rname = f[0].upper(), f[1]
            if kind == "N":
                if obj_row is None:
                    obj_row = rname
                else:
                    free_rows.add(rname)
                    warnings.append(f"free row '{rname}' after the objective row was ignored")
                continue
            if kind not in ("E", "L", "G"):
                raise MpsError(f"bad row type {kind}")
            if rname in row_kind:
                raise MpsError(f"duplicate row name '{rname}'")
            row_kind[rname] = kind
            row_order.append(rname)
        elif section == "COLUMNS":
            # a marker record is NAME 'MARKER' 'INTORG'|'INTEND' (quoted or not);
            # a row or column that is merely called MARKER is ordinary data
            if (len(f) >= 3 and f[1].strip("'\"").upper() == "MARKER"
                    and f[2].strip("'\"").upper() in ("INTORG", "INTEND")):
                raise MpsError("integer columns are not supported (LP only)")
            j = col_index(f[0])
            pairs = f[1:]
            for q in range(0, len(pairs) - 1, 2):
                r, v = pairs[q], float(pairs[q + 1])
                if r == obj_row:
                    cost[j] = cost.get(j, 0.0) + v
                elif r in row_kind:
                    entries.append((r, j, v))
                elif r not in free_rows:
                    warnings.append(f"COLUMNS entry for '{f[0]}' names unknown row '{r}'; ignored")
        elif section in ("RHS", "RANGES"):
            named = len(f) % 2 == 1   # (row, value) pairs, after an optional vector name
            if named:
                if vector[section] is None:
                    vector[section] = f[0]
                if f[0] != vector[section]:
                    warnings.append(f"{section} vector '{f[0]}' ignored; using first vector '{vector[section]}'")
                    continue
            pairs = f[1:] if named else f
            for q in range(0, len(pairs) - 1, 2):
                r, v = pairs[q], float(pairs[q + 1])
                if r == obj_row:
                    if section == "RHS":
                        offset = -v
                    else:
                        warnings.append("RANGES entry on the objective row was ignored")
                elif r in row_kind:
                    (rhs if section == "RHS" else rng)[r] = v
                elif section == "RANGES" or r not in free_rows:
                    warnings.append(f"{section} entry names unknown row '{r}'; ignored")
        elif section == "BOUNDS":
            kind = f[0].upper()
            if kind in ("BV", "LI", "UI"):
                raise MpsError(f"unsupported bound type {kind} (LP only)")
            valued = kind in ("UP", "LO", "FX")
            if kind not in ("UP", "LO", "FX", "FR", "MI", "PL"):
                raise MpsError(f"unsupported bound type {kind}")
            # kind [set] column [value]: the set name is present when the
            # record has one field more than the bound type needs
            need = 3 if valued else 2
            if len(f) not in (need, need + 1):
                raise MpsError(f"malformed BOUNDS record: {line.strip()}")
            cname = f[2] if len(f) == need + 1 else f[1]
            val = float(f[-1]) if valued else None
            j = col_index(cname)
            b = bounds.setdefault(j, [0.0, INF])
            if kind == "UP":
                b[1] = val
                if val < 0.0 and j not in lower_set and b[0] == 0.0:
                    b[0] = -INF
                    warnings.append(f"negative UP bound on '{cname}' opened its lower bound to -infinity")
            elif kind == "LO":
                b[0] = val
                lower_set.add(j)
            elif kind == "FX":
                b[0] = b[1] = val
                lower_set.add(j)
            elif kind == "FR":
                b[0], b[1] = -INF, INF
                lower_set.add(j)
            elif kind == "MI":
                b[0] = -INF
                lower_set.add(j)
            else:   # PL
                b[1] = INF
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
    return MpsModel(lp, maximize, row_order, col_order, name, warnings)
