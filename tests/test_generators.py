"""Instance generators: determinism (SHA-256 fixtures), the planted point's
feasibility, the families' structure, and the deterministic log/exp."""
import json
import math
from pathlib import Path

import numpy as np
import pytest

from optimisation_solver_amd import _native, batch_sizes, block_lp, box_lp, zipf_lp

GOLD = Path(__file__).resolve().parent / "golden"


def test_hashes_match_fixtures():
    h = json.loads((GOLD / "instance_hashes.json").read_text())
    assert box_lp(2000, 5000, 5).sha256() == h["c0_box_2000x5000"]
    assert zipf_lp(3000, 7000, seed=5).sha256() == h["zipf_3000x7000_s5"]
    assert zipf_lp(20000, 50000, seed=11).sha256() == h["zipf_20000x50000_s11"]


def _rows_times(lp, x):
    return np.array([lp.rv[lp.rp[i]:lp.rp[i + 1]] @ x[lp.ci[lp.rp[i]:lp.rp[i + 1]]] for i in range(lp.m)])


@pytest.mark.parametrize("maker", [lambda: box_lp(500, 1250, 5, 1), lambda: zipf_lp(800, 2000, seed=2),
                                   lambda: block_lp(3, 200, 500, 5, 100, 6, 3)])
def test_planted_point_is_feasible_and_structure_canonical(maker):
    lp = maker()
    assert np.all(np.diff(lp.rp) >= 0) and lp.rp[-1] == lp.nnz == lp.cp[-1]
    for i in range(lp.m):   # columns sorted inside rows
        assert np.all(np.diff(lp.ci[lp.rp[i]:lp.rp[i + 1]]) > 0)
    for j in range(lp.n):   # rows sorted inside columns
        assert np.all(np.diff(lp.ri[lp.cp[j]:lp.cp[j + 1]]) > 0)
    assert np.all(lp.rv != 0)
    ax = _rows_times(lp, lp.x0)
    assert np.all(ax >= lp.rlo - 1e-9 * (1 + np.abs(ax))) and np.all(ax <= lp.rhi + 1e-9 * (1 + np.abs(ax)))
    assert np.all(lp.x0 >= np.where(np.isfinite(lp.lo), lp.lo, -np.inf))
    eq = np.arange(lp.m) % 5 < 3
    assert np.all(lp.rlo[eq] == lp.rhi[eq]) and np.all(np.isinf(lp.rlo[~eq]))


def test_zipf_family_shape_and_dual_feasible_costs():
    lp = zipf_lp(4000, 10000, seed=7)
    deg = np.diff(lp.cp)
    assert deg.min() >= 1 and deg.max() <= 295 and 8 < deg.mean() < 13
    free = np.arange(lp.n) % 10 == 0
    assert np.all(np.isinf(lp.lo[free])) and np.all(lp.lo[~free] == 0) and np.all(lp.hi[~free] == 20)
    lit = zipf_lp(4000, 10000, literal=True, seed=7)
    assert np.diff(lit.cp).max() > 295


def test_deterministic_log_exp():
    lib = _native.lpgen()
    for x in [1e-300, 1e-5, 0.3, 0.7071, 1.0, 1.5, 2.0, 10.0, 1e5, 1e300]:
        assert abs(lib.lpgen_log(x) - math.log(x)) <= 4e-16 * max(1.0, abs(math.log(x)))
    for x in [-700.0, -20.0, -1.0, -1e-8, 0.0, 0.5, 1.0, 3.3, 50.0, 700.0]:
        assert abs(lib.lpgen_exp(x) - math.exp(x)) <= 4e-16 * math.exp(x)


def test_streams_distribution():
    lib = _native.lpgen()
    n = 200000
    u, z, k = np.zeros(n), np.zeros(n), np.zeros(n)
    lib.lpgen_stream(1, 0, n, u.ctypes.data)
    lib.lpgen_stream(1, 1, n, z.ctypes.data)
    lib.lpgen_stream(1, 2, n, k.ctypes.data)
    assert abs(u.mean() - 0.5) < 0.005 and u.min() >= 0 and u.max() < 1
    assert abs(z.mean()) < 0.01 and abs(z.std() - 1) < 0.01
    assert abs((k == 1).mean() - 1 / 2.6123753486854883) < 0.005 and k.max() <= 2000


def test_batch_sizes():
    s = batch_sizes(256)
    ms = np.array([a for a, _ in s])
    assert len(s) == 256 and ms.min() >= 500 and ms.max() <= 20000
    assert all(b == round(2.5 * a) for a, b in s)
