"""Parametric multi-period LP family for the learned warm start.

Base LP: block-angular, `periods` blocks of (bm rows, bn columns, per_col
nonzeros per column, rows uniform in the block) plus `link` linking rows of
`link_nnz` nonzeros over two adjacent periods (csrc/lpgen.cpp, makeBlock);
60% equality rows, 40% <= rows with slack U[0,1]; box bounds [0, 20]; a
planted point x0 ~ U[0, 10].

Instance parameter theta in R^dim, theta ~ N(0, I).  Fixed seeded Gaussian
maps P, Q (n x dim) turn it into (defaults rhs_scale 0.1, cost_sigma 0.05)
  x0(theta) = clip(x0 * (1 + rhs_scale tanh(P theta / sqrt(dim))), lo, hi)
  rows:  equality rows b = A x0(theta); <= rows u = A x0(theta) + slack0
  costs: c(theta) = c0 * exp(cost_sigma Q theta / sqrt(dim))   (lognormal)
so every instance is feasible (x0(theta) is inside the box and satisfies every
row) and bounded (box-bounded columns): the right-hand sides move by up to
+-10% through the planted point rather than by scaling b directly, which could
make equality rows inconsistent.  All instances share A.

Seeds: base LP BASE_SEED + 4; maps BASE_SEED + 4 + 7919; instance i draws theta
from seed BASE_SEED + 4000 + i.
"""
from __future__ import annotations

import math

import numpy as np
import scipy.sparse as sp

from ..lp import BASE_SEED, Lp, block_lp

INF = math.inf


class BlockFamily:
    def __init__(self, periods: int = 24, bm: int = 2000, bn: int = 5000, per_col: int = 5, link: int = 4000,
                 link_nnz: int = 12, dim: int = 256, seed: int = BASE_SEED + 4, rhs_scale: float = 0.1,
                 cost_sigma: float = 0.05):
        self.base = block_lp(periods, bm, bn, per_col, link, link_nnz, seed)
        b = self.base
        self.dim = dim
        self.rhs_scale = rhs_scale
        self.cost_sigma = cost_sigma
        self.A = sp.csr_matrix((b.rv, b.ci, b.rp), shape=(b.m, b.n))
        g = np.random.Generator(np.random.PCG64(seed + 7919))
        self.P = g.standard_normal((b.n, dim), dtype=np.float32)
        self.Q = g.standard_normal((b.n, dim), dtype=np.float32)
        ax0 = self.A @ b.x0
        self.eq = b.rlo == b.rhi
        self.slack = np.where(self.eq, 0.0, b.rhi - ax0)
        self.theta_seed = BASE_SEED + 4000

    @property
    def m(self) -> int:
        return self.base.m

    @property
    def n(self) -> int:
        return self.base.n

    def theta(self, i: int) -> np.ndarray:
        return np.random.Generator(np.random.PCG64(self.theta_seed + i)).standard_normal(self.dim)

    def thetas(self, start: int, count: int) -> np.ndarray:
        return np.stack([self.theta(start + i) for i in range(count)])

    def vectors(self, theta: np.ndarray) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
        """(c, row_lower, row_upper) of the instance with parameter theta."""
        b = self.base
        s = 1.0 / math.sqrt(self.dim)
        t32 = theta.astype(np.float32)
        px = (self.P @ t32).astype(np.float64) * s
        qx = (self.Q @ t32).astype(np.float64) * s
        x = np.clip(b.x0 * (1.0 + self.rhs_scale * np.tanh(px)), b.lo, b.hi)
        ax = self.A @ x
        rlo = np.where(self.eq, ax, -INF)
        rhi = ax + self.slack
        c = b.c * np.exp(self.cost_sigma * qx)
        return c, rlo, rhi

    def instance(self, theta: np.ndarray, name: str = "") -> Lp:
        c, rlo, rhi = self.vectors(theta)
        return self.base.with_costs_and_rows(c, rlo, rhi, name=name or "block_family_instance")
