"""Shared workload builders for the tests (small sizes the oracle finishes in seconds)."""

from __future__ import annotations

import numpy as np

from egg_amd import grids
from egg_amd.smoothing import build_structured_context_from_block_maps, build_sweep_context


class Case:
    def __init__(self, grid, w_inv=None):
        self.grid = grid
        self.ctx = build_sweep_context(grid.blocks, grid.free, grid.d, w_inv=w_inv)
        self.bsc = build_structured_context_from_block_maps(grid.d, grid.blocks)

    @property
    def d(self):
        return self.grid.d

    @property
    def X(self):
        return self.grid.X

    def patches(self, O):
        return O.materialize_patches(self.ctx.wire, self.grid.blocks, self.grid.d)


def single(n=33, seed=1, amp=0.3):
    return Case(grids.single_block_2d(n, seed=seed, amp=amp))


def ogrid(n=33, m=17, tangle_seed=None):
    g = grids.ogrid_2d(n, m)
    if tangle_seed is not None:
        g = grids.tangle(g, seed=tangle_seed)
    return Case(g)


def cube(n=9, seed=1, amp=0.2):
    return Case(grids.single_block_3d(n, seed=seed, amp=amp))


def lr_pair(n=9, seed=3):
    """Two blocks sharing a face (conforming), the reference's L/R pair shape."""
    h = 1.0 / (n - 1)
    a = np.arange(n * n, dtype=np.int32).reshape(n, n)
    b = a[-1:, :].copy()
    rest = np.arange(n * n, n * n + (n - 1) * n, dtype=np.int32).reshape(n - 1, n)
    b = np.concatenate([b, rest], axis=0)
    N = n * n + (n - 1) * n
    X = np.zeros((N, 2))
    for blk, x0 in ((a, 0.0), (b, 1.0)):
        i, j = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
        X[blk.reshape(-1)] = np.stack([x0 + i * h, j * h], -1).reshape(-1, 2)
    free = np.ones(N, bool)
    bnd = (np.isclose(X[:, 0], 0) | np.isclose(X[:, 0], 2) | np.isclose(X[:, 1], 0) | np.isclose(X[:, 1], 1))
    free[bnd] = False
    rng = np.random.default_rng(seed)
    X[free] += rng.uniform(-0.2 * h, 0.2 * h, size=(int(free.sum()), 2))
    return Case(grids.Grid(2, [a, b], X, free))
