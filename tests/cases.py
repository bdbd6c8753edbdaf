"""Shared workload builders for the tests (small sizes the oracle finishes in seconds)."""

from __future__ import annotations

import numpy as np

from egg_amd import grids
from egg_amd.smoothing import build_structured_context_from_block_maps, build_sweep_context


class Case:
    def __init__(self, grid, w_inv=None, weight=None, entities=None, curvature=None, directional=None):
        self.grid = grid
        self.ctx = build_sweep_context(grid.blocks, grid.free, grid.d, w_inv=w_inv, weight=weight,
                                       entities=entities, curvature=curvature, directional=directional)
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


def block_grid(blocks_spec, fixed_outer=True):
    """Blocks given as (x0, y0, nx, ny, h) rectangles laid out conformingly on
    a common lattice; coincident nodes merge; the outer boundary is fixed."""
    from egg_amd.grids import Grid, _merge_fast

    coords = []
    for x0, y0, nx, ny, h in blocks_spec:
        i, j = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
        coords.append(np.stack([x0 + i * h, y0 + j * h], -1))
    ids, X = _merge_fast(coords, 2)
    # a node is free when at least three cells surround it (interior and
    # conforming-interface nodes); block faces on the domain boundary are fixed
    cells = np.zeros(X.shape[0], np.int64)
    for b in ids:
        for sl in ((slice(0, -1), slice(0, -1)), (slice(1, None), slice(0, -1)),
                   (slice(0, -1), slice(1, None)), (slice(1, None), slice(1, None))):
            np.add.at(cells, b[sl].reshape(-1), 1)
    return Grid(2, ids, X, cells >= 3)


def bump(grid, amp):
    """Smooth low-frequency displacement vanishing on the boundary (the error
    component plain Jacobi damps slowest; reference tests/smoothing/
    test_fas.py:51-61)."""
    X = grid.X.copy()
    lo = X.min(0)
    span = X.max(0) - lo
    s = np.sin(np.pi * (X - lo) / span)
    b = s[:, 0] * s[:, 1]
    X[:, 0] += amp * b
    X[:, 1] += 0.7 * amp * b
    from egg_amd.grids import Grid

    return Grid(2, grid.blocks, X, grid.free)


def fas_single(res=16, amp=0.08):
    return Case(bump(block_grid([(0.0, 0.0, res + 1, res + 1, 4.0 / res)]), amp))


def fas_two_blocks(res=16, amp=0.15):
    h = 4.0 / res
    return Case(bump(block_grid([(0.0, 0.0, res + 1, res + 1, h), (4.0, 0.0, res + 1, res + 1, h)]), amp))


def halo_pair(n0=4, n1=3):
    """Two n0 x n1 blocks sharing a face along axis 0, seeded with the
    continuous field node (i, j) -> (x0 + i, j): block 0 has x0 = 0, block 1
    x0 = n0 - 1 (its row i = 0 is block 0's row n0 - 1). The known-answer
    layout of the reference's tests/cpp/test_halo_device.cpp:28-76."""
    a = np.arange(n0 * n1, dtype=np.int32).reshape(n0, n1)
    b = np.concatenate([a[-1:, :], np.arange(n0 * n1, n0 * n1 + (n0 - 1) * n1, dtype=np.int32).reshape(n0 - 1, n1)])
    N = n0 * n1 + (n0 - 1) * n1
    X = np.zeros((N, 2))
    for blk, x0 in ((a, 0.0), (b, float(n0 - 1))):
        i, j = np.meshgrid(np.arange(n0), np.arange(n1), indexing="ij")
        X[blk.reshape(-1)] = np.stack([x0 + i, j], -1).reshape(-1, 2)
    free = np.zeros(N, bool)
    free[a[1:, 1:-1].reshape(-1)] = True
    free[b[1:-1, 1:-1].reshape(-1)] = True
    return Case(grids.Grid(2, [a, b], X, free))


def halo_pair_expected(n0=4, n1=3):
    """(block, padded index) -> expected coordinate after one exchange: the
    ghost past block 0's high face holds block 1's first interior layer
    (x = n0, the extrapolation of block 0's own field), the ghost before block
    1's low face block 0's layer n0 - 2 (x = n0 - 2)."""
    out = {}
    for j in range(n1):
        out[(0, (n0 + 1, j + 1))] = (float(n0), float(j))
        out[(1, (0, j + 1))] = (float(n0 - 2), float(j))
    return out


def star(k=5, n=9, layers=0, seed=7, amp=0.2):
    """k quadrilateral blocks around a point (a regular k-gon cut by its
    centre-to-edge-midpoint spokes; block i spans spokes i and i+1 and holds
    vertex i+1), n x n nodes each, bilinear. The centre is a singular node
    of valence k: its patch has 4k occurrences (20 for k = 5, more than one
    round of a 16-lane group). layers > 0 extrudes it into a hex grid of
    `layers` + 1 node planes; the centre line then has 2k hexes around each
    interior node (8 * 2k occurrences: 80 for k = 5, two rounds of 64).
    Interior and spoke nodes are free (the k-gon's edges and the end planes
    are fixed) and moved by U(+-amp h)."""
    from egg_amd.grids import Grid, _merge_fast, corner_dets

    th = 2 * np.pi * np.arange(k + 1) / k
    V = np.stack([np.cos(th), np.sin(th)], -1)
    M = 0.5 * (V[:-1] + V[1:])  # M[i]: midpoint of edge (V_i, V_{i+1})
    M = np.concatenate([M, M[:1]])
    C = np.zeros(2)
    a = np.linspace(0.0, 1.0, n)
    A, B = np.meshgrid(a, a, indexing="ij")
    coords = []
    for i in range(k):
        P = ((1 - A) * (1 - B))[..., None] * C + (A * (1 - B))[..., None] * M[i] \
            + (A * B)[..., None] * V[i + 1] + ((1 - A) * B)[..., None] * M[i + 1]
        if layers:
            z = np.linspace(0.0, 0.5, layers + 1)
            P = np.concatenate([np.repeat(P[:, :, None, :], layers + 1, 2),
                                np.broadcast_to(z[None, None, :, None], (n, n, layers + 1, 1))], -1)
        coords.append(P)
    d = 3 if layers else 2
    ids, X = _merge_fast(coords, d)
    fixed = np.zeros(X.shape[0], bool)
    for b in ids:
        fixed[b[-1].reshape(-1)] = True       # a = 1: edge M_i -> V_{i+1}
        fixed[b[:, -1].reshape(-1)] = True    # b = 1: edge M_{i+1} -> V_{i+1}
        if layers:
            fixed[b[:, :, 0].reshape(-1)] = True
            fixed[b[:, :, -1].reshape(-1)] = True
    free = ~fixed
    h = 0.5 / (n - 1)
    rng = np.random.default_rng(seed)
    for _ in range(100):
        Xp = X.copy()
        Xp[free] += rng.uniform(-amp * h, amp * h, size=(int(free.sum()), d))
        if corner_dets(ids, Xp, d).min() > 0:
            return Case(Grid(d, ids, Xp, free))
    raise RuntimeError("could not draw a valid perturbation")
