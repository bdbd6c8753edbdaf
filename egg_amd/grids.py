"""Synthetic multi-block structured grids for the benchmark workloads.

Every generator returns a :class:`Grid`: per-block global-id arrays (the
``block_global_dof`` of the structured tables), positions ``X`` (N, d) in
global order and the moving-DOF mask. Blocks are oriented so that every
cell's corner Jacobian has positive determinant before any perturbation.

Workloads (seeds are arguments; the benchmark fixes them):

* :func:`single_block_2d` — one n x n block on the unit square, interior
  nodes moved by U(-amp h, amp h) per coordinate, re-drawn until every
  corner sample keeps det A > 0; boundary fixed.
* :func:`ogrid_2d` — 5-block O-grid: a central n x n block whose four edges
  are 90-degree arcs of the circle r = 0.5 (built by transfinite
  interpolation) inside four n x m ring blocks reaching the square
  [-2, 2]^2. The circle is an internal interface of free nodes; the outer
  square is fixed; the four central-block corners are valence-3 nodes.
* :func:`tangle` — displace a seeded fraction of the free nodes by
  ``factor`` times their local edge length in a seeded direction.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass
class Grid:
    d: int
    blocks: list  # per-block int32 global-id arrays
    X: np.ndarray  # (N, d)
    free: np.ndarray  # (N,) bool

    @property
    def n_nodes(self) -> int:
        return int(self.X.shape[0])

    @property
    def n_free(self) -> int:
        return int(self.free.sum())


def _det(c: list) -> np.ndarray:
    """Closed-form det of the matrices with columns c[k] (arrays [..., d])."""
    if len(c) == 2:
        return c[0][..., 0] * c[1][..., 1] - c[1][..., 0] * c[0][..., 1]
    return (c[0][..., 0] * (c[1][..., 1] * c[2][..., 2] - c[2][..., 1] * c[1][..., 2])
            - c[1][..., 0] * (c[0][..., 1] * c[2][..., 2] - c[2][..., 1] * c[0][..., 2])
            + c[2][..., 0] * (c[0][..., 1] * c[1][..., 2] - c[1][..., 1] * c[0][..., 2]))


def corner_dets(blocks, X: np.ndarray, d: int) -> np.ndarray:
    """det A of every (cell, corner) sample of the blocks, in the wire's
    sample order: blocks in order, cells row-major, corner b with axis 0 the
    most significant bit; column k of A is s_k (X[b ^ bit k] - X[b])."""
    out = []
    nc = 1 << d
    for ids in blocks:
        ids = np.asarray(ids)
        if min(ids.shape) < 2:
            continue
        P = X[ids]  # (n0, .., n_{d-1}, d)
        corner = [P[tuple(slice((b >> (d - 1 - k)) & 1, ids.shape[k] - 1 + ((b >> (d - 1 - k)) & 1))
                          for k in range(d))] for b in range(nc)]
        dets = []
        for b in range(nc):
            cols = []
            for k in range(d):
                bit = 1 << (d - 1 - k)
                cols.append((corner[b ^ bit] - corner[b]) * (-1.0 if b & bit else 1.0))
            dets.append(_det(cols).reshape(-1))
        out.append(np.stack(dets, 1).reshape(-1))
    return np.concatenate(out) if out else np.zeros(0)


def min_det(grid: Grid) -> float:
    return float(corner_dets(grid.blocks, grid.X, grid.d).min())


def _merge(block_coords: list, d: int, scale: float = 1e8):
    """Global ids by merging coincident block nodes (rounded coordinates)."""
    keys = {}
    blocks = []
    pts = []
    for P in block_coords:
        shp = P.shape[:-1]
        flat = P.reshape(-1, d)
        ids = np.empty(flat.shape[0], np.int32)
        rk = np.round(flat * scale).astype(np.int64)
        for n in range(flat.shape[0]):
            key = tuple(rk[n])
            g = keys.get(key)
            if g is None:
                g = len(pts)
                keys[key] = g
                pts.append(flat[n])
            ids[n] = g
        blocks.append(ids.reshape(shp))
    return blocks, np.asarray(pts, np.float64)


def single_block_2d(n: int = 129, *, seed: int = 0, amp: float = 0.3) -> Grid:
    """n x n nodes on [0, 1]^2, node (i, j) at (i h, j h); interior perturbed
    by U(-amp h, amp h) per coordinate (re-drawn until min det A > 0)."""
    h = 1.0 / (n - 1)
    ids = np.arange(n * n, dtype=np.int32).reshape(n, n)
    i, j = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    X0 = np.stack([i * h, j * h], -1).reshape(-1, 2)
    free = np.zeros((n, n), bool)
    free[1:-1, 1:-1] = True
    free = free.reshape(-1)
    rng = np.random.default_rng(seed)
    for _ in range(100):
        X = X0.copy()
        X[free] += rng.uniform(-amp * h, amp * h, size=(int(free.sum()), 2))
        if corner_dets([ids], X, 2).min() > 0:
            return Grid(2, [ids], X, free)
    raise RuntimeError("could not draw a valid perturbation")


def wavy_block_2d(n: int = 13, *, wave: float = 0.35, perturb: float = 0.04, seed: int = 11) -> Grid:
    """n x n nodes on [0, 4]^2 with the top edge raised to y = 4 + wave
    sin(pi x / 4), the interior filled by transfinite interpolation of the
    boundary and then moved by perturb * N(0, 1) per coordinate (the control-
    net smoother's test block, reference tests/smoothing/test_control_ref.py:
    41-86); the boundary is fixed."""
    t = np.linspace(0.0, 1.0, n)
    P = np.zeros((n, n, 2))
    P[:, 0] = np.stack([4 * t, np.zeros(n)], -1)  # bottom
    P[:, -1] = np.stack([4 * t, 4 + wave * np.sin(np.pi * t)], -1)  # top (x = 4 t)
    P[0, :] = np.stack([np.zeros(n), 4 * t], -1)  # left
    P[-1, :] = np.stack([np.full(n, 4.0), 4 * t], -1)  # right
    u, v = t[:, None, None], t[None, :, None]
    P[1:-1, 1:-1] = ((1 - u) * P[:1] + u * P[-1:] + (1 - v) * P[:, :1] + v * P[:, -1:]
                     - ((1 - u) * (1 - v) * P[0, 0] + u * (1 - v) * P[-1, 0] + (1 - u) * v * P[0, -1]
                        + u * v * P[-1, -1]))[1:-1, 1:-1]
    P[1:-1, 1:-1] += perturb * np.random.default_rng(seed).standard_normal((n - 2, n - 2, 2))
    ids = np.arange(n * n, dtype=np.int32).reshape(n, n)
    free = np.zeros((n, n), bool)
    free[1:-1, 1:-1] = True
    return Grid(2, [ids], P.reshape(-1, 2), free.reshape(-1))


def single_block_3d(n: int = 17, *, seed: int = 0, amp: float = 0.2) -> Grid:
    """n^3 nodes on [0, 1]^3, interior perturbed by U(-amp h, amp h)."""
    h = 1.0 / (n - 1)
    ids = np.arange(n**3, dtype=np.int32).reshape(n, n, n)
    g = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    X0 = np.stack([c * h for c in g], -1).reshape(-1, 3)
    free = np.zeros((n, n, n), bool)
    free[1:-1, 1:-1, 1:-1] = True
    free = free.reshape(-1)
    rng = np.random.default_rng(seed)
    for _ in range(100):
        X = X0.copy()
        X[free] += rng.uniform(-amp * h, amp * h, size=(int(free.sum()), 3))
        if corner_dets([ids], X, 3).min() > 0:
            return Grid(3, [ids], X, free)
    raise RuntimeError("could not draw a valid perturbation")


def _ang(deg):
    return np.deg2rad(deg)


def ogrid_2d(n: int = 129, m: int = 65, *, r: float = 0.5, half: float = 2.0) -> Grid:
    """5-block O-grid: central n x n disc block + four n x m ring blocks."""
    u = np.linspace(0.0, 1.0, n)

    def circ(theta):
        return np.stack([r * np.cos(theta), r * np.sin(theta)], -1)

    # central block, node (i, j) -> (u_i, v_j); edges are arcs of the circle
    U, V = np.meshgrid(u, u, indexing="ij")
    Bu = circ(_ang(225.0 + 90.0 * u))  # v = 0
    Tu = circ(_ang(135.0 - 90.0 * u))  # v = 1
    Lv = circ(_ang(225.0 - 90.0 * u))  # u = 0
    Rv = circ(_ang(-45.0 + 90.0 * u))  # u = 1
    P00, P10, P01, P11 = Bu[0], Bu[-1], Tu[0], Tu[-1]
    C = ((1 - V)[..., None] * Bu[:, None, :] + V[..., None] * Tu[:, None, :]
         + (1 - U)[..., None] * Lv[None, :, :] + U[..., None] * Rv[None, :, :]
         - ((1 - U) * (1 - V))[..., None] * P00 - (U * (1 - V))[..., None] * P10
         - ((1 - U) * V)[..., None] * P01 - (U * V)[..., None] * P11)
    # exact boundary positions
    C[:, 0] = Bu
    C[:, -1] = Tu
    C[0, :] = Lv
    C[-1, :] = Rv
    blocks = [C]
    # rings: axis 0 tangential (clockwise), axis 1 radial outward
    w = np.linspace(0.0, 1.0, m)
    sides = [
        (315.0, lambda t: np.stack([half - 2 * half * t, -half * np.ones_like(t)], -1)),  # bottom
        (225.0, lambda t: np.stack([-half * np.ones_like(t), -half + 2 * half * t], -1)),  # left
        (135.0, lambda t: np.stack([-half + 2 * half * t, half * np.ones_like(t)], -1)),  # top
        (45.0, lambda t: np.stack([half * np.ones_like(t), half - 2 * half * t], -1)),  # right
    ]
    for a0, side in sides:
        inner = circ(_ang(a0 - 90.0 * u))
        outer = side(u)
        P = (1 - w)[None, :, None] * inner[:, None, :] + w[None, :, None] * outer[:, None, :]
        blocks.append(P)
    ids, X = _merge(blocks, 2)
    free = np.ones(X.shape[0], bool)
    free[np.isclose(np.abs(X[:, 0]), half) | np.isclose(np.abs(X[:, 1]), half)] = False
    return Grid(2, ids, X, free)


def _smoothstep(t: float) -> float:
    t = min(max(t, 0.0), 1.0)
    return t * t * (3.0 - 2.0 * t)


def layer_spacing(k: int, *, first_height: float, growth: float, interior_spacing: float,
                  n_layers: int = 8) -> float:
    """Wall-normal target spacing of off-wall layer ``k``: geometric growth
    from ``first_height``, capped at ``interior_spacing`` and, past
    ``n_layers``, blended to it over another ``n_layers`` with a Hermite
    smoothstep (the profile of the reference's
    BoundaryLayerTarget.normal_spacing, egg/smoothing/targets.py:196-212)."""
    s = min(first_height * growth**k, interior_spacing)
    if k > n_layers:
        w = _smoothstep((k - n_layers) / max(n_layers, 1))
        s = (1.0 - w) * s + w * interior_spacing
    return s


def boundary_layer_w_inv(grid: Grid, walls: dict, *, first_height: float, growth: float,
                         tangential_spacing: float, n_layers: int = 8) -> np.ndarray:
    """Per energy-stencil sample inverse target W^-1, shape (ns, d, d), of
    wall clustering (reference egg/smoothing/targets.py:59-212,
    BoundaryLayerTarget): a cell of block b in ``walls`` ({b: (wall_axis,
    wall_side)}) at off-wall layer k asks for spacing ``layer_spacing(k)``
    along its wall axis and ``tangential_spacing`` along the others; every
    other block is isotropic at ``tangential_spacing``. The target is given
    in the block's logical frame: the reference's W is that diagonal turned
    by the wall frame (tangent, normal), a proper rotation on the right of
    A W^-1 that the shape metrics do not see. Samples are ordered as the
    flat wire's energy stencil (blocks in order, cells row-major, 2^d
    corners per cell)."""
    d = grid.d
    nc = 1 << d
    rows = []
    for b, ids in enumerate(grid.blocks):
        shp = np.asarray(ids).shape
        if any(n < 2 for n in shp):
            continue
        cshape = tuple(n - 1 for n in shp)
        spacing = np.full(cshape + (d,), float(tangential_spacing))
        if b in walls:
            ax, side = walls[b]
            nk = cshape[ax]
            prof = np.array([layer_spacing(k, first_height=first_height, growth=growth,
                                           interior_spacing=tangential_spacing, n_layers=n_layers)
                             for k in range(nk)])
            k = np.arange(nk) if side == 0 else nk - 1 - np.arange(nk)
            shape = [1] * d
            shape[ax] = nk
            spacing[..., ax] = np.broadcast_to(prof[k].reshape(shape), cshape)
        winv = np.zeros(cshape + (d, d))
        for a in range(d):
            winv[..., a, a] = 1.0 / spacing[..., a]
        rows.append(np.repeat(winv.reshape(-1, d, d), nc, axis=0))
    return np.concatenate(rows) if rows else np.zeros((0, d, d))


def ogrid_boundary_layer_w_inv(grid: Grid, n: int, *, r: float = 0.5, first_ratio: float = 0.3,
                               growth: float = 1.2, n_layers: int = 8) -> np.ndarray:
    """The boundary-layer target of an ``ogrid_2d(n, m)`` grid: its four ring
    blocks cluster towards the circle (wall axis 1, low side), first layer
    ``first_ratio`` of the tangential spacing at the circle (a quarter arc
    over n - 1 cells)."""
    s_t = 0.5 * np.pi * r / (n - 1)
    return boundary_layer_w_inv(grid, {b: (1, 0) for b in range(1, 5)}, first_height=first_ratio * s_t,
                                growth=growth, tangential_spacing=s_t, n_layers=n_layers)


def annulus_2d(n: int = 65, m: int = 33, *, r: float = 0.5, half: float = 2.0, seed: int = 0,
               amp: float = 0.2, aspect: float = 1.0) -> tuple[Grid, dict]:
    """The four ring blocks of ``ogrid_2d(n, m)`` alone: the circle becomes a
    wall whose nodes slide on it (returned as ``{node: ("circle", (0, 0, r))}``
    for the wire's ``entities``), the outer square stays fixed. The circle
    nodes start rotated by up to ``amp`` of their angular spacing and the
    ring-interior nodes moved by up to ``amp`` of the local spacing (seeded),
    so the sweeps have both to even out. ``aspect`` != 1 scales every y
    coordinate by it: the wall is then the ellipse of radii (r, aspect r)."""
    g = ogrid_2d(n, m, r=r, half=half)
    used = np.unique(np.concatenate([np.asarray(b).reshape(-1) for b in g.blocks[1:]]))
    relabel = np.full(g.n_nodes, -1, np.int64)
    relabel[used] = np.arange(used.size)
    blocks = [relabel[np.asarray(b)].astype(np.int32) for b in g.blocks[1:]]
    X = g.X[used].copy()
    rad = np.linalg.norm(X, axis=1)
    on_circle = np.isclose(rad, r)
    fixed = np.isclose(np.abs(X[:, 0]), half) | np.isclose(np.abs(X[:, 1]), half)
    free = ~fixed
    rng = np.random.default_rng(seed)
    dth = 0.5 * np.pi / (n - 1)
    th = np.arctan2(X[on_circle, 1], X[on_circle, 0]) + amp * dth * rng.uniform(-1, 1, int(on_circle.sum()))
    X[on_circle] = r * np.stack([np.cos(th), np.sin(th)], -1)
    inner = free & ~on_circle
    h = local_edge_length(Grid(2, blocks, g.X[used], free))
    X[inner] += amp * h[inner, None] * rng.uniform(-1, 1, (int(inner.sum()), 2))
    if aspect != 1.0:
        X[:, 1] *= aspect
        ents = {int(i): ("ellipse", (0.0, 0.0, float(r), float(aspect * r))) for i in np.flatnonzero(on_circle)}
    else:
        ents = {int(i): ("circle", (0.0, 0.0, float(r))) for i in np.flatnonzero(on_circle)}
    return Grid(2, blocks, X, free), ents


def interface_c2_windows(grid: Grid, weight: float = 1.0) -> tuple[np.ndarray, np.ndarray]:
    """Curvature windows across the block interfaces of a 2D grid: every node
    on the faces of exactly two blocks (not a face corner) gets the windows
    (a2, a1, n, b1) and (a1, n, b1, b2), with a1, a2 and b1, b2 its first two
    inward neighbours across each block's face: the grid line through the
    interface should not kink there (the reference's block-interface C2
    term). Returns (nodes (W, 4) global ids, weight (W,))."""
    inward: dict[int, list] = {}
    for ids in grid.blocks:
        ids = np.asarray(ids)
        nx, ny = ids.shape
        if nx < 3 or ny < 3:
            continue
        for ax, n in ((0, nx), (1, ny)):
            other = ny if ax == 0 else nx
            for side in (0, 1):
                i0, i1, i2 = (0, 1, 2) if side == 0 else (n - 1, n - 2, n - 3)
                for t in range(1, other - 1):
                    idx = (lambda i: (i, t)) if ax == 0 else (lambda i: (t, i))
                    inward.setdefault(int(ids[idx(i0)]), []).append((int(ids[idx(i1)]), int(ids[idx(i2)])))
    rows = []
    for node, lst in inward.items():
        if len(lst) != 2:
            continue
        (a1, a2), (b1, b2) = lst
        rows += [(a2, a1, node, b1), (a1, node, b1, b2)]
    nodes = np.asarray(rows, np.int32).reshape(-1, 4)
    return nodes, np.full(nodes.shape[0], float(weight))


def wall_directional_samples(grid: Grid, weight: float = 1.0, eps: float = 1e-12):
    """Directional samples on a single block (the reference's three kinds,
    egg/smoothing/directional.py:166-301, placed by hand): the first row of
    nodes above the j = 0 wall as a chain of parallel segments with the
    wall's normal (0, 1) as reference; every third node of that row a stem to
    the next row up with the through axis (1, 0); the middle row as line
    triples. A 3D block takes them in its middle k-plane (references with a
    zero third component). Returns a ``DirectionalSamples``."""
    from egg_amd.smoothing.directional import KIND_LINE, KIND_PARALLEL, KIND_STEM, DirectionalSamples

    ids = np.asarray(grid.blocks[0])
    if grid.d == 3:
        ids = ids[:, :, ids.shape[2] // 2]
    n = ids.shape[0]
    rows = []
    for i in range(n - 1):
        rows.append((int(ids[i, 1]), int(ids[i + 1, 1]), -1, KIND_PARALLEL, (0.0, 1.0)))
    for i in range(2, n - 2, 3):
        rows.append((int(ids[i, 2]), int(ids[i, 1]), -1, KIND_STEM, (1.0, 0.0)))
    m = ids.shape[1] // 2
    for i in range(1, n - 1, 2):
        rows.append((int(ids[i - 1, m]), int(ids[i, m]), int(ids[i + 1, m]), KIND_LINE, (0.0, 0.0)))
    nodes = np.full((len(rows), 4), -1, np.int32)
    nodes[:, :3] = [r[:3] for r in rows]
    k = np.asarray([r[3] for r in rows], np.int32)
    ref = np.asarray([r[4] + (0.0,) * (grid.d - 2) for r in rows], np.float64)
    return DirectionalSamples(nodes, k, ref, np.zeros_like(ref), np.full(len(rows), float(weight)), eps)


def local_edge_length(grid: Grid) -> np.ndarray:
    """Shortest incident block edge per node."""
    h = np.full(grid.n_nodes, np.inf)
    for ids in grid.blocks:
        ids = np.asarray(ids)
        for ax in range(grid.d):
            a = np.take(ids, np.arange(ids.shape[ax] - 1), axis=ax).reshape(-1)
            b = np.take(ids, np.arange(1, ids.shape[ax]), axis=ax).reshape(-1)
            L = np.linalg.norm(grid.X[a] - grid.X[b], axis=1)
            np.minimum.at(h, a, L)
            np.minimum.at(h, b, L)
    return h


def tangle(grid: Grid, *, frac: float = 0.02, factor: float = 1.5, seed: int = 0) -> Grid:
    """Displace ``frac`` of the free nodes by ``factor`` x local edge length in
    a seeded random direction (re-drawn until some det A <= 0)."""
    rng = np.random.default_rng(seed)
    free_ids = np.flatnonzero(grid.free)
    h = local_edge_length(grid)
    for _ in range(100):
        pick = rng.choice(free_ids, size=max(1, int(round(frac * free_ids.size))), replace=False)
        dirs = rng.normal(size=(pick.size, grid.d))
        dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
        X = grid.X.copy()
        X[pick] += factor * h[pick, None] * dirs
        if corner_dets(grid.blocks, X, grid.d).min() <= 0:
            return Grid(grid.d, grid.blocks, X, grid.free.copy())
    raise RuntimeError("tangling did not invert any cell")


def cubed_sphere_3d(n: int = 65, *, r: float = 1.0, half: float = 2.0) -> Grid:
    """7-block hex grid: a central n^3 block filling the ball of radius r
    (the cube [-1, 1]^3 scaled radially onto the ball, so its boundary is
    the sphere) inside six n^3 shell blocks from the sphere to the cube
    [-half, half]^3. The sphere is an internal interface; the outer cube is
    fixed; the central block's 12 edges and 8 corners are singular."""
    u = np.linspace(-1.0, 1.0, n)
    A, B, C = np.meshgrid(u, u, u, indexing="ij")
    xi = np.stack([A, B, C], -1)
    inf = np.abs(xi).max(-1, keepdims=True)
    l2 = np.linalg.norm(xi, axis=-1, keepdims=True)
    with np.errstate(invalid="ignore", divide="ignore"):
        core = np.where(l2 > 0, r * inf * xi / l2, 0.0)
    blocks = [core]
    w = np.linspace(0.0, 1.0, n)
    for axis in range(3):
        for sgn in (-1.0, 1.0):
            a, b = np.meshgrid(u, u, indexing="ij")
            face = np.zeros(a.shape + (3,))
            others = [k for k in range(3) if k != axis]
            face[..., axis] = sgn
            face[..., others[0]] = a
            face[..., others[1]] = b
            inner = r * face / np.linalg.norm(face, axis=-1, keepdims=True)
            outer = half * face
            P = (1 - w)[None, None, :, None] * inner[:, :, None, :] + w[None, None, :, None] * outer[:, :, None, :]
            # orient (tangential a, tangential b, radial) to a positive Jacobian
            J = np.linalg.det(np.stack([P[1, 0, 0] - P[0, 0, 0], P[0, 1, 0] - P[0, 0, 0], P[0, 0, 1] - P[0, 0, 0]], 1))
            if J < 0:
                P = P[::-1]
            blocks.append(P)
    ids, X = _merge_fast(blocks, 3)
    free = np.ones(X.shape[0], bool)
    free[(np.abs(X) >= half - 1e-9).any(axis=1)] = False
    return Grid(3, ids, X, free)


def spherical_shell_3d(n: int = 17, *, r: float = 1.0, half: float = 2.0, seed: int = 0, amp: float = 0.1,
                       axis=(1.0, 0.37, 0.61)) -> tuple[Grid, dict]:
    """The six shell blocks of ``cubed_sphere_3d(n)`` alone: the sphere of
    radius r becomes a wall whose nodes slide on it (returned as
    ``{node: ("sphere", (c, r, ax, ay))}`` for the wire's ``entities``, the
    sphere's polar axis along ``axis`` so that no node sits on a pole of its
    chart), the outer cube stays fixed. The sphere nodes start rotated by a
    small seeded rotation's worth (up to ``amp`` of their spacing), the
    others moved by up to ``amp`` of the local spacing."""
    g = cubed_sphere_3d(n, r=r, half=half)
    used = np.unique(np.concatenate([np.asarray(b).reshape(-1) for b in g.blocks[1:]]))
    relabel = np.full(g.n_nodes, -1, np.int64)
    relabel[used] = np.arange(used.size)
    blocks = [relabel[np.asarray(b)].astype(np.int32) for b in g.blocks[1:]]
    X = g.X[used].copy()
    on = np.isclose(np.linalg.norm(X, axis=1), r)
    free = ~(np.abs(X) >= half - 1e-9).any(axis=1)
    rng = np.random.default_rng(seed)
    h = local_edge_length(Grid(3, blocks, g.X[used], free))
    step = amp * h[on, None] * rng.uniform(-1, 1, (int(on.sum()), 3))
    P = X[on] + step
    X[on] = r * P / np.linalg.norm(P, axis=1, keepdims=True)
    inner = free & ~on
    X[inner] += amp * h[inner, None] * rng.uniform(-1, 1, (int(inner.sum()), 3))
    az = np.asarray(axis, float) / np.linalg.norm(axis)
    ax = np.cross(az, (0.0, 0.0, 1.0))
    ax /= np.linalg.norm(ax)
    ay = np.cross(az, ax)
    rec = (0.0, 0.0, 0.0, float(r), *ax, *ay)
    return Grid(3, blocks, X, free), {int(i): ("sphere", rec) for i in np.flatnonzero(on)}


def _merge_fast(block_coords: list, d: int, scale: float = 1e8):
    """Vectorized coincident-node merge (rounded coordinates)."""
    flat = [P.reshape(-1, d) for P in block_coords]
    allp = np.concatenate(flat)
    keys = np.round(allp * scale).astype(np.int64)
    _, first, inv = np.unique(keys, axis=0, return_index=True, return_inverse=True)
    # number global ids in order of first appearance
    order = np.argsort(first, kind="stable")
    rank = np.empty_like(order)
    rank[order] = np.arange(order.size)
    gid = rank[inv.reshape(-1)].astype(np.int32)
    X = allp[np.sort(first)]
    ids = []
    o = 0
    for P in block_coords:
        m = int(np.prod(P.shape[:-1]))
        ids.append(gid[o:o + m].reshape(P.shape[:-1]))
        o += m
    return ids, X


def strip_2d(sides: list[int], ny: int, *, x0: float = 0.0, y0: float = 0.0, h: float | None = None) -> Grid:
    """Conforming chain of 2D blocks side by side along x: block k has
    sides[k] x ny nodes (axis 0 = x), consecutive blocks share a whole face.
    Uniform spacing h (default 1/(ny-1)); the outer boundary is fixed."""
    h = 1.0 / (ny - 1) if h is None else h
    # global node (column c, row j) has id c ny + j: the numbering of a merge
    # that names nodes in order of first appearance (block 0 row-major, then
    # each next block's columns past the shared one)
    cols = 1 + sum(n - 1 for n in sides)
    first = np.concatenate([[0], np.cumsum([n - 1 for n in sides])[:-1]]).astype(np.int64)
    grid_ids = np.arange(cols * ny, dtype=np.int64).reshape(cols, ny)
    ids = [grid_ids[c:c + n] for c, n in zip(first, sides)]
    xcol = np.empty(cols)
    xs = x0
    for c0, n in zip(first, sides):
        xcol[c0:c0 + n] = xs + np.arange(n) * h
        xs += (n - 1) * h
    X = np.stack(np.meshgrid(xcol, y0 + np.arange(ny) * h, indexing="ij"), -1).reshape(-1, 2)
    ids = [b.astype(np.int32) for b in ids]
    x_lo, x_hi = x0, xs
    y_lo, y_hi = y0, y0 + (ny - 1) * h
    free = ~(np.isclose(X[:, 0], x_lo) | np.isclose(X[:, 0], x_hi) | np.isclose(X[:, 1], y_lo)
             | np.isclose(X[:, 1], y_hi))
    return Grid(2, ids, X, free)


def batched_2d(n_grids: int = 64, *, seed: int = 0, min_blocks: int = 3, max_blocks: int = 9,
               sides=tuple(range(33, 514, 8)), frac: float = 0.02, factor: float = 1.5,
               jitter: float = 0.0) -> tuple[Grid, list]:
    """Many independent multi-block grids in one Grid (disjoint global ids):
    grid g is a conforming strip of 3..9 blocks whose sides are drawn from
    {33, 41, ..., 513} (n = 1 mod 8, so three multigrid levels exist), each
    2 % tangled by 1.5 x the local edge length (frac = 0: not tangled; free
    nodes then move by U(+-jitter h) instead, which folds no cell for
    jitter < 0.25). Returns the union grid and the per-grid (first block,
    block count, first node, node count)."""
    rng = np.random.default_rng(seed)
    blocks, Xs, frees, index = [], [], [], []
    node0 = 0
    for gi in range(n_grids):
        nb = int(rng.integers(min_blocks, max_blocks + 1))
        ny = int(rng.choice(sides))
        nx = [int(rng.choice(sides)) for _ in range(nb)]
        g = strip_2d(nx, ny, x0=0.0, y0=0.0)
        if frac > 0:
            g = tangle(g, frac=frac, factor=factor, seed=int(rng.integers(1 << 30)))
        elif jitter > 0:
            h = 1.0 / (ny - 1)
            X = g.X.copy()
            X[g.free] += rng.uniform(-jitter * h, jitter * h, size=(int(g.free.sum()), 2))
            g = Grid(2, g.blocks, X, g.free)
        index.append((len(blocks), nb, node0, g.n_nodes))
        blocks += [b + node0 for b in g.blocks]
        Xs.append(g.X)
        frees.append(g.free)
        node0 += g.n_nodes
    return Grid(2, [b.astype(np.int32) for b in blocks], np.concatenate(Xs), np.concatenate(frees)), index
