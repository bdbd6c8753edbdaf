"""Control-net front end of the reduced Gauss-Newton smoother: one block's
nodes as a fixed linear map of a tensor-product B-spline control net,
X = M C + b (reference egg/geometry/control_net.py:1-37, :330-447), the
fitted net, the Boolean-sum boundary offset b and the ``control`` wire the
session takes (reference egg/smoothing/control_ref.py:96-190,
egg/smoothing/control_backend.py:48-136).

Each axis k samples a clamped-uniform B-spline basis (degree p, n_ctrl[k]
functions, knots 0 (p+1 times), j/(m+1) for j = 1..m, 1 (p+1 times), m =
n_ctrl - p - 1) at the uniform node parameters u_i = i / (n_node - 1); M is
the Kronecker product of the per-axis matrices and is never formed. The
clamped ends make a boundary node depend on boundary controls only, so the
interior controls are the free ones and b (built from the boundary
mismatch) keeps every boundary node exactly where it was.
"""

from __future__ import annotations

from dataclasses import dataclass
from itertools import product

import numpy as np

__all__ = ["ControlMap", "bspline_matrix", "boolean_sum_offset", "boundary_ctrl_mask", "build_control_wire",
           "clamped_knots", "control_map", "fit_control_net"]


def clamped_knots(n_ctrl: int, degree: int) -> np.ndarray:
    """Clamped-uniform knots on [0, 1]: ``degree + 1`` copies of each end and
    ``n_ctrl - degree - 1`` evenly spaced interior knots."""
    if n_ctrl < degree + 1:
        raise ValueError(f"a degree-{degree} net needs at least {degree + 1} controls per axis, got {n_ctrl}")
    m = n_ctrl - degree - 1
    return np.concatenate([np.zeros(degree + 1), np.arange(1, m + 1) / (m + 1), np.ones(degree + 1)])


def bspline_matrix(n_node: int, n_ctrl: int, degree: int = 3) -> np.ndarray:
    """(n_node, n_ctrl) values of the clamped B-spline basis at the uniform
    node parameters, by the Cox-de Boor recursion (the parameter u = 1
    belongs to the last non-empty span, so the last function is 1 there)."""
    U = clamped_knots(n_ctrl, degree)
    u = np.linspace(0.0, 1.0, n_node)
    # degree 0: the indicator of the span holding u (half-open, last span closed)
    last = n_ctrl - 1  # the last non-empty span [U[n_ctrl - 1], U[n_ctrl])
    N = np.zeros((n_node, len(U) - 1))
    for j in range(len(U) - 1):
        if U[j + 1] > U[j]:
            N[:, j] = (u >= U[j]) & ((u < U[j + 1]) | ((j == last) & (u == U[j + 1])))
    for p in range(1, degree + 1):
        nxt = np.zeros((n_node, len(U) - 1 - p))
        for j in range(len(U) - 1 - p):
            a = U[j + p] - U[j]
            c = U[j + p + 1] - U[j + 1]
            left = (u - U[j]) / a * N[:, j] if a > 0 else 0.0
            right = (U[j + p + 1] - u) / c * N[:, j + 1] if c > 0 else 0.0
            nxt[:, j] = left + right
        N = nxt
    return N


@dataclass
class ControlMap:
    """X = M C (+ b) for one block: ``basis[k]`` is axis k's (n_node, n_ctrl)
    matrix, C has shape ``ctrl_shape + (d,)``, X ``node_shape + (d,)``."""

    d: int
    node_shape: tuple
    ctrl_shape: tuple
    degree: int
    basis: list

    def prolong(self, C: np.ndarray) -> np.ndarray:
        X = np.asarray(C, np.float64)
        for k, B in enumerate(self.basis):
            X = np.moveaxis(np.tensordot(B, X, axes=([1], [k])), 0, k)
        return X

    def gather(self, G: np.ndarray) -> np.ndarray:
        """The adjoint: a node field to the controls (M^T G)."""
        Y = np.asarray(G, np.float64)
        for k, B in enumerate(self.basis):
            Y = np.moveaxis(np.tensordot(B.T, Y, axes=([1], [k])), 0, k)
        return Y


def control_map(node_shape, ctrl_shape, degree: int = 3) -> ControlMap:
    node_shape = tuple(int(n) for n in node_shape)
    ctrl_shape = tuple(int(n) for n in ctrl_shape)
    if len(node_shape) != len(ctrl_shape) or len(node_shape) not in (2, 3):
        raise ValueError("node_shape and ctrl_shape must both be 2D or both 3D")
    return ControlMap(len(node_shape), node_shape, ctrl_shape, int(degree),
                      [bspline_matrix(n, c, degree) for n, c in zip(node_shape, ctrl_shape)])


def boundary_ctrl_mask(shape) -> np.ndarray:
    """True on the lattice points that lie on the lattice's boundary."""
    idx = np.indices(tuple(shape))
    return np.any([(i == 0) | (i == n - 1) for i, n in zip(idx, shape)], axis=0)


def _modes(mats, T):
    """Apply mats[j] along the j-th leading axis of T."""
    for j, A in enumerate(mats):
        T = np.moveaxis(np.tensordot(A, T, axes=([1], [j])), 0, j)
    return T


def fit_control_net(cmap: ControlMap, X0: np.ndarray) -> np.ndarray:
    """Least-squares net of the node field X0, facet by facet from the
    corners inward (reference control_ref.py:96-151): a corner control takes
    its corner node (the clamped basis interpolates there); every edge, face
    and finally the interior fits its interior controls to its own nodes
    with its boundary controls (set before) held, one pseudo-inverse per
    spanned axis (the free set of a slab is a tensor product of per-axis
    interior ranges, and the pseudo-inverse of a Kronecker product is the
    Kronecker product of the pseudo-inverses)."""
    d = cmap.d
    X0 = np.asarray(X0, np.float64)
    C = np.zeros(cmap.ctrl_shape + (d,))
    pinv = [np.linalg.pinv(B[:, 1:-1]) for B in cmap.basis]
    # slabs by dimension: offsets in {-1, 0, 1}^d, 0 = the axis is spanned
    slabs = sorted((o for o in product((-1, 0, 1), repeat=d)), key=lambda o: sum(v == 0 for v in o))
    for off in slabs:
        span = [k for k in range(d) if off[k] == 0]
        pick = lambda shape: tuple(slice(None) if o == 0 else (0 if o < 0 else n - 1)  # noqa: E731
                                   for o, n in zip(off, shape))
        cs, ns = pick(cmap.ctrl_shape), pick(cmap.node_shape)
        if not span:
            C[cs] = X0[ns]
            continue
        slab = C[cs].copy()
        inner = (slice(1, -1),) * len(span)
        slab[inner] = 0.0
        rhs = X0[ns] - _modes([cmap.basis[k] for k in span], slab)
        slab[inner] = _modes([pinv[k] for k in span], rhs)
        C[cs] = slab
    return C


def _boolean_sum_fill(F: np.ndarray) -> np.ndarray:
    """Transfinite fill of a node field from its boundary values, the
    reference's D-general form (egg/init/tfi.py:66-85): the sum over axes of
    the linear blends between the two end faces, minus (D - 1) times the
    multilinear blend of the 2^D corners, at the uniform node parameters."""
    d = F.ndim - 1

    def blend(G, ax):
        t = np.linspace(0.0, 1.0, F.shape[ax]).reshape([-1 if a == ax else 1 for a in range(d + 1)])
        return (1.0 - t) * np.take(G, [0], axis=ax) + t * np.take(G, [-1], axis=ax)

    total = sum(blend(F, ax) for ax in range(d))
    corner = F
    for ax in range(d):
        corner = blend(corner, ax)
    return total - (d - 1) * corner


def boolean_sum_offset(cmap: ControlMap, C: np.ndarray, X0: np.ndarray) -> np.ndarray:
    """b: the boundary mismatch X0 - M C extended inward by the transfinite
    fill (reference control_ref.py:154-171), so M C + b equals X0 on the
    block boundary; b depends on the boundary controls only."""
    X0 = np.asarray(X0, np.float64)
    delta = X0 - cmap.prolong(C)
    bnd = boundary_ctrl_mask(cmap.node_shape)
    delta[~bnd] = 0.0
    b = _boolean_sum_fill(delta)
    b[bnd] = delta[bnd]
    return b


def _window(B: np.ndarray, k_supp: int):
    """Per node row of a banded basis matrix: the first control of a
    k_supp-wide window holding its nonzeros (clamped into the lattice) and
    the window's weights."""
    n_node, n_ctrl = B.shape
    if n_ctrl < k_supp:
        raise ValueError(f"an axis with {n_ctrl} controls cannot hold a {k_supp}-wide support window")
    nz = B != 0.0
    first = np.where(nz.any(axis=1), nz.argmax(axis=1), 0)
    first = np.minimum(first, n_ctrl - k_supp)
    width = np.where(nz.any(axis=1), n_ctrl - 1 - nz[:, ::-1].argmax(axis=1) - first + 1, 0)
    if np.any(width > k_supp):
        raise ValueError(f"a basis row spans more than {k_supp} controls")
    cols = first[:, None] + np.arange(k_supp)[None, :]
    return first.astype(np.int32), np.take_along_axis(B, cols, axis=1).reshape(-1)


def build_control_wire(cmap: ControlMap, C0: np.ndarray, free_ctrl: np.ndarray, b_field: np.ndarray, *,
                       block: int = 0) -> dict:
    """The session's ``control`` dict for one block (reference
    control_backend.py:76-136 keys): block, ctrl_shape, k_supp, C0 (lattice
    row-major, last axis fastest, d values per control), free_idx (flat
    lattice indices of the free controls), b (node order of the block, d
    values per node), and per axis k ``off{k}`` (first control of each
    node's window) and ``wt{k}`` (the window's k_supp weights per node)."""
    d = cmap.d
    k_supp = cmap.degree + 1
    C0 = np.asarray(C0, np.float64)
    if C0.shape != cmap.ctrl_shape + (d,):
        raise ValueError(f"C0 has shape {C0.shape}, expected {cmap.ctrl_shape + (d,)}")
    if np.shape(free_ctrl) != cmap.ctrl_shape:
        raise ValueError(f"free_ctrl has shape {np.shape(free_ctrl)}, expected {cmap.ctrl_shape}")
    if np.shape(b_field) != cmap.node_shape + (d,):
        raise ValueError(f"b has shape {np.shape(b_field)}, expected {cmap.node_shape + (d,)}")
    wire = {"block": int(block), "ctrl_shape": np.asarray(cmap.ctrl_shape, np.int32), "k_supp": int(k_supp),
            "C0": np.ascontiguousarray(C0).reshape(-1),
            "free_idx": np.flatnonzero(np.asarray(free_ctrl, bool).reshape(-1)).astype(np.int32),
            "b": np.ascontiguousarray(b_field, np.float64).reshape(-1)}
    for k in range(d):
        wire[f"off{k}"], wire[f"wt{k}"] = _window(cmap.basis[k], k_supp)
    return wire
