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

__all__ = ["ControlMap", "bspline_matrix", "boolean_sum_offset", "boundary_ctrl_mask", "build_control_topology",
           "build_control_wire", "clamped_knots", "control_map", "fit_control_net", "topology_mismatch"]


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


def boolean_sum_offset(cmap: ControlMap, C: np.ndarray, X0: np.ndarray, keep: np.ndarray | None = None) -> np.ndarray:
    """b: the boundary mismatch X0 - M C extended inward by the transfinite
    fill (reference control_ref.py:154-171), so M C + b equals X0 on the
    block boundary; b depends on the boundary controls only. ``keep``
    (node-shaped bool): the boundary nodes whose mismatch is kept, the
    others' is taken as 0 (seam faces of a multi-block topology, whose
    nodes stay pure spline, reference control_topology.py:1386-1420)."""
    X0 = np.asarray(X0, np.float64)
    delta = X0 - cmap.prolong(C)
    bnd = boundary_ctrl_mask(cmap.node_shape)
    delta[~bnd] = 0.0
    if keep is not None:
        delta[~np.asarray(keep, bool)] = 0.0
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


def _face_controls(cmap: ControlMap, axis: int, side: int):
    """Flat lattice indices of the controls on one face of the lattice and,
    per control, its weights on the face's nodes (node-lattice order of the
    face): the clamped basis puts the face's nodes on the face's controls
    only, with the tangential basis products as weights."""
    d = cmap.d
    cs = cmap.ctrl_shape
    tang = [k for k in range(d) if k != axis]
    a_face = 0 if side == 0 else cs[axis] - 1
    out = []
    for t in product(*[range(cs[k]) for k in tang]):
        a = [0] * d
        a[axis] = a_face
        for k, v in zip(tang, t):
            a[k] = v
        w = np.ones([cmap.node_shape[k] for k in tang])
        for j, k in enumerate(tang):
            shape = [1] * len(tang)
            shape[j] = -1
            w = w * cmap.basis[k][:, a[k]].reshape(shape)
        out.append((int(np.ravel_multi_index(a, cs)), tuple(a), w))
    return out


def build_control_topology(blocks, X: np.ndarray, ctrl_shapes, *, degree: int = 3, net_blocks=None,
                           c1: bool = False, seam_weight: float = 0.0):
    """A multi-block control wire (the reference's "blocks" wire,
    control_backend.py / bindings.cpp:593-740 keys) for conforming blocks
    given as per-block global node id arrays over the positions X.

    Each net block gets a fitted net and its Boolean-sum offset, built from
    the mismatch on the faces no other net block shares (seam nodes are
    pure spline, so the blocks agree on them). Controls on
    a block face shared by two or more net blocks are identified across the
    blocks by their weights on the face's global nodes (equal weight maps =
    one control; needs equal control and node counts along the seam) and
    share one reduced DOF, so the blocks stay watertight; controls on faces
    no other net block shares are fixed (no R entries); the rest are free.
    ``c1``: across a seam shared by exactly two faces, the control next to
    the seam on the second block follows 2 q_seam - q_first (a two-entry R
    row, tangent continuity for equal spacing). ``seam_weight`` > 0: one
    penalty row per seam control and component, weight seam_weight, on the
    second difference across the seam C_first - 2 C_seam + C_second.

    Returns ``(wire, info)``: info holds ``c_off`` (stacked control offset
    per net), ``net_blocks``, ``q0`` (nq,) and ``root_of`` (reduced root per
    stacked control, -1 fixed, -2 dependent).
    """
    X = np.asarray(X, np.float64)
    d = X.shape[1]
    nb = len(blocks)
    net_blocks = list(range(nb)) if net_blocks is None else [int(b) for b in net_blocks]
    net_of = {b: i for i, b in enumerate(net_blocks)}
    cmaps, C_fit, X0s, c_off = [], [], [], [0]
    for i, b in enumerate(net_blocks):
        ids = np.asarray(blocks[b])
        cm = control_map(ids.shape, ctrl_shapes[i], degree)
        X0 = X[ids]
        cmaps.append(cm)
        X0s.append(X0)
        C_fit.append(fit_control_net(cm, X0).reshape(-1, d))
        c_off.append(c_off[-1] + int(np.prod(cm.ctrl_shape)))
    n_stack = c_off[-1]
    Cf = np.concatenate(C_fit, axis=0)
    parent = np.arange(n_stack)

    def find(a):
        while parent[a] != a:
            parent[a] = parent[parent[a]]
            a = parent[a]
        return a

    fixed = np.zeros(n_stack, bool)
    faces = {}
    for b in range(nb):
        ids = np.asarray(blocks[b])
        for k in range(d):
            for side in (0, 1):
                f = np.take(ids, 0 if side == 0 else -1, axis=k)
                faces.setdefault(frozenset(f.ravel().tolist()), []).append((b, k, side, f))
    seams = []  # (first net face controls, second) for the two-face seams
    outer = [np.zeros(np.shape(blocks[b]), bool) for b in net_blocks]  # nodes on faces no other net shares
    for members in faces.values():
        net_members = [m for m in members if m[0] in net_of]
        shared = len(net_members) >= 2 and len(net_members) == len(members)
        if not shared:
            for b, k, side, _ in net_members:
                idx = [slice(None)] * d
                idx[k] = 0 if side == 0 else -1
                outer[net_of[b]][tuple(idx)] = True
        sig_maps = []
        for b, k, side, f in net_members:
            i = net_of[b]
            ctrls = _face_controls(cmaps[i], k, side)
            if not shared:
                for flat, _, _ in ctrls:
                    fixed[c_off[i] + flat] = True
                continue
            sig = []  # per face control: its weights as {global node: weight}
            for flat, a, w in ctrls:
                nz = np.abs(w) > 1e-14
                sig.append((dict(zip(f[nz].tolist(), w[nz].tolist())), c_off[i] + flat, a))
            sig_maps.append((i, k, side, sig))
        if not shared:
            continue

        def same(u, v):
            return u.keys() == v.keys() and all(abs(u[g] - v[g]) <= 1e-9 for g in u)

        i0, k0, s0, sig0 = sig_maps[0]
        pairs = []
        for i1, k1, s1, sig1 in sig_maps[1:]:
            match = []
            for u, m0, a0 in sig0:
                hit = [(m1, a1) for v, m1, a1 in sig1 if same(u, v)]
                if len(hit) != 1:
                    raise ValueError("seam faces do not share their control rows: equal control and node counts "
                                     "along the seam are required")
                parent[find(hit[0][0])] = find(m0)
                match.append((m0, a0, hit[0][0], hit[0][1]))
            pairs.append((i1, k1, s1, match))
        if len(sig_maps) == 2:
            i1, k1, s1, match = pairs[0]
            seams.append(((i0, k0, s0), (i1, k1, s1), match))
    roots = np.array([find(a) for a in range(n_stack)])
    group_fixed = np.zeros(n_stack, bool)
    np.logical_or.at(group_fixed, roots, fixed)
    # group means of the fitted controls
    cnt = np.bincount(roots, minlength=n_stack).astype(np.float64)
    mean = np.zeros((n_stack, d))
    np.add.at(mean, roots, Cf)
    mean[cnt > 0] /= cnt[cnt > 0, None]
    C0 = mean[roots].copy()
    root_of = np.full(n_stack, -1, np.int64)
    free_groups = sorted(set(roots[~group_fixed[roots]].tolist()), key=lambda g: np.flatnonzero(roots == g)[0])
    gidx = {g: r for r, g in enumerate(free_groups)}
    for a in range(n_stack):
        if not group_fixed[roots[a]]:
            root_of[a] = gidx[roots[a]]
    dep = {}  # dependent control -> (seam control, first-side control)
    pen = []  # (first-side control, seam control, second-side control)

    def inward(cm, k, side, a):
        a = list(a)
        a[k] = 1 if side == 0 else cm.ctrl_shape[k] - 2
        return int(np.ravel_multi_index(a, cm.ctrl_shape))

    for (iA, kA, sA), (iB, kB, sB), match in seams:
        for mA, aA, mB, aB in match:
            a1 = c_off[iA] + inward(cmaps[iA], kA, sA, aA)
            b1 = c_off[iB] + inward(cmaps[iB], kB, sB, aB)
            pen.append((a1, mA, b1))
            if not c1 or root_of[mA] < 0:
                continue
            single = lambda m: root_of[m] >= 0 and cnt[roots[m]] == 1 and m not in dep  # noqa: E731
            if single(a1) and single(b1) and a1 != b1 and b1 not in [v[1] for v in dep.values()]:
                dep[b1] = (mA, a1)
    # renumber the roots without the dependent controls' own
    dead = {int(root_of[b1]) for b1 in dep}
    alive = [r for r in range(len(free_groups)) if r not in dead]
    renum = {r: j for j, r in enumerate(alive)}
    for a in range(n_stack):
        if a in dep:
            root_of[a] = -2
        elif root_of[a] >= 0:
            root_of[a] = renum[int(root_of[a])]
    nr = len(alive)
    q0 = np.zeros((nr, d))
    for a in range(n_stack):
        if root_of[a] >= 0:
            q0[root_of[a]] = C0[a]
    for b1, (s, a1) in dep.items():
        C0[b1] = 2.0 * q0[root_of[s]] - q0[root_of[a1]]
    # R over the stacked components
    r_off, r_col, r_coef = [0], [], []
    for a in range(n_stack):
        for c in range(d):
            if root_of[a] >= 0:
                r_col.append(int(root_of[a]) * d + c)
                r_coef.append(1.0)
            elif root_of[a] == -2:
                s, a1 = dep[a]
                for m, w in ((s, 2.0), (a1, -1.0)):
                    r_col.append(int(root_of[m]) * d + c)
                    r_coef.append(w)
            r_off.append(len(r_col))
    wire = {"blocks": [], "C0": np.ascontiguousarray(C0).reshape(-1), "nq": nr * d,
            "r_off": np.asarray(r_off, np.int32), "r_col": np.asarray(r_col, np.int32),
            "r_coef": np.asarray(r_coef, np.float64)}
    k_supp = degree + 1
    for i, b in enumerate(net_blocks):
        cm = cmaps[i]
        Cb = C0[c_off[i]:c_off[i + 1]].reshape(cm.ctrl_shape + (d,))
        entry = {"block": int(b), "ctrl_shape": np.asarray(cm.ctrl_shape, np.int32), "k_supp": int(k_supp),
                 "b": np.ascontiguousarray(boolean_sum_offset(cm, Cb, X0s[i], keep=outer[i])).reshape(-1)}
        for k in range(d):
            entry[f"off{k}"], entry[f"wt{k}"] = _window(cm.basis[k], k_supp)
        wire["blocks"].append(entry)
    if seam_weight > 0.0 and pen:
        p_off, p_col, p_coef = [0], [], []
        for a1, s, b1 in pen:
            for c in range(d):
                for m, w in ((a1, 1.0), (s, -2.0), (b1, 1.0)):
                    p_col.append(m * d + c)
                    p_coef.append(w)
                p_off.append(len(p_col))
        wire.update(p_off=np.asarray(p_off, np.int32), p_col=np.asarray(p_col, np.int32),
                    p_coef=np.asarray(p_coef, np.float64), p_w=np.full(len(p_off) - 1, float(seam_weight)))
    info = {"c_off": np.asarray(c_off), "net_blocks": net_blocks, "q0": q0.reshape(-1), "root_of": root_of,
            "cmaps": cmaps}
    return wire, info


def topology_mismatch(wire: dict, blocks, C: np.ndarray) -> float:
    """The largest spread, over the global nodes several net blocks share,
    of the positions their nets give the node (M C + b per block; 0 for a
    watertight topology)."""
    C = np.asarray(C, np.float64).reshape(-1)
    vals = {}
    off = 0
    for e in wire["blocks"]:
        cs = tuple(int(v) for v in e["ctrl_shape"])
        ids = np.asarray(blocks[e["block"]])
        d = len(cs)
        cm = ControlMap(d, ids.shape, cs, int(e["k_supp"]) - 1,
                        [_window_matrix(e[f"off{k}"], e[f"wt{k}"], cs[k], int(e["k_supp"])) for k in range(d)])
        n = int(np.prod(cs))
        Xb = cm.prolong(C[off * d:(off + n) * d].reshape(cs + (d,))) + np.asarray(e["b"]).reshape(ids.shape + (d,))
        off += n
        for gid, x in zip(ids.reshape(-1).tolist(), Xb.reshape(-1, d)):
            vals.setdefault(gid, []).append(x)
    spread = [np.max(np.ptp(np.array(v), axis=0)) for v in vals.values() if len(v) > 1]
    return float(max(spread)) if spread else 0.0


def _window_matrix(off, wt, n_ctrl: int, k_supp: int) -> np.ndarray:
    off = np.asarray(off, np.int64)
    wt = np.asarray(wt, np.float64).reshape(len(off), k_supp)
    B = np.zeros((len(off), n_ctrl))
    for i, a in enumerate(off):
        B[i, a:a + k_supp] = wt[i]
    return B
