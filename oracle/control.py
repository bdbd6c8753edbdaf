"""CPU restatement of the control-net reduced Gauss-Newton smoother
(TEST INFRASTRUCTURE ONLY: only tests/ use it).

Dense and FP64 throughout: the block map M (the Kronecker product of the
wire's per-axis windows, in global node rows), the reduced gradient
G = M_f^T g and the exact reduced Hessian A = sum over samples of
W_p^T H_p W_p, W_p = d vec(A_p) / d C_free (reference
egg/smoothing/control_ref.py:296-365), with the per-sample metric value,
gradient and Hessian from the golden-pinned C oracle (`oracle.metric`). The
outer loop follows the device loop's rules (reference
src/control_solve.hpp:533-676): the damping floor 1e-8 trace(P) / n from the
Jacobi blocks P_f = sum over nodes of M[n, f]^2 times the contracted node
Hessian, the step solved exactly (a dense solve in place of the PCG), the
line search alpha = 1, 1/2, 1/4, ... down to alpha_min, accept iff finite,
E <= E0 + tol and min det > 0, damping x10 on rejection, / 3 after a full
step.

The multi-block loop (`run_control_topo`) restates the reference's
ControlTopoSessionT::run (src/control_solve.hpp:2209-2425; NumPy twin
egg/smoothing/control_topology.py:2149-2364) on the same dense footing:
the nets' maps stacked into one global map Mg (a global node takes the row
of its owner block's net, the owner being the lowest-index block holding
the node, as the session's remap assigns it; nodes owned by a block
without a net stay put), the reduced gradient R^T (Mg^T g + P^T W P C) and
Hessian R^T (Mg^T K Mg + P^T W P) R with K the assembled node-space Hessian
of the fine energy, the penalty energy 1/2 sum w (P C)^2 added to the
line-search energy, the damping floor 1e-8 trace(P_stack) / (n_stack dim)
from the stacked Jacobi blocks.
"""

from __future__ import annotations

import numpy as np

from . import oracle as O


def window_matrix(off, wt, n_ctrl: int, k_supp: int) -> np.ndarray:
    """One axis's (n_node, n_ctrl) basis matrix from its window table."""
    off = np.asarray(off, np.int64)
    wt = np.asarray(wt, np.float64).reshape(len(off), k_supp)
    B = np.zeros((len(off), n_ctrl))
    for i, a in enumerate(off):
        B[i, a:a + k_supp] = wt[i]
    return B


def block_map(wire: dict, dim: int) -> np.ndarray:
    """Dense M (block nodes row-major, controls row-major)."""
    cs = [int(c) for c in np.asarray(wire["ctrl_shape"]).reshape(-1)[:dim]]
    M = None
    for k in range(dim):
        B = window_matrix(wire[f"off{k}"], wire[f"wt{k}"], cs[k], int(wire["k_supp"]))
        M = B if M is None else np.kron(M, B)
    return M


def _samples(es: dict, dim: int, X: np.ndarray):
    gc = np.asarray(es["gc"], np.int64)
    gn = [np.asarray(es[f"gn{k}"], np.int64) for k in range(dim)]
    s = [np.asarray(es[f"s{k}"], np.float64) for k in range(dim)]
    A = np.stack([s[k][:, None] * (X[gn[k]] - X[gc]) for k in range(dim)], axis=2)  # A[p, i, k]
    return gc, gn, s, A


def metric_field(dim: int, kind: str, A: np.ndarray):
    """Per sample: value, gradient (dim^2) and Hessian (dim^2 x dim^2) of the
    metric in vec(A) = row-major A."""
    V = dim * dim
    val = np.zeros(len(A))
    g = np.zeros((len(A), V))
    H = np.zeros((len(A), V, V))
    for p, a in enumerate(A.reshape(len(A), V)):
        val[p], g[p], H[p] = O.metric(dim, kind, a)
    return val, g, H


def reduced_system(dim: int, kind: str, X: np.ndarray, es: dict, Mf: np.ndarray):
    """(G (F, dim), A (F*dim, F*dim), P (F, dim, dim)) at the global state X;
    Mf: (n_nodes_global, F) columns of the free controls."""
    gc, gn, s, A = _samples(es, dim, X)
    _, g, H = metric_field(dim, kind, A)
    P_ = len(gc)
    F = Mf.shape[1]
    # Q[p, k, f] = d A[p, i, k] / d C[f, i]
    Q = np.stack([s[k][:, None] * (Mf[gn[k]] - Mf[gc]) for k in range(dim)], axis=1)
    gA = g.reshape(P_, dim, dim)  # [p, i, k]
    G = np.einsum("pik,pkf->fi", gA, Q)
    H4 = H.reshape(P_, dim, dim, dim, dim)  # [p, i, k, j, l]
    Ared = np.einsum("pkf,pikjl,plg->figj", Q, H4, Q).reshape(F * dim, F * dim)
    # contracted node Hessians -> Jacobi blocks
    n_nodes = Mf.shape[0]
    hn = np.zeros((n_nodes, dim, dim))
    u_c = -np.stack(s, axis=1)  # corner: u_k = -s_k
    np.add.at(hn, gc, np.einsum("pk,pikjl,pl->pij", u_c, H4, u_c))
    for k in range(dim):
        u = np.zeros((P_, dim))
        u[:, k] = s[k]
        np.add.at(hn, gn[k], np.einsum("pk,pikjl,pl->pij", u, H4, u))
    P = np.einsum("nf,nij->fij", Mf**2, hn)
    return G, Ared, P


def run_control(dim: int, kind: str, X: np.ndarray, es: dict, wire: dict, block_ids: np.ndarray, *,
                n_outer: int = 30, grad_tol: float = 1e-8, alpha_min: float = 1e-4, lm0: float = 0.0,
                lm_tries: int = 8, tol: float = 1e-6):
    """The reduced Gauss-Newton loop from the wire's net. Returns (X, C,
    report dict with energies, mindets, alphas, iters, converged)."""
    X = np.array(X, np.float64)
    ids = np.asarray(block_ids, np.int64).reshape(-1)
    M = block_map(wire, dim)
    C = np.asarray(wire["C0"], np.float64).reshape(-1, dim).copy()
    b = np.asarray(wire["b"], np.float64).reshape(-1, dim) if wire.get("b") is not None else 0.0
    free = np.asarray(wire["free_idx"], np.int64)
    Mf = np.zeros((len(X), len(free)))
    Mf[ids] = M[:, free]
    X[ids] = M @ C + b
    e0, md0 = O.energy_mindet(dim, kind, 0.0, es, X)
    rep = {"energies": [e0], "mindets": [md0], "alphas": [], "iters": 0, "converged": False}
    lam = float(lm0)
    nfd = len(free) * dim
    for _ in range(n_outer):
        G, Ared, P = reduced_system(dim, kind, X, es, Mf)
        if np.abs(G).max() < grad_tol:
            rep["converged"] = True
            break
        lam_floor = 1e-8 * float(np.trace(P, axis1=1, axis2=2).sum()) / nfd
        accepted = False
        for _try in range(lm_tries):
            dC = np.linalg.solve(Ared + lam * np.eye(nfd), -G.reshape(-1)).reshape(-1, dim)
            corr = Mf @ dC
            alpha = 1.0
            while alpha >= alpha_min:
                e, md = O.energy_mindet(dim, kind, 0.0, es, X + alpha * corr)
                if np.isfinite(e) and e <= e0 + tol and md > 0.0:
                    accepted = True
                    break
                alpha *= 0.5
            if accepted:
                break
            lam = max(10.0 * lam, lam_floor)
        if not accepted:
            break
        X = X + alpha * corr
        C[free] += alpha * dC
        e0, md0 = e, md
        rep["energies"].append(e0)
        rep["mindets"].append(md0)
        rep["alphas"].append(alpha)
        rep["iters"] += 1
        if alpha == 1.0:
            lam /= 3.0
    return X, C, rep


def reduced_system_sparse(dim: int, kind: str, X: np.ndarray, es: dict, Mf: np.ndarray):
    """reduced_system through the assembled node-space Hessian: K (N dim x N
    dim, sparse) = sum over samples of J_p^T H_p J_p, A = (Mf x I)^T K
    (Mf x I); the same (G, A, P) for map matrices too wide for the
    per-sample contraction."""
    import scipy.sparse as sp

    gc, gn, s, A = _samples(es, dim, X)
    _, g, H = metric_field(dim, kind, A)
    P_ = len(gc)
    N, F = Mf.shape
    H4 = H.reshape(P_, dim, dim, dim, dim)  # [p, i, k, j, l]
    gA = g.reshape(P_, dim, dim)  # [p, i, k]
    # node roles of a sample: the corner (u = -s) and the axis-k neighbours (u = s_k e_k)
    nodes = [gc] + gn
    us = [-np.stack(s, axis=1)]
    for k in range(dim):
        u = np.zeros((P_, dim))
        u[:, k] = s[k]
        us.append(u)
    grad = np.zeros((N, dim))
    rows, cols, vals = [], [], []
    ii, jj = np.meshgrid(np.arange(dim), np.arange(dim), indexing="ij")
    hn = np.zeros((N, dim, dim))
    for r, (nr, ur) in enumerate(zip(nodes, us)):
        np.add.at(grad, nr, np.einsum("pik,pk->pi", gA, ur))
        for q, (nq_, uq) in enumerate(zip(nodes, us)):
            B = np.einsum("pk,pikjl,pl->pij", ur, H4, uq)
            rows.append((nr[:, None, None] * dim + ii[None]).reshape(-1))
            cols.append((nq_[:, None, None] * dim + jj[None]).reshape(-1))
            vals.append(B.reshape(-1))
            if r == q:
                np.add.at(hn, nr, B)
    K = sp.coo_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                      shape=(N * dim, N * dim)).tocsr()
    Mx = sp.kron(sp.csr_matrix(Mf), sp.identity(dim, format="csr"), format="csr")
    G = (Mf.T @ grad)
    Ared = np.asarray((Mx.T @ (K @ Mx)).todense())
    P = np.einsum("nf,nij->fij", Mf**2, hn)
    return G, Ared, P


def _csr_dense(off, col, coef, n_rows: int, n_cols: int) -> np.ndarray:
    A = np.zeros((n_rows, n_cols))
    off = np.asarray(off, np.int64)
    col = np.asarray(col, np.int64)
    coef = np.asarray(coef, np.float64)
    for r in range(n_rows):
        for e in range(off[r], off[r + 1]):
            A[r, col[e]] += coef[e]
    return A


def topo_map(wire: dict, blocks, n_global: int, dim: int):
    """(Mg (n_global, n_stack), bg (n_global, dim), net_owned (n_global,)
    bool): the stacked map in global node rows, each node on its owner
    block's net (owner = the lowest-index block holding it; its first
    occurrence there in row-major order)."""
    owner_block = np.full(n_global, -1, np.int64)
    first = np.zeros(n_global, np.int64)
    for b in range(len(blocks) - 1, -1, -1):
        ids = np.asarray(blocks[b]).reshape(-1)
        rev = np.arange(len(ids))[::-1]
        owner_block[ids[rev]] = b  # the first occurrence wins within the block
        first[ids[rev]] = rev
    n_stack = sum(int(np.prod(np.asarray(e["ctrl_shape"]).reshape(-1)[:dim])) for e in wire["blocks"])
    Mg = np.zeros((n_global, n_stack))
    bg = np.zeros((n_global, dim))
    owned = np.zeros(n_global, bool)
    off = 0
    for e in wire["blocks"]:
        b = int(e["block"])
        M = block_map(e, dim)
        sel = np.flatnonzero(owner_block == b)
        Mg[sel, off:off + M.shape[1]] = M[first[sel]]
        bb = np.asarray(e["b"], np.float64).reshape(-1, dim) if e.get("b") is not None else np.zeros((M.shape[0], dim))
        bg[sel] = bb[first[sel]]
        owned[sel] = True
        off += M.shape[1]
    return Mg, bg, owned


def topo_operators(wire: dict, dim: int):
    """(R (n_stack dim, nq), P (n_pen, n_stack dim), w (n_pen,)) dense."""
    n_stack = sum(int(np.prod(np.asarray(e["ctrl_shape"]).reshape(-1)[:dim])) for e in wire["blocks"])
    nsd = n_stack * dim
    R = _csr_dense(wire["r_off"], wire["r_col"], wire["r_coef"], nsd, int(wire["nq"]))
    if wire.get("p_off") is not None and len(wire["p_off"]) > 1:
        Pm = _csr_dense(wire["p_off"], wire["p_col"], wire["p_coef"], len(wire["p_off"]) - 1, nsd)
        w = np.asarray(wire["p_w"], np.float64)
    else:
        Pm, w = np.zeros((0, nsd)), np.zeros(0)
    return R, Pm, w


def topo_system(dim: int, kind: str, X: np.ndarray, es: dict, Mg: np.ndarray, R: np.ndarray, Pm: np.ndarray,
                w: np.ndarray, C: np.ndarray):
    """(Gq (nq,), Aq (nq, nq), lam_floor) at the state X with stacked
    controls C: R^T (Mg^T g + P^T W P C), R^T (Mg^T K Mg + P^T W P) R and
    1e-8 trace(P_stack) / (n_stack dim)."""
    G, A, Pb = reduced_system_sparse(dim, kind, X, es, Mg)
    Gq = R.T @ (G.reshape(-1) + Pm.T @ (w * (Pm @ C)))
    Aq = R.T @ (A + Pm.T @ (w[:, None] * Pm)) @ R
    return Gq, Aq, 1e-8 * float(np.trace(Pb, axis1=1, axis2=2).sum()) / (Mg.shape[1] * dim)


def run_control_topo(dim: int, kind: str, X: np.ndarray, es: dict, wire: dict, blocks, *, n_outer: int = 30,
                     grad_tol: float = 1e-8, alpha_min: float = 1e-4, lm0: float = 0.0, lm_tries: int = 8,
                     tol: float = 1e-6):
    """The multi-block reduced Gauss-Newton loop of the wire's nets. Returns
    (X, C, report dict with energies, mindets, alphas, iters, converged)."""
    X = np.array(X, np.float64)
    Mg, bg, owned = topo_map(wire, blocks, len(X), dim)
    R, Pm, w = topo_operators(wire, dim)
    nq = R.shape[1]
    C = np.asarray(wire["C0"], np.float64).copy()  # stacked components

    def pen(Cv):
        r = Pm @ Cv
        return 0.5 * float(np.sum(w * r * r))

    X[owned] = (Mg @ C.reshape(-1, dim) + bg)[owned]
    e_f, md0 = O.energy_mindet(dim, kind, 0.0, es, X)
    e0 = e_f + pen(C)
    rep = {"energies": [e0], "mindets": [md0], "alphas": [], "iters": 0, "converged": False}
    if nq == 0:
        return X, C, rep
    lam = float(lm0)
    for _ in range(n_outer):
        Gq, Aq, lam_floor = topo_system(dim, kind, X, es, Mg, R, Pm, w, C)
        if np.abs(Gq).max() < grad_tol:
            rep["converged"] = True
            break
        accepted = False
        for _try in range(lm_tries):
            dq = np.linalg.solve(Aq + lam * np.eye(nq), -Gq)
            dC = R @ dq
            corr = Mg @ dC.reshape(-1, dim)
            alpha = 1.0
            while alpha >= alpha_min:
                e_f, md = O.energy_mindet(dim, kind, 0.0, es, X + alpha * corr)
                e = e_f + pen(C + alpha * dC)
                if np.isfinite(e) and e <= e0 + tol and md > 0.0:
                    accepted = True
                    break
                alpha *= 0.5
            if accepted:
                break
            lam = max(10.0 * lam, lam_floor)
        if not accepted:
            break
        X = X + alpha * corr
        C = C + alpha * dC
        e0, md0 = e, md
        rep["energies"].append(e0)
        rep["mindets"].append(md0)
        rep["alphas"].append(alpha)
        rep["iters"] += 1
        if alpha == 1.0:
            lam /= 3.0
    return X, C, rep
