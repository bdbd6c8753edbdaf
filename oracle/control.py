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
