"""Jacobi-PCG oracle (TEST INFRASTRUCTURE ONLY): NumPy FP64 restatement of
the reference's pcg_solve (src/control_solve.hpp:761-818): x0 = 0, r = b,
z = D^-1 r, p = z; alpha = rz / pAp (stop when pAp <= 0), x += alpha p,
r -= alpha Ap, stop at |r|^2 <= rtol^2 |b|^2 (the iteration counts), beta =
rz' / rz, p = z + beta p. Returns (x, iterations, relative residual)."""

from __future__ import annotations

import numpy as np
import scipy.sparse as sp


def csr(rowptr, col, val):
    n = len(rowptr) - 1
    return sp.csr_matrix((val, col, rowptr), shape=(n, n))


def pcg(A, b, rtol=1e-8, max_iter=1000):
    x = np.zeros_like(b)
    r = b.copy()
    b2 = float(r @ r)
    if b2 == 0.0:
        return x, 0, 0.0
    stop2 = rtol * rtol * b2
    dinv = 1.0 / A.diagonal()
    z = dinv * r
    p = z.copy()
    rz = float(r @ z)
    rr = b2
    it = 0
    while it < max_iter:
        ap = A @ p
        pap = float(p @ ap)
        if not pap > 0.0:
            break
        a = rz / pap
        x += a * p
        r -= a * ap
        rr = float(r @ r)
        if rr <= stop2:
            it += 1
            break
        z = dinv * r
        rz_new = float(r @ z)
        p = z + (rz_new / rz) * p
        rz = rz_new
        it += 1
    return x, it, float(np.sqrt(rr / b2))
