"""Jacobi-preconditioned conjugate gradients on the MI355X (FP64).

:func:`pcg_solve` runs the device solver behind ``eggmi_pcg_csr`` (the PCG
skeleton of the reference's control-net step, src/control_solve.hpp:761-818)
on a CSR operator, :func:`pcg_solve_lattice` the same solver matrix-free
(``eggmi_pcg_lattice``); :func:`lattice_laplacian` builds the benchmark's SPD
system: the graph Laplacian of an n^d node lattice with Dirichlet closure
(every lattice edge to a node outside the lattice adds to the diagonal only),
i.e. the 2d+1-point stencil 2d / -1.
"""

from __future__ import annotations

import numpy as np


def lattice_laplacian(n: int = 33, d: int = 3):
    """CSR (rowptr int64, col int32, val float64) of the Dirichlet lattice Laplacian."""
    N = n**d
    idx = np.arange(N).reshape((n,) * d)
    rows = [np.arange(N)]
    cols = [np.arange(N)]
    vals = [np.full(N, 2.0 * d)]
    for ax in range(d):
        for sh in (-1, 1):
            src = np.take(idx, np.arange(max(0, -sh), n - max(0, sh)), axis=ax).reshape(-1)
            dst = np.take(idx, np.arange(max(0, sh), n - max(0, -sh)), axis=ax).reshape(-1)
            rows.append(src)
            cols.append(dst)
            vals.append(np.full(src.size, -1.0))
    r = np.concatenate(rows)
    c = np.concatenate(cols)
    v = np.concatenate(vals)
    order = np.lexsort((c, r))
    r, c, v = r[order], c[order], v[order]
    rowptr = np.zeros(N + 1, np.int64)
    np.add.at(rowptr, r + 1, 1)
    rowptr = np.cumsum(rowptr)
    return rowptr, c.astype(np.int32), v.astype(np.float64)


def pcg_solve_lattice(n: int, d: int, b, *, rtol: float = 1e-8, max_iter: int = 1000,
                      recurrence: str = "standard"):
    """Solve the Dirichlet lattice Laplacian system of :func:`lattice_laplacian`
    matrix-free on the GPU; returns (x, iterations, relative residual, status)."""
    from egg_amd._cpp import cpp_core

    return cpp_core.pcg_lattice(np.full(d, n, np.int32), 2.0 * d, -1.0, np.ascontiguousarray(b, np.float64),
                                rtol=rtol, max_iter=max_iter, recurrence=recurrence)


def pcg_solve(rowptr, col, val, b, *, rtol: float = 1e-8, max_iter: int = 1000, recurrence: str = "standard"):
    """Solve A x = b on the GPU; returns (x, iterations, relative residual, status).
    ``recurrence``: "standard" (the reference's, two launches per iteration) or
    "one_reduction" (Chronopoulos-Gear, one launch per iteration)."""
    from egg_amd._cpp import cpp_core

    return cpp_core.pcg_csr(np.ascontiguousarray(rowptr, np.int64), np.ascontiguousarray(col, np.int32),
                            np.ascontiguousarray(val, np.float64), np.ascontiguousarray(b, np.float64),
                            rtol=rtol, max_iter=max_iter, recurrence=recurrence)
