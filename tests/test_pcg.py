"""Jacobi-PCG: the oracle against scipy (CPU) and the FP64 device solver
against the oracle (GPU), on the benchmark's 33^3 Dirichlet lattice
Laplacian and smaller systems (reference parity gate
tests/smoothing/test_control_cpp.py:122-171: iteration count +-2)."""

from __future__ import annotations

import numpy as np
import pytest
import scipy.sparse.linalg as spla


def _system(n, d, seed=0):
    from egg_amd.pcg import lattice_laplacian
    from oracle import pcg as P

    rp, c, v = lattice_laplacian(n, d)
    A = P.csr(rp, c, v)
    b = np.random.default_rng(seed).normal(size=A.shape[0])
    return rp, c, v, A, b


@pytest.mark.parametrize("n,d", [(9, 2), (12, 3)])
def test_oracle_pcg_matches_direct_solve(n, d):
    from oracle import pcg as P

    _, _, _, A, b = _system(n, d)
    x, it, rr = P.pcg(A, b, rtol=1e-10)
    xs = spla.spsolve(A.tocsc(), b)
    assert rr <= 1e-10 and it > 0
    np.testing.assert_allclose(x, xs, rtol=0, atol=1e-8 * np.abs(xs).max())


def test_lattice_laplacian_is_spd():
    _, _, _, A, _ = _system(6, 3)
    assert abs(A - A.T).max() == 0
    assert np.linalg.eigvalsh(A.toarray()).min() > 0


RECURRENCES = ["standard", "one_reduction"]


@pytest.mark.gpu
@pytest.mark.parametrize("rec", RECURRENCES)
@pytest.mark.parametrize("n,d,rtol", [(33, 3, 1e-8), (17, 2, 1e-10), (9, 3, 1e-12)])
def test_device_pcg_parity(n, d, rtol, rec):
    from egg_amd.pcg import pcg_solve
    from oracle import pcg as P

    rp, c, v, A, b = _system(n, d, seed=n)
    x, it, rr, status = pcg_solve(rp, c, v, b, rtol=rtol, max_iter=2000, recurrence=rec)
    xo, ito, rro = P.pcg(A, b, rtol=rtol, max_iter=2000)
    print(f"{rec} n={n} d={d} device iters={it} relres={rr:.3e} oracle iters={ito} relres={rro:.3e}", flush=True)
    assert status == 1 and rr <= rtol
    assert abs(it - ito) <= 2
    np.testing.assert_allclose(x, xo, rtol=0, atol=10 * rtol * np.abs(xo).max())
    # the residual the device reports is the true one
    r = b - A @ x
    assert np.sqrt(r @ r / (b @ b)) == pytest.approx(rr, rel=1e-6, abs=1e-14)


@pytest.mark.gpu
@pytest.mark.parametrize("rec", RECURRENCES)
def test_device_pcg_edge_cases(rec):
    from egg_amd.pcg import pcg_solve

    rp, c, v, A, b = _system(5, 2)
    x, it, rr, status = pcg_solve(rp, c, v, np.zeros_like(b), rtol=1e-8, recurrence=rec)
    assert status == 3 and it == 0 and np.all(x == 0)
    x, it, rr, status = pcg_solve(rp, c, v, b, rtol=1e-8, max_iter=3, recurrence=rec)
    assert status == 0 and it == 3
    # negative definite operator: stops on non-positive curvature at once
    x, it, rr, status = pcg_solve(rp, c, -v, b, rtol=1e-8, recurrence=rec)
    assert status == 2 and it == 0 and np.all(x == 0) and rr == pytest.approx(1.0)


def test_pcg_rejects_bad_csr():
    from egg_amd.pcg import pcg_solve

    rp, c, v, A, b = _system(4, 2)
    c2 = c.copy()
    c2[0] = 10**6
    with pytest.raises(ValueError, match="column index"):
        pcg_solve(rp, c2, v, b)


@pytest.mark.gpu
@pytest.mark.parametrize("rec", RECURRENCES)
@pytest.mark.parametrize("n,d,rtol", [(33, 3, 1e-8), (17, 2, 1e-10), (5, 1, 1e-12)])
def test_device_pcg_lattice_parity(n, d, rtol, rec):
    """The matrix-free lattice operator solves the same system as the CSR
    Laplacian: iterations within +-2 of the FP64 oracle, solution within
    rtol, reported residual the true one."""
    from egg_amd.pcg import pcg_solve_lattice
    from oracle import pcg as P

    rp, c, v, A, b = _system(n, d, seed=n + 1)
    x, it, rr, status = pcg_solve_lattice(n, d, b, rtol=rtol, max_iter=2000, recurrence=rec)
    xo, ito, rro = P.pcg(A, b, rtol=rtol, max_iter=2000)
    print(f"lattice {rec} n={n} d={d} device iters={it} relres={rr:.3e} oracle iters={ito}", flush=True)
    assert status == 1 and rr <= rtol and abs(it - ito) <= 2
    np.testing.assert_allclose(x, xo, rtol=0, atol=10 * rtol * np.abs(xo).max())
    r = b - A @ x
    assert np.sqrt(r @ r / (b @ b)) == pytest.approx(rr, rel=1e-6, abs=1e-14)


@pytest.mark.gpu
@pytest.mark.parametrize("rec", RECURRENCES)
def test_device_pcg_iteration_limits(rec):
    """max_iter = 0 returns x = 0 with relres 1 (the reference's loop never
    runs); a limit that is not a multiple of the graph chunk stops exactly
    there; rtol >= 1 still runs one iteration (no check before the first)."""
    from egg_amd.pcg import pcg_solve

    rp, c, v, A, b = _system(9, 3, seed=5)
    x, it, rr, status = pcg_solve(rp, c, v, b, rtol=1e-14, max_iter=0, recurrence=rec)
    assert status == 0 and it == 0 and np.all(x == 0) and rr == pytest.approx(1.0)
    for lim in (1, 15, 17, 33):
        x, it, rr, status = pcg_solve(rp, c, v, b, rtol=1e-14, max_iter=lim, recurrence=rec)
        assert status == 0 and it == lim, (lim, it, status)
    x, it, rr, status = pcg_solve(rp, c, v, b, rtol=2.0, max_iter=50, recurrence=rec)
    assert status == 1 and it == 1
