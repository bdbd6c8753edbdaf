"""Edge cases of the sweep path on the MI355X (through the C ABI): report
cadences that do not divide the sweep count, the smallest blocks (one DOF in
2D and 3D), a block with no DOF, and zero-count calls.

Report cadence follows the reference's run_block_jacobi
(src/sweep.hpp:2037-2085): a report after every k-th sweep and after the
last one, one report at the end when report_every <= 0."""

from __future__ import annotations

import numpy as np
import pytest

from egg_amd import grids
from tests.cases import Case, single

pytestmark = pytest.mark.gpu


def _session(case):
    from egg_amd.smoothing import CppStructuredSweepSession

    return CppStructuredSweepSession(case.ctx, case.bsc, case.X)


def _fp32(X):
    return np.asarray(X, np.float32).astype(np.float64)


@pytest.mark.parametrize("n,every", [(10, 4), (10, 0), (10, 25), (7, 7), (1, 3)])
def test_ragged_report_cadence(oracle, n, every):
    """Report count and every reported energy / min det against the FP64
    oracle run with the same cadence."""
    c = single(n=17, seed=3)
    s = _session(c)
    e, m = s.run(n, omega=0.8, report_every=every)
    Xo, eo, mo = oracle.jacobi(2, "barrier", 0.0, 0.8, n, c.X, c.patches(oracle), c.ctx.energy_stencil,
                               report_every=every)
    k = n if every <= 0 else every
    assert len(e) == len(m) == len(eo) == (n + k - 1) // k
    np.testing.assert_allclose(e, eo, rtol=2e-3)
    np.testing.assert_allclose(m, mo, rtol=2e-3)
    np.testing.assert_allclose(s.get_X(), Xo, atol=1e-5)


@pytest.mark.parametrize("graphs", [0, 256])
@pytest.mark.parametrize("dim", [2, 3])
def test_one_dof_block(oracle, dim, graphs):
    """A 3^D block has one DOF, surrounded by the uniform unit-cube boundary:
    by symmetry the barrier energy is minimised at the centre, which the
    sweeps reach; each step along the way matches the FP64 oracle."""
    g = grids.single_block_2d(3, seed=1, amp=0.3) if dim == 2 else grids.single_block_3d(3, seed=1, amp=0.3)
    c = Case(g)
    assert int(g.free.sum()) == 1
    s = _session(c)
    s.core.set_graph_chunk(graphs)
    e, m = s.run(6, omega=0.8, report_every=1)
    Xo, eo, mo = oracle.jacobi(dim, "barrier", 0.0, 0.8, 6, c.X, c.patches(oracle), c.ctx.energy_stencil,
                               report_every=1)
    # the energy falls towards zero at the minimiser, where FP32 cancellation
    # in mu - 1 sets an absolute floor
    np.testing.assert_allclose(e, eo, rtol=1e-5, atol=1e-6 * float(eo[0]))
    np.testing.assert_allclose(s.get_X(), Xo, atol=1e-6)
    s.run(40, omega=0.8)
    X = s.get_X().reshape(-1, dim)
    dof = int(np.flatnonzero(g.free.reshape(-1))[0])
    np.testing.assert_allclose(X[dof], np.full(dim, 0.5), atol=1e-5)
    fixed = ~g.free.reshape(-1)
    np.testing.assert_array_equal(X[fixed], _fp32(c.X.reshape(-1, dim)[fixed]))


def test_block_without_dofs(oracle):
    """A 2 x 2 block has no DOF: sweeps and FAS (no coarse level: plain
    sweeps) leave it as it is and report its energy each time."""
    c = Case(grids.single_block_2d(2, seed=1))
    assert int(c.grid.free.sum()) == 0
    s = _session(c)
    e, m = s.run(3, omega=0.8, report_every=1)
    E, M = oracle.energy_mindet(2, "barrier", 0.0, c.ctx.energy_stencil, c.X.astype(np.float32), prec="f32")
    assert len(e) == 3 and np.all(e == e[0]) and np.all(m == m[0])
    assert e[0] == pytest.approx(E, rel=1e-9) and m[0] == pytest.approx(M, rel=1e-9)
    ef, mf = s.run_fas(2, nu_pre=1, nu_post=1, omega=0.8)
    assert len(ef) == 2 and np.all(ef == e[0])
    np.testing.assert_array_equal(s.get_X(), _fp32(c.X))


def test_zero_counts():
    """Zero sweeps or cycles return no reports and leave the state alone."""
    c = single(n=9, seed=2)
    s = _session(c)
    X0 = s.get_X()
    for e, m in (s.run(0, omega=0.8), s.run_fas(0, omega=0.8)):
        assert len(e) == 0 and len(m) == 0
    np.testing.assert_array_equal(s.get_X(), X0)
