"""Per-sample energy weights and per-sample (boundary-layer) targets on the
MI355X (stored stencils and the explicit sample reduction) against the FP64
oracle on the same wire. A
weight scales a sample's mu, gradient and Hessian, not its det (reference
src/patch.hpp:45-77, :250-292); a weighted wire ships every DOF's stored
stencil (egg/smoothing/flat_context.py:186-218).

Bar as for the unweighted stored path: positions within 1e-5 of the domain
extent of the FP64 oracle, reported energies within 2e-3 along the
trajectory, and each report equal to the FP64 weighted reduction of the
returned FP32 state."""

from __future__ import annotations

import numpy as np
import pytest

from egg_amd import grids
from tests.cases import Case, lr_pair, ogrid

pytestmark = pytest.mark.gpu

POS_TOL = 1e-5
TRAJ_RTOL = 2e-3


def _session(case):
    from egg_amd.smoothing import CppStructuredSweepSession

    return CppStructuredSweepSession(case.ctx, case.bsc, case.X)


def _weighted(grid, seed=5, lo=0.2, hi=1.8):
    ns = Case(grid).ctx.energy_stencil["num_samples"]
    return Case(grid, weight=np.random.default_rng(seed).uniform(lo, hi, ns))


def _extent(c):
    return float((c.X.max(0) - c.X.min(0)).max())


@pytest.mark.parametrize("make", [lambda: grids.single_block_2d(17, seed=5, amp=0.2), lambda: lr_pair().grid,
                                  lambda: ogrid(33, 17).grid, lambda: grids.single_block_3d(7, seed=2, amp=0.2)])
@pytest.mark.parametrize("small", [False, True])
def test_weighted_parity(oracle, make, small):
    """`small`: the tile variant with the fewest lanes per stored DOF."""
    g = make()
    c = _weighted(g)
    s = _session(c)
    if small:
        s.core.set_tile_variant(12 if g.d == 2 else 4)
    info = s.core.info()
    assert info["n_synth"] == 0 and info["n_stored"] == g.n_free and info["structured_energy"] == 0
    e, m = s.run(12, omega=0.8, report_every=3)
    Xo, eo, mo = oracle.jacobi(g.d, "barrier", 0.0, 0.8, 12, c.X, c.patches(oracle), c.ctx.energy_stencil,
                               report_every=3)
    X = s.get_X()
    assert np.abs(X - Xo).max() / _extent(c) < POS_TOL
    np.testing.assert_allclose(e, eo, rtol=TRAJ_RTOL)
    E, M = oracle.energy_mindet(g.d, "barrier", 0.0, c.ctx.energy_stencil, X.astype(np.float32), prec="f32")
    assert e[-1] == pytest.approx(E, rel=1e-9) and m[-1] == pytest.approx(M, rel=1e-9)


def test_zero_and_scalar_weights(oracle):
    """Zero weights on a third of the samples keep oracle parity; a weight
    of 2 everywhere doubles every reported energy and leaves the trajectory
    (Newton steps are invariant under a power-of-two scaling) as the unit
    weights give it."""
    g = grids.single_block_2d(17, seed=6, amp=0.2)
    ns = Case(g).ctx.energy_stencil["num_samples"]
    w = np.ones(ns)
    w[np.random.default_rng(1).permutation(ns)[: ns // 3]] = 0.0
    c = Case(g, weight=w)
    e, m = (s := _session(c)).run(10, omega=0.8, report_every=2)
    Xo, eo, _ = oracle.jacobi(2, "barrier", 0.0, 0.8, 10, c.X, c.patches(oracle), c.ctx.energy_stencil,
                              report_every=2)
    assert np.abs(s.get_X() - Xo).max() / _extent(c) < POS_TOL
    np.testing.assert_allclose(e, eo, rtol=TRAJ_RTOL)
    one, two = Case(g, weight=np.ones(ns)), Case(g, weight=np.full(ns, 2.0))
    s1, s2 = _session(one), _session(two)
    e1, m1 = s1.run(10, omega=0.8, report_every=2)
    e2, m2 = s2.run(10, omega=0.8, report_every=2)
    np.testing.assert_allclose(e2, 2.0 * e1, rtol=1e-6)
    np.testing.assert_allclose(m2, m1, rtol=1e-6)
    np.testing.assert_allclose(s2.get_X(), s1.get_X(), atol=1e-6)


def test_fas_on_a_weighted_wire_is_plain_sweeps():
    """A weighted wire ships every DOF's stored stencil, so the multigrid
    hierarchy (built on the synthesized interior DOFs) has no coarse level and
    run_fas is nu_pre + nu_post plain sweeps per cycle, bit for bit."""
    c = _weighted(grids.single_block_2d(17, seed=5, amp=0.2))
    a, b = _session(c), _session(c)
    assert a.mg_levels() == []
    ef, mf = a.run_fas(3, nu_pre=2, nu_post=1, omega=0.8)
    ej, mj = b.run(9, omega=0.8, report_every=3)
    np.testing.assert_array_equal(ef, ej)
    np.testing.assert_array_equal(a.get_X(), b.get_X())


@pytest.mark.parametrize("variant", [0, 12, 13])
def test_boundary_layer_target_parity(oracle, variant):
    """The O-grid's rings under a boundary-layer target (anisotropic W^-1 per
    sample, reference BoundaryLayerTarget): every DOF on the stored path
    (16, 4 or 8 lanes per DOF), parity with the FP64 oracle, and the sweeps
    pull the first ring layer towards the circle."""
    g = grids.ogrid_2d(33, 17)
    W = grids.ogrid_boundary_layer_w_inv(g, 33)
    c = Case(g, w_inv=W)
    s = _session(c)
    s.core.set_tile_variant(variant)
    assert s.core.info()["n_synth"] == 0
    e, m = s.run(10, omega=0.8, report_every=5)
    Xo, eo, mo = oracle.jacobi(2, "barrier", 0.0, 0.8, 10, c.X, c.patches(oracle), c.ctx.energy_stencil,
                               report_every=5)
    assert np.abs(s.get_X() - Xo).max() / _extent(c) < POS_TOL
    np.testing.assert_allclose(e, eo, rtol=TRAJ_RTOL)
    s.run(400, omega=0.8)
    ring = np.asarray(g.blocks[1])
    X0, X1 = c.X.reshape(-1, 2), s.get_X().reshape(-1, 2)

    def first_layer(X):
        return np.linalg.norm(X[ring[1:-1, 1]] - X[ring[1:-1, 0]], axis=1).mean()

    assert first_layer(X1) < 0.7 * first_layer(X0) and s.run(1, omega=0.8)[1][-1] > 0
