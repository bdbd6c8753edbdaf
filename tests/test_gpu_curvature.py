"""Block-interface C2 curvature windows on the MI355X (reference
src/curvature.hpp, src/sweep.hpp:1461-1487 and :1595-1612) against the FP64
oracle on the same wire: the windows' energy w r^2, gradient and stabilised
Gauss-Newton Hessian on top of the shape metric, their energy with the DOF
moved in each line-search trial. DOFs with windows take the special stored
path (a synthesized interior DOF is expanded to its stored patch); the energy
reports stay the shape metric's, as in the reference.

Bar: positions within 1e-5 of the domain extent, energies within 2e-3 per
report."""

from __future__ import annotations

import numpy as np
import pytest

from egg_amd import grids
from tests.cases import Case, ogrid

pytestmark = pytest.mark.gpu

POS_TOL = 1e-5
TRAJ_RTOL = 2e-3


def _session(case):
    from egg_amd.smoothing import CppStructuredSweepSession

    return CppStructuredSweepSession(case.ctx, case.bsc, case.X)


def _extent(c):
    return float((c.X.max(0) - c.X.min(0)).max())


def _ogrid_c2(w=0.5, tangle_seed=None):
    g = ogrid(33, 17, tangle_seed=tangle_seed).grid
    return Case(g, curvature=grids.interface_c2_windows(g, w))


@pytest.mark.parametrize("variant", [0, 8, 12, 13])
def test_curvature_parity(oracle, variant):
    c = _ogrid_c2()
    s = _session(c)
    plain = _session(Case(c.grid)).core.info()
    info = s.core.info()
    assert info["n_synth"] < plain["n_synth"] and info["n_stored"] > plain["n_stored"]
    s.core.set_tile_variant(variant)
    e, m = s.run(20, omega=0.8, report_every=5)
    Xo, eo, mo = oracle.jacobi(2, "barrier", 0.0, 0.8, 20, c.X, c.patches(oracle), c.ctx.energy_stencil,
                               report_every=5)
    X = s.get_X()
    assert np.abs(X - Xo).max() / _extent(c) < POS_TOL
    np.testing.assert_allclose(e, eo, rtol=TRAJ_RTOL)
    X0, _, _ = oracle.jacobi(2, "barrier", 0.0, 0.8, 20, c.X, Case(c.grid).patches(oracle),
                             c.ctx.energy_stencil)
    assert np.abs(X - X0).max() / _extent(c) > 1e-3  # the windows change the result


def test_curvature_with_constrained_wall(oracle):
    """Curvature windows across the annulus' radial block faces with the
    circle wall constrained: both kinds of special DOF in one launch."""
    g, ents = grids.annulus_2d(33, 17, seed=1)
    c = Case(g, entities=ents, curvature=grids.interface_c2_windows(g, 0.3))
    s = _session(c)
    e, m = s.run(15, omega=0.8, report_every=5)
    Xo, eo, mo = oracle.jacobi(2, "barrier", 0.0, 0.8, 15, c.X, c.patches(oracle), c.ctx.energy_stencil,
                               report_every=5)
    X = s.get_X()
    assert np.abs(X - Xo).max() / _extent(c) < POS_TOL
    np.testing.assert_allclose(e, eo, rtol=TRAJ_RTOL)
    con = np.array(sorted(ents))
    np.testing.assert_allclose(np.linalg.norm(X[con], axis=1), 0.5, rtol=1e-6)


def test_curvature_graphs_and_fas():
    """Graph replay gives the bits of single launches (sweeps and whole FAS
    cycles, whose gradient passes also run the special DOFs), and FAS with
    the windows keeps every cell valid while the shape energy falls."""
    c = _ogrid_c2(tangle_seed=None)
    a, b = _session(c), _session(c)
    b.core.set_graph_chunk(8)
    ea, _ = a.run(16, omega=0.8, report_every=8)
    eb, _ = b.run(16, omega=0.8, report_every=8)
    np.testing.assert_array_equal(eb, ea)
    np.testing.assert_array_equal(b.get_X(), a.get_X())
    fa, fb = _session(c), _session(c)
    fb.core.set_graph_chunk(256)
    e1, m1 = fa.run_fas(4, nu_pre=2, nu_post=2, omega=0.8, max_levels=3)
    e2, m2 = fb.run_fas(4, nu_pre=2, nu_post=2, omega=0.8, max_levels=3)
    np.testing.assert_array_equal(e2, e1)
    np.testing.assert_array_equal(fb.get_X(), fa.get_X())
    assert np.all(m1 > 0) and e1[-1] < e1[0]
