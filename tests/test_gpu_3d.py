"""3D hex path on the MI355X: the 7-block cubed-sphere-in-cube workload
(central ball block + six shells, interfaces with stored-stencil DOFs,
singular edges and corners) against the FP64 oracle, and 3D untangling."""

from __future__ import annotations

import numpy as np
import pytest

from tests.cases import Case

pytestmark = pytest.mark.gpu


def _session(c):
    from egg_amd.smoothing import CppStructuredSweepSession

    return CppStructuredSweepSession(c.ctx, c.bsc, c.X)


def test_cubed_sphere_parity(oracle):
    from egg_amd import grids

    c = Case(grids.cubed_sphere_3d(9))
    s = _session(c)
    info = s.core.info()
    assert info["n_stored"] > 0 and info["n_synth"] > 0
    e, m = s.run(8, omega=0.8, report_every=1)
    Xo, eo, mo = oracle.jacobi(3, "barrier", 0.0, 0.8, 8, c.X, c.patches(oracle), c.ctx.energy_stencil,
                               report_every=1)
    err = np.abs(s.get_X() - Xo).max() / 4.0
    print(f"cubed sphere max|dX|/extent={err:.2e} E={e[-1]:.9g} E_f64={eo[-1]:.9g}")
    assert err < 1e-5
    np.testing.assert_allclose(e, eo, rtol=2e-3)


def test_cubed_sphere_halo_bit_exact():
    from egg_amd import grids

    c = Case(grids.cubed_sphere_3d(9))
    s = _session(c)
    s.run(3, omega=0.8)
    P = s.core.packed()
    tab = s.core.tables()
    assert np.array_equal(P[:, tab["copy_dst"]].view(np.uint32), P[:, tab["copy_src"]].view(np.uint32))


def test_cubed_sphere_untangle(oracle):
    from egg_amd import grids
    from egg_amd.smoothing import cpp_untangle

    g = grids.tangle(grids.cubed_sphere_3d(9), frac=0.01, factor=1.5, seed=5)
    c = Case(g)
    assert grids.min_det(g) < 0
    X, md, outer, delta = cpp_untangle(c.ctx, c.bsc, c.X)
    assert md > 1e-9
    E, M = oracle.energy_mindet(3, "barrier", 0.0, c.ctx.energy_stencil, X.astype(np.float32), prec="f32")
    assert M == pytest.approx(md, rel=1e-9)


def test_cubed_sphere_fas_monotone():
    from egg_amd import grids

    c = Case(grids.cubed_sphere_3d(17))
    s = _session(c)
    e, m = s.run_fas(3, nu_pre=2, nu_post=2, omega=0.8, max_levels=3)
    assert np.all(np.diff(e) <= 1e-6 * abs(e[0])) and np.all(m > 0)
