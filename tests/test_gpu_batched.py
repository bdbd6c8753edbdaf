"""Batched workload on the MI355X: many independent multi-block grids of
uneven block sizes in one session (one sweep launch for all of them, halo
copies pushed in-kernel), against the FP64 oracle; FAS V(4,4) on three
levels; the tile-shape autotuner."""

from __future__ import annotations

import numpy as np
import pytest

from egg_amd import grids
from tests.cases import Case

pytestmark = pytest.mark.gpu


def _session(c):
    from egg_amd.smoothing import CppStructuredSweepSession

    return CppStructuredSweepSession(c.ctx, c.bsc, c.X)


def _small_batch(tangled, seed=3):
    g, index = grids.batched_2d(5, seed=seed, min_blocks=3, max_blocks=5, sides=(33, 41, 49),
                                frac=0.02 if tangled else 0.0, jitter=0.2)
    return Case(g), index


def test_batched_plain_parity(oracle):
    c, index = _small_batch(False)
    assert len(index) == 5 and c.grid.n_free > 20000 and grids.min_det(c.grid) > 0
    s = _session(c)
    info = s.core.info()
    assert info["n_stored"] > 0 and info["n_share"] > 0
    e, m = s.run(20, omega=0.8, report_every=5)
    Xo, eo, mo = oracle.jacobi(2, "barrier", 0.0, 0.8, 20, c.X, c.patches(oracle), c.ctx.energy_stencil,
                               report_every=5)
    ext = float((c.X.max(0) - c.X.min(0)).max())
    assert np.abs(s.get_X() - Xo).max() / ext < 1e-5
    np.testing.assert_allclose(e, eo, rtol=2e-4)
    np.testing.assert_allclose(m, mo, rtol=1e-3)


def test_batched_untangle_parity(oracle):
    c, _ = _small_batch(True)
    assert grids.min_det(c.grid) < 0
    s = _session(c)
    delta = 2.0 * abs(grids.min_det(c.grid))
    e, m = s.run(10, phase="untangle", delta=delta, omega=0.5, report_every=5)
    Xo, eo, mo = oracle.jacobi(2, "untangle", delta, 0.5, 10, c.X, c.patches(oracle), c.ctx.energy_stencil,
                               report_every=5)
    ext = float((c.X.max(0) - c.X.min(0)).max())
    assert np.abs(s.get_X() - Xo).max() / ext < 1e-5
    np.testing.assert_allclose(e, eo, rtol=2e-4)


def test_batched_untangle_then_fas():
    from egg_amd.smoothing import cpp_untangle

    c, _ = _small_batch(True, seed=4)
    s = _session(c)
    X, md, outer, _ = cpp_untangle(c.ctx, c.bsc, c.X, session=s)
    assert md > 1e-9
    levels = s.mg_levels()
    assert len(levels) >= 2
    e, m = s.run_fas(6, nu_pre=4, nu_post=4, omega=0.8, max_levels=3)
    assert np.all(m > 0)
    assert np.all(np.diff(e) <= 1e-6 * abs(e[0]))
    # FAS beats plain Jacobi at the same number of fine sweeps
    j = _session(c)
    j.set_X(X)
    ej, _ = j.run(6 * 8, omega=0.8)
    assert e[-1] < ej[-1]


def test_autotuner_keeps_state_and_picks_a_variant():
    c, _ = _small_batch(False, seed=5)
    s = _session(c)
    s.run(3, omega=0.8)
    X0 = s.get_X()
    best, ms = s.core.autotune(5, 0.8)
    assert 0 <= best < len(ms) and np.all(ms > 0)
    assert ms[best] == ms.min()
    np.testing.assert_array_equal(s.get_X(), X0)
    # the tuned session continues exactly like one that ran the same
    # variants without tuning (the default for the first sweeps, then the pick)
    r = _session(c)
    r.run(3, omega=0.8)
    r.core.set_tile_variant(int(best))
    e1, _ = s.run(4, omega=0.8)
    e2, _ = r.run(4, omega=0.8)
    np.testing.assert_array_equal(e1, e2)
    np.testing.assert_array_equal(s.get_X(), r.get_X())
