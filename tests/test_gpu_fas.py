"""FAS V-cycles on the MI355X (HIP transfers + tau-shifted coarse sweeps),
against the single-block FAS oracle and the reference's behavioural gates
(tests/smoothing/test_fas.py:72-256)."""

from __future__ import annotations

import numpy as np
import pytest

from tests.cases import Case, ogrid, single

pytestmark = pytest.mark.gpu


def _session(case):
    from egg_amd.smoothing import CppStructuredSweepSession

    return CppStructuredSweepSession(case.ctx, case.bsc, case.X)


def test_fas_single_block_parity(oracle):
    """Per-cycle energies and the final state against the FP64 FAS oracle."""
    from oracle import fas

    c = single(n=33, seed=1)
    s = _session(c)
    e, m = s.run_fas(6, nu_pre=2, nu_post=2, omega=0.8)
    Xo, eo, mo = fas.run_fas(2, "barrier", c.grid.blocks[0], c.grid.free, c.X, n_cycles=6, nu_pre=2, nu_post=2,
                             omega=0.8)
    print("gpu", e, "\noracle", eo)
    assert len(e) == 6 and np.all(m > 0)
    e0 = float(eo[0])
    np.testing.assert_allclose(e, eo, rtol=1e-3, atol=1e-5 * e0)
    np.testing.assert_allclose(s.get_X(), Xo, atol=1e-5)


def test_fas_beats_jacobi_and_is_monotone():
    """Energy is monotone per cycle and FAS reaches at most half the Jacobi
    energy at an equal fine-sweep budget (reference test_fas.py:94-109)."""
    c = ogrid(65, 33)
    s = _session(c)
    e, m = s.run_fas(8, nu_pre=2, nu_post=2, omega=0.8)
    assert np.all(np.diff(e) <= 1e-6 * abs(e[0])) and np.all(m > 0)
    j = _session(c)
    ej, mj = j.run(8 * 4, omega=0.8)
    print("fas", e[-1], "jacobi", ej[-1])
    assert e[-1] <= 0.5 * ej[-1]


def test_fas_same_minimiser():
    """FAS and a long Jacobi run reach the same minimiser (test_fas.py:112-125)."""
    c = ogrid(17, 9)
    s = _session(c)
    s.run_fas(30, nu_pre=2, nu_post=2, omega=0.8)
    j = _session(c)
    j.run(6000, omega=0.8)
    np.testing.assert_allclose(s.get_X(), j.get_X(), atol=2e-4)


def test_fas_fallback_without_levels():
    """max_levels=1 falls back to block-Jacobi with the same fine budget and
    a report per cycle."""
    c = single(n=17, seed=3)
    s = _session(c)
    e, m = s.run_fas(3, nu_pre=2, nu_post=1, max_levels=1, omega=0.8)
    j = _session(c)
    ej, mj = j.run(9, omega=0.8, report_every=3)
    np.testing.assert_array_equal(e, ej)


def test_fas_rejects_untangle():
    c = single(n=9)
    with pytest.raises(ValueError, match="untangle"):
        _session(c).run_fas(1, phase="untangle")


def test_fas_3d_block():
    """3D FAS on a perturbed hex block: monotone and converging."""
    from egg_amd import grids

    c = Case(grids.single_block_3d(17, seed=2, amp=0.2))
    s = _session(c)
    e, m = s.run_fas(4, nu_pre=2, nu_post=2, omega=0.8)
    assert np.all(np.diff(e) <= 1e-6 * abs(e[0])) and np.all(m > 0)
    assert e[-1] < 0.1 * e[0]
