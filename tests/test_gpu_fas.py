"""FAS V-cycles on the MI355X (HIP transfers + tau-shifted coarse sweeps):
per-cycle parity with the single-block FAS oracle, and the reference's
behavioural gates on the same grids (tests/smoothing/test_fas.py:72-256).
Tolerance floors follow the reference's FP32 build (tests/real_tol.py:
5e-4)."""

from __future__ import annotations

import numpy as np
import pytest

from tests.cases import Case, block_grid, bump, fas_single, fas_two_blocks, ogrid, single

pytestmark = pytest.mark.gpu

FLOOR = 5e-4  # real_tol floor of the reference's FP32 build


def _session(case):
    from egg_amd.smoothing import CppStructuredSweepSession

    return CppStructuredSweepSession(case.ctx, case.bsc, case.X)


def _monotone(e):
    slack = FLOOR * np.maximum(1.0, np.abs(e[:-1]))
    return bool(np.all(e[1:] <= e[:-1] + slack))


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
    np.testing.assert_allclose(e, eo, rtol=1e-3, atol=1e-5 * float(eo[0]))
    np.testing.assert_allclose(s.get_X(), Xo, atol=1e-5)


@pytest.mark.parametrize("graphs", [False, True])
def test_fas_bump_parity(oracle, graphs):
    """Same on the reference's smooth-bump grid (4 x 4 square, 17 x 17), with
    cycles launched one kernel at a time (line search decided on the host)
    and replayed as captured graphs (decided on the device)."""
    from oracle import fas

    c = fas_single()
    s = _session(c)
    s.core.set_graph_chunk(256 if graphs else 0)
    e, m = s.run_fas(5, nu_pre=2, nu_post=2, omega=0.8)
    Xo, eo, mo = fas.run_fas(2, "barrier", c.grid.blocks[0], c.grid.free, c.X, n_cycles=5, nu_pre=2, nu_post=2,
                             omega=0.8)
    np.testing.assert_allclose(e, eo, rtol=1e-3, atol=1e-5 * float(eo[0]))
    np.testing.assert_allclose(s.get_X(), Xo, atol=4e-5)


def test_fas_monotone_and_barrier():
    c = fas_single()
    e, m = _session(c).run_fas(6, omega=0.8)
    assert e.shape == (6,) and np.all(np.isfinite(e)) and np.all(m > 0) and _monotone(e)


def test_fas_beats_jacobi_at_equal_fine_budget():
    c = fas_single()
    e_fas, _ = _session(c).run_fas(5, nu_pre=2, nu_post=2, omega=0.8)
    e_jac, _ = _session(c).run(5 * 4, omega=0.8)
    print("fas", e_fas[-1], "jacobi", e_jac[-1])
    assert e_fas[-1] <= 0.5 * e_jac[-1] + FLOOR


def test_fas_agrees_with_jacobi_minimiser():
    c = fas_single()
    e_fas, _ = _session(c).run_fas(12, omega=0.8)
    e_jac, _ = _session(c).run(600, omega=0.8)
    assert e_fas[-1] <= e_jac[-1] + FLOOR * max(1.0, abs(e_jac[-1]))


def test_fas_two_blocks():
    c = fas_two_blocks()
    s = _session(c)
    assert s.mg_levels() == [[(9, 9), (9, 9)], [(5, 5), (5, 5)], [(3, 3), (3, 3)]]
    e, m = s.run_fas(6, omega=0.8)
    assert np.all(np.isfinite(e)) and np.all(m > 0) and _monotone(e)
    assert np.all(np.isfinite(s.get_X()))


def test_fas_two_blocks_beats_jacobi():
    """Coarse interface DOFs relax: two-block FAS reaches a long Jacobi run's energy."""
    c = fas_two_blocks(res=32, amp=0.08)
    e_fas, m_fas = _session(c).run_fas(12, omega=0.8)
    e_jac, _ = _session(c).run(1200, omega=0.8)
    print("fas", e_fas[-1], "jacobi", e_jac[-1])
    assert np.all(m_fas > 0)
    assert e_fas[-1] <= e_jac[-1] + FLOOR * max(1.0, abs(e_jac[-1]))


def test_fas_max_levels():
    c = fas_single()
    e2, m2 = _session(c).run_fas(4, omega=0.8, max_levels=2)
    assert e2.shape == (4,) and np.all(m2 > 0) and _monotone(e2)
    e1, m1 = _session(c).run_fas(4, omega=0.8, max_levels=1)
    assert e1.shape == (4,) and np.all(m1 > 0)


def test_fas_fallback_matches_jacobi():
    """max_levels=1 is plain Jacobi with the same fine budget, one report per cycle."""
    c = single(n=17, seed=3)
    e, m = _session(c).run_fas(3, nu_pre=2, nu_post=1, max_levels=1, omega=0.8)
    ej, mj = _session(c).run(9, omega=0.8, report_every=3)
    np.testing.assert_array_equal(e, ej)


def test_fas_carries_uncoarsenable_block():
    g = bump(block_grid([(0.0, 0.0, 17, 17, 0.25), (6.0, 0.0, 6, 6, 0.4)]), 0.08)
    c = Case(g)
    s = _session(c)
    assert s.mg_levels() == [[(9, 9), (6, 6)], [(5, 5), (6, 6)], [(3, 3), (6, 6)]]
    e, m = s.run_fas(6, omega=0.8)
    assert np.all(np.isfinite(e)) and np.all(m > 0) and _monotone(e)


def test_fas_falls_back_without_coarse_level():
    c = Case(bump(block_grid([(0.0, 0.0, 6, 6, 0.8)]), 0.05))
    s = _session(c)
    assert s.mg_levels() == []
    e, m = s.run_fas(3, omega=0.8)
    assert e.shape == (3,) and np.all(m > 0)


def test_fas_rejects_untangle():
    with pytest.raises(ValueError, match="untangle"):
        _session(single(n=9)).run_fas(1, phase="untangle")


def test_fas_ogrid_monotone():
    """5-block O-grid with valence-3 corners: monotone and valid per cycle."""
    c = ogrid(65, 33)
    e, m = _session(c).run_fas(8, nu_pre=2, nu_post=2, omega=0.8)
    e_jac, _ = _session(c).run(8 * 4, omega=0.8)
    print("fas", e[-1], "jacobi", e_jac[-1])
    assert np.all(m > 0) and _monotone(e) and e[-1] <= e_jac[-1]


def test_fas_3d_block():
    from egg_amd import grids

    c = Case(grids.single_block_3d(17, seed=2, amp=0.2))
    e, m = _session(c).run_fas(4, nu_pre=2, nu_post=2, omega=0.8)
    assert np.all(m > 0) and _monotone(e) and e[-1] < 0.1 * e[0]


def _cube_case():
    from egg_amd import grids

    return Case(grids.single_block_3d(17, seed=3, amp=0.15))


@pytest.mark.parametrize("make,nu", [(lambda: fas_single(amp=0.3), (2, 2)), (lambda: fas_two_blocks(), (1, 2)),
                                     (lambda: ogrid(33, 17, tangle_seed=None), (3, 3)), (_cube_case, (2, 1))])
def test_fas_graph_matches_launches(make, nu):
    """A FAS cycle replayed as a captured hipGraph (line-search ladder and
    its pick on the device, report appended at a device counter) gives the
    bits of the same cycle launched kernel by kernel with the ladder decided
    on the host: over several calls, an odd sweep count per cycle (the
    replay alternates between the two captured parities), plain sweeps in
    between, and a changed cycle shape (recapture)."""
    c = make()
    a = _session(c)
    b = _session(c)
    a.core.set_graph_chunk(0)
    b.core.set_graph_chunk(256)
    for rnd in range(3):
        ea, ma = a.run_fas(4, nu_pre=nu[0], nu_post=nu[1], omega=0.8, max_levels=3)
        eb, mb = b.run_fas(4, nu_pre=nu[0], nu_post=nu[1], omega=0.8, max_levels=3)
        np.testing.assert_array_equal(eb, ea)
        np.testing.assert_array_equal(mb, ma)
        np.testing.assert_array_equal(b.get_X(), a.get_X())
        a.run(3, omega=0.8)
        b.run(3, omega=0.8)
    ea, _ = a.run_fas(2, nu_pre=2, nu_post=1, omega=0.7, max_levels=2)
    eb, _ = b.run_fas(2, nu_pre=2, nu_post=1, omega=0.7, max_levels=2)
    np.testing.assert_array_equal(eb, ea)
    np.testing.assert_array_equal(b.get_X(), a.get_X())


def test_fas_graph_with_proposal_matches_launches():
    """The captured cycle also carries the learned proposal of the fine
    smoothing (a random network, so some proposals are accepted)."""
    from egg_amd._cpp import cpp_core
    from oracle import mlp as M

    c = ogrid(33, 17)
    net = cpp_core.ProposalMLP(2, 0, 0)
    net.set_params((np.random.default_rng(7).standard_normal(M.n_params(2)) * 0.05).astype(np.float32))
    a = _session(c)
    b = _session(c)
    a.core.set_graph_chunk(0)
    b.core.set_graph_chunk(256)
    for s in (a, b):
        s.set_proposal(net, mode="predict_correct", scale=0.5)
    for _ in range(2):
        ea, ma = a.run_fas(3, nu_pre=2, nu_post=2, omega=0.8, max_levels=3)
        eb, mb = b.run_fas(3, nu_pre=2, nu_post=2, omega=0.8, max_levels=3)
        np.testing.assert_array_equal(eb, ea)
        np.testing.assert_array_equal(b.get_X(), a.get_X())


@pytest.mark.parametrize("graphs", [0, 256])
def test_level_variants_one_lane_per_cell_same_bits(graphs):
    """Each coarse level can run its own tile shape. The one-lane-per-cell
    shapes (2D variants 0-3, 6, 7) sum a patch in the same order, so cycles
    whose coarse levels use several of them give the bits of variant 0
    throughout; one-lane-per-sample shapes on the coarse levels (a different
    summation order) change the bits, which shows the choice is applied."""
    c = ogrid(65, 33)
    a, b, d = _session(c), _session(c), _session(c)
    for s in (a, b, d):
        s.core.set_graph_chunk(graphs)
        s.core.set_tile_variant(0)
    b.core.set_level_variants([6, 7, 2])
    d.core.set_level_variants([4, 5, 4])
    for _ in range(2):
        ea, ma = a.run_fas(3, nu_pre=2, nu_post=2, omega=0.8, max_levels=4)
        eb, mb = b.run_fas(3, nu_pre=2, nu_post=2, omega=0.8, max_levels=4)
        ed, md = d.run_fas(3, nu_pre=2, nu_post=2, omega=0.8, max_levels=4)
        np.testing.assert_array_equal(eb, ea)
        np.testing.assert_array_equal(mb, ma)
        np.testing.assert_array_equal(b.get_X(), a.get_X())
        np.testing.assert_allclose(ed, ea, rtol=1e-4)
    assert not np.array_equal(d.get_X(), a.get_X())
    np.testing.assert_allclose(d.get_X(), a.get_X(), atol=5e-4)


def test_level_variants_fas_parity(oracle):
    """Coarse levels on one-lane-per-sample and cells-per-lane shapes keep
    per-cycle parity with the FP64 FAS oracle (levels 17, 9 and 5 here)."""
    from oracle import fas

    c = single(n=33, seed=1)
    s = _session(c)
    s.core.set_level_variants([4, 9, 5])
    e, m = s.run_fas(6, nu_pre=2, nu_post=2, omega=0.8)
    Xo, eo, mo = fas.run_fas(2, "barrier", c.grid.blocks[0], c.grid.free, c.X, n_cycles=6, nu_pre=2, nu_post=2,
                             omega=0.8)
    assert len(e) == 6 and np.all(m > 0)
    np.testing.assert_allclose(e, eo, rtol=1e-3, atol=1e-5 * float(eo[0]))
    np.testing.assert_allclose(s.get_X(), Xo, atol=1e-5)


@pytest.mark.parametrize("graphs", [0, 256])
@pytest.mark.parametrize("make", [lambda: ogrid(65, 33), _cube_case])
def test_one_workgroup_levels_same_bits(make, graphs):
    """A coarse level smoothed by one single-workgroup launch per smoothing
    call (level variant V + v) walks variant v's tiles with the one-sweep
    kernel's per-DOF arithmetic and Jacobi buffers, so its cycles give the
    bits of the per-sweep launches of v, odd and even sweep counts alike."""
    c = make()
    nv = _session(c).core.set_tile_variant(-1)
    for v in ([0, 2, 4] if c.grid.d == 2 else [0, 1, 3]):
        a, b = _session(c), _session(c)
        for s in (a, b):
            s.core.set_graph_chunk(graphs)
        a.core.set_level_variants([v] * 3)
        b.core.set_level_variants([nv + v] * 3)
        print(f"one-workgroup levels: dim {c.grid.d} variant {v} graphs {graphs}", flush=True)
        for nu, nuc in ((2, 5), (3, 4)):
            ea, ma = a.run_fas(2, nu_pre=nu, nu_post=nu, nu_coarse=nuc, omega=0.8, max_levels=4)
            eb, mb = b.run_fas(2, nu_pre=nu, nu_post=nu, nu_coarse=nuc, omega=0.8, max_levels=4)
            np.testing.assert_array_equal(eb, ea)
            np.testing.assert_array_equal(mb, ma)
            np.testing.assert_array_equal(b.get_X(), a.get_X())


@pytest.mark.parametrize("make", [lambda: ogrid(65, 33), _cube_case])
def test_autotune_levels(make):
    """The level autotuner times every variant on each coarse level, as
    per-sweep launches and (small levels) as one single-workgroup launch,
    leaves the fine state as it was, and the cycles that follow give the
    bits of a session told the same picks directly; out-of-range picks and
    the untangle phase are refused."""
    c = make()
    s, t = _session(c), _session(c)
    for x in (s, t):
        x.run(5, omega=0.8)
    nv = s.core.set_tile_variant(-1)
    X0 = s.get_X()
    best, ms = s.core.autotune_levels(4, 0.8, 4)
    n_lv = len(s.mg_levels())
    assert ms.shape == (min(3, n_lv), 2 * nv) and np.all(ms[:, :nv] > 0)
    assert np.all((ms[:, nv:] > 0) | (ms[:, nv:] == -1)) and np.any(ms[-1, nv:] > 0)
    timed = np.where(ms > 0, ms, np.inf)
    assert len(best) == ms.shape[0] and all(int(np.argmin(timed[i])) == best[i] for i in range(len(best)))
    np.testing.assert_array_equal(s.get_X(), X0)
    t.core.set_level_variants(best)
    for _ in range(2):
        es, ms_ = s.run_fas(2, nu_pre=2, nu_post=2, omega=0.8, max_levels=4)
        et, mt = t.run_fas(2, nu_pre=2, nu_post=2, omega=0.8, max_levels=4)
        np.testing.assert_array_equal(es, et)
        np.testing.assert_array_equal(s.get_X(), t.get_X())
    with pytest.raises(ValueError, match="level tile variant"):
        s.core.set_level_variants([2 * nv])
    with pytest.raises(ValueError, match="barrier"):
        s.core.autotune_levels(2, 0.8, 3, "untangle")
