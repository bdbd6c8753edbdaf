"""Multi-sweep launches: k block-Jacobi sweeps per launch over LDS-staged
dependency closures.

CPU: the closure tables of many topologies (interfaces, valence-3 corners,
valence-5/6 stars, singular hosts moved to other blocks, a batch of strips,
FAS coarse levels) pass the host-side schedule check: k sweeps of a
synthetic nonlinear update of every DOF from its stencil, run through the
tables' schedule (regions, compute rectangles per distance, stored entries,
LDS pushes, core writes and replica pushes), equal k plain synchronous
sweeps of the same update bit for bit.

GPU: k sweeps per launch agree with k one-sweep launches to FP32 rounding
noise (the same per-DOF device functions; the compiler fuses multiply-adds
differently in the two kernels, which moves about 1 % of the updates by one
ulp per sweep), plain and FAS, across kernel variants, tile shapes, report
cadences and captured graph chunks. A stale or missing neighbour in any
sub-sweep would move a DOF by a fraction of its step, orders of magnitude
more."""

from __future__ import annotations

import numpy as np
import pytest

from egg_amd import grids
from egg_amd.smoothing import structured_arrays
from egg_amd.smoothing import structured_tables as T
from tests.cases import Case, block_grid, fas_two_blocks, lr_pair, ogrid, single, star


def _forced_hosts(g):
    """Interface tables with every singular node hosted by the
    highest-numbered block of its fan (owner slots in another block)."""
    bsc = T.build_block_structured_context(2, g.blocks, g.free)
    fan = {}
    for b, a in enumerate(g.blocks):
        a = np.asarray(a)
        for corner in np.ndindex(2, 2):
            lg = tuple(0 if c == 0 else s - 1 for c, s in zip(corner, a.shape))
            fan.setdefault(int(a[lg]), []).append((b, lg))
    sb, sl = [], []
    for b, lg in zip(bsc.sing_block, bsc.sing_logical):
        gid = int(np.asarray(g.blocks[b])[tuple(lg)])
        hb, hl = max(fan[gid], key=lambda c: c[0])
        sb.append(hb)
        sl.append(hl)
    bsc.sing_block = np.asarray(sb, np.int32)
    bsc.sing_logical = np.asarray(sl, np.int32).reshape(-1, 2)
    return bsc


def _batch():
    g, _ = grids.batched_2d(3, seed=5, min_blocks=3, max_blocks=4, sides=(33, 41), frac=0.0, jitter=0.2)
    return Case(g)


CASES = {
    "single17": lambda: single(n=17),
    "ogrid33": lambda: ogrid(33, 17),
    "lr_pair": lambda: lr_pair(9),
    "star5": lambda: star(k=5, n=9),
    "star6": lambda: star(k=6, n=7),
    "tee": lambda: Case(block_grid([(0.0, 0.0, 9, 17, 0.25), (2.0, 0.0, 9, 17, 0.25), (2.0, 4.0, 9, 9, 0.25)])),
    "batch3": _batch,
}


def _core_session(c, device, bsc=None):
    from egg_amd._cpp import cpp_core

    return cpp_core.CppStructuredSweepSession(c.ctx.wire, structured_arrays(bsc if bsc is not None else c.bsc),
                                              np.ascontiguousarray(c.X).ravel(), device=device, dim=c.d)


@pytest.mark.parametrize("name", sorted(CASES))
@pytest.mark.parametrize("k,ty,tx", [(2, 8, 8), (2, 8, 4), (3, 4, 16), (4, 16, 8), (8, 4, 4)])
def test_schedule_check_fine(name, k, ty, tx):
    c = CASES[name]()
    s = _core_session(c, "tables")
    s.set_multi_sweep(0, k, ty, tx, 0)
    info = s.multi_sweep_info(0, check=True, seed=k + 7)
    assert info["check_err"] == 0.0, info
    assert info["k"] == k
    assert info["core_dofs"] == c.grid.n_free
    # every DOF updated at least once per sub-sweep, the redundant rings on top
    assert info["updates"] >= k * info["core_dofs"]


def test_schedule_check_singular_hosts():
    g = grids.ogrid_2d(33, 17)
    c = Case(g)
    s = _core_session(c, "tables", _forced_hosts(g))
    for k, ty, tx in ((2, 8, 8), (4, 4, 32), (6, 16, 16)):
        s.set_multi_sweep(0, k, ty, tx, 1)
        assert s.multi_sweep_info(0, check=True, seed=k)["check_err"] == 0.0


@pytest.mark.parametrize("level", [1, 2])
def test_schedule_check_coarse_levels(level):
    c = fas_two_blocks(res=32)
    s = _core_session(c, "tables")
    assert len(s.mg_levels()) > level
    for k, ty, tx in ((2, 4, 4), (4, 8, 8), (8, 16, 16)):
        s.set_multi_sweep(level, k, ty, tx, 0)
        info = s.multi_sweep_info(level, check=True, seed=3)
        assert info["check_err"] == 0.0, info


def test_rejects_what_it_cannot_run():
    c = CASES["ogrid33"]()
    s = _core_session(c, "tables")
    with pytest.raises(ValueError, match="sweeps per launch"):
        s.set_multi_sweep(0, 9, 8, 8, 0)
    with pytest.raises(ValueError, match="kernel variant"):
        s.set_multi_sweep(0, 2, 8, 8, 99)
    big = _core_session(ogrid(129, 65), "tables")
    with pytest.raises(ValueError, match="outgrows the LDS"):
        big.set_multi_sweep(0, 8, 128, 128, 0)
    assert big.multi_sweep_info(0)["k"] == 0  # a failed build leaves it off
    c3 = Case(grids.single_block_3d(9))
    s3 = _core_session(c3, "tables")
    with pytest.raises(ValueError, match="2D"):
        s3.set_multi_sweep(0, 2, 8, 8, 0)


# ------------------------------------------------------------------ GPU

# one-sweep tile variant with the same lanes per node as each multi-sweep
# kernel variant (cells per lane 1 / 2 / 4): variants 0, 8, 9
_ONE_SWEEP_PARTNER = {1: 0, 2: 8, 4: 9}
_KERNEL_CPL = {0: 1, 1: 2, 2: 4, 3: 2, 4: 4, 5: 1}


def _gpu_session(c, bsc=None):
    from egg_amd.smoothing import CppStructuredSweepSession

    return CppStructuredSweepSession(c.ctx, bsc if bsc is not None else c.bsc, c.X, device="gpu")


def _rel(a, b):
    """Largest coordinate difference over the grid's extent."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.abs(a - b).max() / max(1.0, float((a.max(0) - a.min(0)).max())))


# FP32 rounding noise after 10-20 sweeps: ~1e-7 of the extent (measured
# 6e-7 absolute on a 3.25-wide grid whose nodes move ~5e-4 per sweep); a
# neighbour one sweep stale would differ by ~1e-4
TOL = 1e-6


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["ogrid33", "star5", "tee", "batch3"])
@pytest.mark.parametrize("kernel", [0, 1, 2, 3, 4, 5])
def test_multi_equals_single_launches(name, kernel):
    c = CASES[name]()
    k, ty, tx = (3, 8, 16) if kernel % 2 else (4, 16, 8)
    print(f"multisweep: {name} kernel {kernel} k {k} tile {ty}x{tx}", flush=True)
    ref = _gpu_session(c)
    ref.core.set_tile_variant(_ONE_SWEEP_PARTNER[_KERNEL_CPL[kernel]])
    ref.core.set_graph_chunk(0)
    e0, m0 = ref.run(14, omega=0.8, report_every=7)
    s = _gpu_session(c)
    s.core.set_tile_variant(_ONE_SWEEP_PARTNER[_KERNEL_CPL[kernel]])
    s.core.set_graph_chunk(0)
    s.core.set_multi_sweep(0, k, ty, tx, kernel)
    e1, m1 = s.run(14, omega=0.8, report_every=7)  # launches of k, then the cadence's remainder
    assert _rel(s.get_X(), ref.get_X()) <= TOL
    np.testing.assert_allclose(e1, e0, rtol=1e-6)
    np.testing.assert_allclose(m1, m0, rtol=1e-4)
    _, launches, _ = s.core.last_timing()
    assert launches == 2 * -(-7 // k)


@pytest.mark.gpu
def test_multi_singular_hosts_and_untangle():
    g = grids.tangle(grids.ogrid_2d(33, 17), seed=4)
    c = Case(g)
    bsc = _forced_hosts(g)
    print("multisweep: forced singular hosts, untangle", flush=True)
    runs = []
    for k in (0, 5):
        s = _gpu_session(c, bsc)
        s.core.set_graph_chunk(0)
        if k:
            s.core.set_multi_sweep(0, k, 8, 8, 0)
        _, m_before = s.core.energy(phase="untangle", delta=0.0)
        delta = 2.0 * abs(m_before)
        e, m = s.run(10, phase="untangle", delta=delta, omega=0.5, report_every=10)
        runs.append((s.get_X(), e, m))
    assert _rel(runs[0][0], runs[1][0]) <= TOL
    np.testing.assert_allclose(runs[1][1], runs[0][1], rtol=1e-6)


@pytest.mark.gpu
def test_multi_in_captured_graphs():
    """Graph chunks of 10 sweeps as launches of 3, 3, 3, 1 (an even count)
    and of 4, 4, 2 (odd: the chunk ends in the other buffer)."""
    c = CASES["ogrid33"]()
    ref = _gpu_session(c)
    ref.core.set_graph_chunk(0)
    ref.run(40, omega=0.8, report_every=20)
    for k in (3, 4):
        print(f"multisweep: graph chunks of 10, k {k}", flush=True)
        s = _gpu_session(c)
        s.core.set_graph_chunk(10)
        s.core.set_multi_sweep(0, k, 16, 8, 0)
        s.run(40, omega=0.8, report_every=20)
        assert _rel(s.get_X(), ref.get_X()) <= TOL


@pytest.mark.gpu
def test_multi_fas_levels():
    """FAS V(4,4) cycles with multi-sweep launches on the fine level and the
    coarse levels agree with one-sweep launches."""
    c = fas_two_blocks(res=32)
    out = []
    for multi in (False, True):
        s = _gpu_session(c)
        s.core.set_graph_chunk(0)
        s.run_fas(1, nu_pre=4, nu_post=4, nu_coarse=8, omega=0.8, max_levels=3)  # prepares the levels
        s.core.set_X(np.ascontiguousarray(c.X).ravel())  # start again from the input
        if multi:
            s.core.set_multi_sweep(0, 4, 16, 8, 0)
            s.core.set_multi_sweep(1, 4, 8, 8, 0)
            s.core.set_multi_sweep(2, 8, 16, 16, 0)
        print(f"multisweep: FAS, multi-sweep levels {multi}", flush=True)
        e, m = s.run_fas(2, nu_pre=4, nu_post=4, nu_coarse=8, omega=0.8, max_levels=3)
        out.append((s.get_X(), e))
    assert _rel(out[0][0], out[1][0]) <= 10 * TOL
    np.testing.assert_allclose(out[1][1], out[0][1], rtol=1e-6)


@pytest.mark.gpu
def test_level_tuner_keeps_the_state_and_picks_a_setting():
    """The per-level multi-sweep tuner leaves the fine state as it was, and
    the cycles after it agree with one-sweep cycles."""
    c = fas_two_blocks(res=32)
    ref = _gpu_session(c)
    e0, _ = ref.run_fas(3, nu_pre=4, nu_post=4, nu_coarse=8, omega=0.8, max_levels=3)
    s = _gpu_session(c)
    s.run_fas(1, nu_pre=4, nu_post=4, nu_coarse=8, omega=0.8, max_levels=3)
    s.core.set_X(np.ascontiguousarray(c.X).ravel())
    X0 = s.get_X()
    cands = [(2, 8, 8, 5), (4, 8, 8, 1), (8, 16, 16, 0), (9, 8, 8, 0)]
    print("multisweep: level tuner", flush=True)
    ms = s.core.autotune_levels_multi(4, 0.8, cands, 3)
    assert ms.shape == (2, 1 + len(cands))
    assert np.all(ms[:, 0] > 0) and np.all(ms[:, 4] == -1.0)  # the last candidate is not buildable (k > 8)
    np.testing.assert_array_equal(s.get_X(), X0)
    e1, _ = s.run_fas(3, nu_pre=4, nu_post=4, nu_coarse=8, omega=0.8, max_levels=3)
    # the picks may fold a patch with other multiply-add fusion (cells per
    # lane 2 or 4): the bound of the one-sweep cells-per-lane FAS comparison
    assert _rel(s.get_X(), ref.get_X()) <= 10 * TOL
    np.testing.assert_allclose(e1, e0, rtol=1e-5)
