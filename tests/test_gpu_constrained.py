"""Constrained boundary DOFs on the MI355X: nodes that slide on a segment, a
circle, an ellipse or a rail (the group's "entities"; reference
src/sweep.hpp:1285-1403 and :1717-1876, src/geometry.hpp:288-464) against
the FP64 oracle on the same wire: a tangent-reduced Newton step damped by
omega, trials projected onto the entity (a rail releases its node beyond its
ends: full-space step, no projection), the free DOFs' accept rule.

Bar: positions within 1e-5 of the domain extent, energies within 2e-3 per
report; constrained nodes on their entity to FP32 rounding (segments along
an axis exactly)."""

from __future__ import annotations

import numpy as np
import pytest

from egg_amd import grids
from tests.cases import Case

pytestmark = pytest.mark.gpu

POS_TOL = 1e-5
TRAJ_RTOL = 2e-3


def _session(case):
    from egg_amd.smoothing import CppStructuredSweepSession

    return CppStructuredSweepSession(case.ctx, case.bsc, case.X)


def _extent(c):
    return float((c.X.max(0) - c.X.min(0)).max())


def _annulus(n=33, m=17, seed=1, **kw):
    g, ents = grids.annulus_2d(n, m, seed=seed)
    return Case(g, entities=ents, **kw), np.array(sorted(ents))


def _slab(n=17, rail=False):
    """A single block whose bottom and top edges (corners excluded) slide on
    their segments; `rail`: on rails instead, the bottom one covering only
    x in [0.2, 0.8], so the nodes beyond it are released."""
    g = grids.single_block_2d(n, seed=4, amp=0.25)
    ids = np.asarray(g.blocks[0])
    free = g.free.copy()
    X = g.X.copy()
    rng = np.random.default_rng(2)
    ents = {}
    for j, y in ((0, 0.0), (n - 1, 1.0)):
        for i in range(1, n - 1):
            node = int(ids[i, j])
            free[node] = True
            X[node, 0] += 0.3 / (n - 1) * rng.uniform(-1, 1)
            if rail:
                ents[node] = ("linerail", (0.2, y, 0.8, y) if j == 0 else (0.0, y, 1.0, y))
            else:
                ents[node] = ("lineseg", (0.0, y, 1.0, y))
    return Case(grids.Grid(2, g.blocks, X, free), entities=ents), np.array(sorted(ents))


def _make(name):
    if name == "annulus":
        return _annulus()
    if name == "ellipse":
        g, ents = grids.annulus_2d(33, 17, seed=1, aspect=0.6)
        return Case(g, entities=ents), np.array(sorted(ents))
    return _slab(rail=name == "rail")


@pytest.mark.parametrize("variant", [0, 8, 12, 13])
@pytest.mark.parametrize("make", ["annulus", "slab", "ellipse", "rail"])
def test_constrained_parity(oracle, make, variant):
    c, con = _make(make)
    s = _session(c)
    s.core.set_tile_variant(variant)
    e, m = s.run(20, omega=0.8, report_every=5)
    Xo, eo, mo = oracle.jacobi(2, "barrier", 0.0, 0.8, 20, c.X, c.patches(oracle), c.ctx.energy_stencil,
                               report_every=5)
    X = s.get_X()
    assert np.abs(X - Xo).max() / _extent(c) < POS_TOL
    np.testing.assert_allclose(e, eo, rtol=TRAJ_RTOL)
    if make == "annulus":
        np.testing.assert_allclose(np.linalg.norm(X[con], axis=1), 0.5, rtol=1e-6)
    elif make == "ellipse":
        np.testing.assert_allclose(np.linalg.norm(X[con] / (0.5, 0.3), axis=1), 1.0, rtol=1e-6)
    elif make == "slab":
        assert np.all(np.isin(X[con, 1], (0.0, 1.0)))
        assert np.all((X[con, 0] >= 0.0) & (X[con, 0] <= 1.0))
    else:  # rails: on them where the foot is inside, released beyond
        on = (X[con, 0] >= 0.2) & (X[con, 0] <= 0.8) | (X[con, 1] > 0.5)
        assert np.all(np.isin(X[con[on], 1], (0.0, 1.0)))
        assert np.any(~on)
    assert np.abs(X[con] - c.X[con]).max() > 1e-4  # they moved


def test_constrained_untangle_and_weights(oracle):
    """The untangle phase and weighted samples take the same constrained
    step (the constrained kernel's weighted instantiation)."""
    c, con = _annulus(seed=5)
    e, _ = (s := _session(c)).run(10, phase="untangle", delta=1e-3, omega=0.5, report_every=5)
    Xo, eo, _ = oracle.jacobi(2, "untangle", 1e-3, 0.5, 10, c.X, c.patches(oracle), c.ctx.energy_stencil,
                              report_every=5)
    assert np.abs(s.get_X() - Xo).max() / _extent(c) < POS_TOL
    np.testing.assert_allclose(e, eo, rtol=TRAJ_RTOL)
    g, ents = grids.annulus_2d(33, 17, seed=1)
    ns = Case(g).ctx.energy_stencil["num_samples"]
    cw = Case(g, entities=ents, weight=np.random.default_rng(3).uniform(0.5, 1.5, ns))
    ew, _ = (sw := _session(cw)).run(10, omega=0.8, report_every=5)
    Xo, eo, _ = oracle.jacobi(2, "barrier", 0.0, 0.8, 10, cw.X, cw.patches(oracle), cw.ctx.energy_stencil,
                              report_every=5)
    assert np.abs(sw.get_X() - Xo).max() / _extent(cw) < POS_TOL
    np.testing.assert_allclose(ew, eo, rtol=TRAJ_RTOL)


def test_constrained_graphs_and_fas():
    """Graph-replayed sweeps give the bits of single launches with the
    constrained launch in each sweep; FAS keeps the constrained DOFs out of
    its hierarchy and corrections, so they stay on the circle while the
    energy falls."""
    c, con = _annulus()
    a, b = _session(c), _session(c)
    b.core.set_graph_chunk(8)
    ea, _ = a.run(24, omega=0.8, report_every=8)
    eb, _ = b.run(24, omega=0.8, report_every=8)
    np.testing.assert_array_equal(eb, ea)
    np.testing.assert_array_equal(b.get_X(), a.get_X())
    f = _session(c)
    assert len(f.mg_levels()) >= 2
    e, m = f.run_fas(6, nu_pre=2, nu_post=2, omega=0.8, max_levels=3)
    X = f.get_X()
    np.testing.assert_allclose(np.linalg.norm(X[con], axis=1), 0.5, rtol=1e-6)
    assert np.all(m > 0) and e[-1] < e[0]
