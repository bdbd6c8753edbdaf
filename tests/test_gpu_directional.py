"""Directional soft-energy terms (2D and 3D) on the MI355X (reference
src/directional.hpp, src/sweep.hpp:1488-1510 and :1746-1760, reports
src/sweep.hpp:1921, FAS line search src/fas.hpp:233) against the FP64 oracle
on the same wire: each window's gradient, Gauss-Newton Hessian and
stabiliser on top of the shape metric, its energy with the DOF moved in each
line-search trial, the whole-sample table added to every reported energy.
DOFs with windows take the special stored path (a synthesized interior DOF
is expanded to its stored patch), constrained ones included.

Bar: positions within 1e-5 of the domain extent, energies within 2e-3 per
report; the reported energy equals the FP64 shape total plus the FP64
directional total of the returned FP32 state to 1e-9."""

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


def _block(w=0.5, eps=1e-3, n=17):
    g = grids.single_block_2d(n, seed=4, amp=0.25)
    return Case(g, directional=grids.wall_directional_samples(g, w, eps)), g


@pytest.mark.parametrize("variant", [0, 8, 12, 13])
def test_directional_parity(oracle, variant):
    c, g = _block()
    s = _session(c)
    plain = _session(Case(g)).core.info()
    info = s.core.info()
    assert info["n_synth"] < plain["n_synth"] and info["n_stored"] > plain["n_stored"]
    s.core.set_tile_variant(variant)
    e, m = s.run(20, omega=0.8, report_every=5)
    Xo, eo, mo = oracle.jacobi(2, "barrier", 0.0, 0.8, 20, c.X, c.patches(oracle), c.ctx.energy_stencil,
                               report_every=5)
    X = s.get_X()
    assert np.abs(X - Xo).max() / _extent(c) < POS_TOL
    np.testing.assert_allclose(e, eo, rtol=TRAJ_RTOL)
    X0, _, _ = oracle.jacobi(2, "barrier", 0.0, 0.8, 20, c.X, Case(g).patches(oracle), c.ctx.energy_stencil)
    assert np.abs(X - X0).max() / _extent(c) > 1e-4  # the windows change the result


@pytest.mark.parametrize("variant", [0, 4])
def test_directional_parity_3d(oracle, variant):
    """A 3D block with the samples in its middle plane: windowed interior
    DOFs expanded to their 64 stored occurrences, 64 or 16 lanes per DOF."""
    g = grids.single_block_3d(7, seed=2, amp=0.2)
    c = Case(g, directional=grids.wall_directional_samples(g, 0.5, 1e-3))
    s = _session(c)
    s.core.set_tile_variant(variant)
    e, m = s.run(12, omega=0.8, report_every=4)
    Xo, eo, mo = oracle.jacobi(3, "barrier", 0.0, 0.8, 12, c.X, c.patches(oracle), c.ctx.energy_stencil,
                               report_every=4)
    X = s.get_X()
    assert np.abs(X - Xo).max() / _extent(c) < POS_TOL
    np.testing.assert_allclose(e, eo, rtol=TRAJ_RTOL)
    f = _session(c)
    ef, mf = f.run_fas(3, nu_pre=2, nu_post=2, omega=0.8, max_levels=2)
    assert np.all(mf > 0) and ef[-1] < ef[0]


def test_reported_energy_is_the_composed_objective(oracle):
    """Each report is the shape total plus the directional total (reference
    tests/test_cpp_directional.py:120-133), both in FP64 of the FP32 state."""
    from egg_amd._cpp import cpp_core

    c, g = _block()
    smp = grids.wall_directional_samples(g, 0.5, 1e-3)
    s = _session(c)
    e, m = s.run(7, omega=0.8, report_every=7)
    X = s.get_X().astype(np.float32).astype(np.float64)
    E_shape, M = oracle.energy_mindet(2, "barrier", 0.0, Case(g).ctx.energy_stencil, X.astype(np.float32),
                                      prec="f32")
    E_dir, _ = cpp_core.directional_eval(X, smp.nodes, smp.kind, smp.ref, smp.weight, smp.eps)
    assert E_dir > 0
    assert e[-1] == pytest.approx(E_shape + E_dir, rel=1e-9) and m[-1] == pytest.approx(M, rel=1e-9)


def test_sweeps_align_the_chain():
    """After 100 sweeps from the perturbed block, the chain above the wall is
    closer to parallel, the stems closer to upright and the line triples
    straighter with the terms than under plain sweeps: the directional total
    is below half of plain sweeps' (the shape metric alone also drifts
    towards the uniform grid, which satisfies every term; reference
    tests/test_cpp_directional.py:136-160)."""
    from egg_amd._cpp import cpp_core

    g = grids.single_block_2d(17, seed=4, amp=0.25)
    smp = grids.wall_directional_samples(g, 2.0, 1e-6)
    c = Case(g, directional=smp)
    a, b = _session(c), _session(Case(g))
    ea, ma = a.run(100, omega=0.8)
    b.run(100, omega=0.8)

    def ed(X):
        return cpp_core.directional_eval(X, smp.nodes, smp.kind, smp.ref, smp.weight, smp.eps)[0]

    assert ma[-1] > 0 and ed(a.get_X()) < 0.5 * ed(b.get_X()) and ed(a.get_X()) < 1e-5 * ed(c.X)


def test_directional_on_constrained_wall(oracle):
    """Stems from the annulus' circle-wall nodes to the next ring outward,
    kept along the radius (through axis = the wall tangent), with the wall
    constrained: the constrained kernel folds the windows into its tangent
    Newton step (reference src/sweep.hpp:1746-1760)."""
    from egg_amd.smoothing.directional import KIND_STEM, DirectionalSamples

    g, ents = grids.annulus_2d(33, 17, seed=1)
    con = set(ents)
    rows = []
    for ids in g.blocks:
        ids = np.asarray(ids)
        for i in range(ids.shape[0]):
            if int(ids[i, 0]) in con:
                c0, n1 = int(ids[i, 0]), int(ids[i, 1])
                t = np.array([-g.X[c0, 1], g.X[c0, 0]])
                rows.append((n1, c0, t / np.linalg.norm(t)))
    assert len(rows) > 8
    nodes = np.full((len(rows), 4), -1, np.int32)
    nodes[:, :2] = [r[:2] for r in rows]
    ref = np.array([r[2] for r in rows])
    smp = DirectionalSamples(nodes, np.full(len(rows), KIND_STEM, np.int32), ref, np.zeros_like(ref),
                             np.full(len(rows), 0.5), 1e-3)
    c = Case(g, entities=ents, directional=smp)
    s = _session(c)
    e, m = s.run(15, omega=0.8, report_every=5)
    Xo, eo, mo = oracle.jacobi(2, "barrier", 0.0, 0.8, 15, c.X, c.patches(oracle), c.ctx.energy_stencil,
                               report_every=5)
    X = s.get_X()
    assert np.abs(X - Xo).max() / _extent(c) < POS_TOL
    np.testing.assert_allclose(e, eo, rtol=TRAJ_RTOL)
    conl = np.array(sorted(con))
    np.testing.assert_allclose(np.linalg.norm(X[conl], axis=1), 0.5, rtol=1e-6)


def test_directional_graphs_and_fas():
    """Graph replay gives the bits of single launches (sweeps, and FAS cycles
    whose device line-search ladder adds the directional table), the host
    pair line search agrees with the device ladder bit for bit, and FAS keeps
    every cell valid while the composed energy falls."""
    c, _ = _block()
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


def test_directional_fas_parity(oracle):
    """FAS V-cycles with directional terms against the FAS oracle fed the
    same wire: the windows enter the fine smoothing, the fine gradient the
    restriction reads (the special-DOF gradient pass) and the line search
    (the directional table in the device ladder); DOFs carrying windows are
    no synthesized interior DOFs, so their coarse images do not relax."""
    from oracle import fas

    c, g = _block()
    grp = c.ctx.wire["groups"][0]
    synth = np.zeros(len(g.X), bool)
    synth[np.asarray(grp["dof_idx"])[np.asarray(grp["interior_block"]) >= 0]] = True
    K = int(grp["dir_k"])
    windowed = np.asarray(grp["dir_slot"]).reshape(-1, K).max(axis=1) >= 0
    synth[np.asarray(grp["dof_idx"])[windowed]] = False
    print("directional FAS parity: 4 cycles, 3 levels", flush=True)
    s = _session(c)
    e, m = s.run_fas(4, nu_pre=2, nu_post=2, omega=0.8, max_levels=3)
    Xo, eo, mo = fas.run_fas(2, "barrier", np.asarray(g.blocks[0]), g.free, c.X, n_cycles=4, nu_pre=2, nu_post=2,
                             omega=0.8, max_levels=3, fine_patches=c.patches(oracle),
                             fine_stencil=c.ctx.energy_stencil, synth=synth)
    print("gpu", e, "\noracle", eo)
    assert np.all(m > 0)
    np.testing.assert_allclose(e, eo, rtol=1e-4)
    assert np.abs(s.get_X() - Xo).max() / _extent(c) < POS_TOL
