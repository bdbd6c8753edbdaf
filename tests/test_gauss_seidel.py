"""The converged state checked by a method independent of the block-Jacobi
restatement: the reference's sequential node baseline (in-place
Gauss-Seidel Newton, tests/smoothing/test_control_ref.py:95-135, restated
as oracle o_gauss_seidel). It shares the patch energy with the sweep but
nothing of its update (no frozen snapshot, no omega, its own solve and
halving ladder), so block-Jacobi, FAS and it agreeing on the minimiser pins
the minimiser, not the sweep's composition."""

from __future__ import annotations

import numpy as np
import pytest

from egg_amd import grids
from tests.cases import Case


def test_gauss_seidel_reaches_the_uniform_lattice(oracle):
    """A perturbed Cartesian block relaxes to the uniform lattice (the
    analytic minimiser of the shape metric with identity target)."""
    g = grids.single_block_2d(13, seed=5, amp=0.3)
    c = Case(g)
    X, n = oracle.gauss_seidel(2, "barrier", g.X, c.patches(oracle))
    ids = np.asarray(g.blocks[0])
    h = 1.0 / 12
    want = np.stack(np.meshgrid(np.arange(13) * h, np.arange(13) * h, indexing="ij"), -1).reshape(-1, 2)
    np.testing.assert_allclose(X[ids.reshape(-1)], want, rtol=0, atol=1e-10)
    assert n < 10000


@pytest.mark.parametrize("name", ["ogrid", "star5", "cube"])
def test_gauss_seidel_and_jacobi_share_the_minimiser(oracle, name):
    from tests.cases import star

    g = {"ogrid": grids.ogrid_2d(9, 5), "star5": star(k=5, n=7).grid,
         "cube": grids.single_block_3d(5, seed=2, amp=0.2)}[name]
    c = Case(g)
    P = c.patches(oracle)
    Xg, _ = oracle.gauss_seidel(g.d, "barrier", g.X, P)
    Xj, e, _ = oracle.jacobi(g.d, "barrier", 0.0, 0.8, 3000, g.X, P, c.ctx.energy_stencil)
    eg = oracle.energy_mindet(g.d, "barrier", 0.0, c.ctx.energy_stencil, Xg)[0]
    assert abs(eg - e[-1]) <= 1e-10 * max(1.0, abs(eg))
    np.testing.assert_allclose(Xj, Xg, rtol=0, atol=1e-9)


@pytest.mark.gpu
def test_device_converges_to_the_gauss_seidel_minimiser(oracle):
    """The HIP path (FAS V(4,4) cycles, then plain sweeps) on the O-grid
    ends at the Gauss-Seidel minimiser: energy to FP32 resolution, positions
    to 1e-5 of the extent."""
    from egg_amd.smoothing import CppStructuredSweepSession

    g = grids.ogrid_2d(13, 7)
    c = Case(g)
    Xg, n = oracle.gauss_seidel(2, "barrier", g.X, c.patches(oracle))
    eg = oracle.energy_mindet(2, "barrier", 0.0, c.ctx.energy_stencil, Xg)[0]
    print(f"Gauss-Seidel: {n} sweeps, E = {eg:.10g}", flush=True)
    s = CppStructuredSweepSession(c.ctx, c.bsc, g.X, device="gpu")
    s.run_fas(60, nu_pre=4, nu_post=4, omega=0.8)
    e, m = s.run(2000, omega=0.8)
    X = s.get_X()
    assert abs(e[-1] - eg) / eg < 2e-6 and m[-1] > 0
    assert np.abs(X - Xg).max() / 4.0 < 1e-5


def test_pinned_converged_state_is_a_fixed_point(oracle):
    """The FP64 converged state X* of the 129-node O-grid
    (tests/golden/xstar_ogrid129.npy, oracle/pin_energy.py) carries the
    pinned energy E* and is a fixed point of the FP64 sweep: one more
    oracle sweep moves no node by more than 1e-9 of the extent and lowers
    the energy by less than 1e-12 relative; every cell is valid."""
    import json
    from pathlib import Path

    from egg_amd.smoothing import build_sweep_context

    gold = Path(__file__).parent / "golden"
    X = np.load(gold / "xstar_ogrid129.npy", allow_pickle=False)
    e_star = json.loads((gold / "estar_ogrid129.json").read_text())["E_star"]
    g = grids.tangle(grids.ogrid_2d(129, 65), frac=0.02, factor=1.5, seed=2026)
    assert X.shape == g.X.shape
    fixed = ~g.free
    np.testing.assert_array_equal(X[fixed], g.X[fixed])  # the boundary never moved
    ctx = build_sweep_context(g.blocks, g.free, 2)
    e, md = oracle.energy_mindet(2, "barrier", 0.0, ctx.energy_stencil, X)
    assert md > 0 and abs(e - e_star) <= 1e-9 * e_star
    X1, e1, _ = oracle.jacobi(2, "barrier", 0.0, 0.8, 1, X, oracle.materialize_patches(ctx.wire, g.blocks, 2),
                              ctx.energy_stencil)
    assert np.abs(X1 - X).max() <= 1e-9 * 4.0
    assert e - e1[-1] <= 1e-12 * e
