"""The benchmark workloads at their stated sizes on the GPU, checked against
the FP64 oracle and by properties that hold at any size:

* 2D O-grid, 129 x 129 disc inside four 129 x 65 ring blocks, 2 % of the
  interior nodes displaced by 1.5 h (seed 2026): untangle, then 150 sweeps
  against 150 FP64 oracle sweeps; FAS V(8,8) on 4 levels reaches 1e-6 of the
  FP64 converged energy E* (tests/golden/estar_ogrid129.json).
* 64 independent multi-block strips of 3-9 blocks, sides 33..513 (28.6 M
  DOFs), 2 % tangled: untangle, FAS V(4,4) on 3 levels keeps every cell
  valid, lowers the energy, reports the FP64 energy of the state it
  returns and reaches the pinned energy floor (estar_batched64.json); 50
  sweeps of the batch match 50 oracle sweeps of one complete grid of it
  (grids are independent).
* 3D cubed sphere, 7 x 65^3, 1 % tangled: untangle, FAS V(4,4) on 3
  levels as above, to 1e-6 of the pinned converged energy
  (estar_cubed_sphere65.json); one sweep matches the oracle on every 40th
  DOF (a Jacobi sweep moves each DOF from frozen neighbours, so a sample
  sweep is exact).
"""

from __future__ import annotations

import json
from pathlib import Path

import numpy as np
import pytest

from egg_amd import grids

pytestmark = pytest.mark.gpu

SEED = 2026


def _build(g):
    from egg_amd.smoothing import (CppStructuredSweepSession, build_structured_context_from_block_maps,
                                   build_sweep_context)

    ctx = build_sweep_context(g.blocks, g.free, g.d)
    bsc = build_structured_context_from_block_maps(g.d, g.blocks)
    return ctx, bsc, CppStructuredSweepSession(ctx, bsc, g.X, device="gpu")


def _extent(X):
    return float((X.max(0) - X.min(0)).max())


def test_ogrid_129_parity_and_fas_convergence(oracle):
    from egg_amd.smoothing import cpp_untangle

    g = grids.tangle(grids.ogrid_2d(129, 65), frac=0.02, factor=1.5, seed=SEED)
    assert grids.min_det(g) < 0
    ctx, bsc, s = _build(g)
    print(f"O-grid 129/65: {g.n_free} DOFs, untangle", flush=True)
    Xu, md, outer, _ = cpp_untangle(ctx, bsc, g.X, session=s)
    assert md > 0 and outer >= 1
    e, m = s.run(150, omega=0.8, report_every=50)
    Xg = s.get_X()
    print("oracle: 150 FP64 sweeps", flush=True)
    Xo, eo, mo = oracle.jacobi(2, "barrier", 0.0, 0.8, 150, Xu, oracle.materialize_patches(ctx.wire, g.blocks, 2),
                               ctx.energy_stencil, report_every=50)
    err = np.abs(Xg - Xo).max() / _extent(Xo)
    print(f"pos {err:.2e} of extent, E {abs(e[-1] - eo[-1]) / eo[-1]:.2e}, md {abs(m[-1] - mo[-1]):.2e}")
    assert err < 1e-6
    np.testing.assert_allclose(e, eo, rtol=5e-7)
    assert abs(m[-1] - mo[-1]) < 1e-9 and m[-1] > 0
    # FAS to the FP64 converged energy
    e_star = json.loads((Path(__file__).parent / "golden" / "estar_ogrid129.json").read_text())["E_star"]
    s.set_X(Xu)
    cycles = 0
    for _ in range(40):
        ef, mf = s.run_fas(10, nu_pre=8, nu_post=8, nu_coarse=16, omega=0.8, max_levels=4)
        cycles += 10
        assert np.all(mf > 0)
        if abs(ef[-1] - e_star) / e_star <= 1e-6:
            break
    print(f"FAS V(8,8): {cycles} cycles, E {ef[-1]:.10g} vs E* {e_star:.10g}")
    assert abs(ef[-1] - e_star) / e_star <= 1e-6 and cycles <= 400


def test_batched_64_grids(oracle):
    from egg_amd.smoothing import build_sweep_context, cpp_untangle

    g, index = grids.batched_2d(64, seed=SEED)
    ctx, bsc, s = _build(g)
    print(f"batched: {len(index)} grids, {g.n_free} DOFs, untangle", flush=True)
    Xu, md, outer, _ = cpp_untangle(ctx, bsc, g.X, session=s)
    assert md > 0
    e, m = s.run_fas(3, nu_pre=4, nu_post=4, omega=0.8, max_levels=3)
    print(f"FAS V(4,4): E {e.tolist()}, min det {m.tolist()}", flush=True)
    assert np.all(m > 0) and np.all(np.diff(e) < 0)
    ec, mc = oracle.energy_mindet(2, "barrier", 0.0, ctx.energy_stencil, s.get_X())
    assert abs(ec - e[-1]) <= 1e-9 * ec and mc == pytest.approx(m[-1], rel=1e-6)
    # on to the pinned energy floor (tests/golden/estar_batched64.json: the
    # strips' optimum has zero energy, E* is the floor FP32 positions set)
    f = json.loads((Path(__file__).parent / "golden" / "estar_batched64.json").read_text())
    ef, mf = s.run_fas(12, nu_pre=4, nu_post=4, omega=0.8, max_levels=3)
    print(f"FAS V(4,4) x15: E {ef[-1]:.6g} vs the floor {f['E_star']:.6g}", flush=True)
    assert np.all(mf > 0) and abs(ef[-1] - f["E_star"]) <= f["convergence_rtol"] * f["E_star"]
    # one complete grid of the batch (grids are independent): the one nearest 300 k DOFs
    frees = [int(g.free[n0:n0 + nn].sum()) for (_, _, n0, nn) in index]
    gi = int(np.argmin([abs(f - 300_000) for f in frees]))
    b0, nb, n0, nn = index[gi]
    blocks = [b - n0 for b in g.blocks[b0:b0 + nb]]
    sub_ctx = build_sweep_context(blocks, g.free[n0:n0 + nn], 2)
    s.set_X(Xu)
    s.run(50, omega=0.8)
    Xg = s.get_X()[n0:n0 + nn]
    print(f"oracle: 50 FP64 sweeps of grid {gi} ({frees[gi]} DOFs)", flush=True)
    Xo, eo, mo = oracle.jacobi(2, "barrier", 0.0, 0.8, 50, Xu[n0:n0 + nn],
                               oracle.materialize_patches(sub_ctx.wire, blocks, 2), sub_ctx.energy_stencil)
    eg, mg = oracle.energy_mindet(2, "barrier", 0.0, sub_ctx.energy_stencil, Xg)
    err = np.abs(Xg - Xo).max() / _extent(Xo)
    print(f"grid {gi}: pos {err:.2e} of extent, E {abs(eg - eo[-1]) / eo[-1]:.2e}")
    assert err < 1e-5 and abs(eg - eo[-1]) / eo[-1] < 5e-4 and mg > 0


def test_cubed_sphere_7x65(oracle):
    from egg_amd.smoothing import cpp_untangle

    g = grids.tangle(grids.cubed_sphere_3d(65), frac=0.01, factor=1.5, seed=SEED)
    ctx, bsc, s = _build(g)
    print(f"cubed sphere: {g.n_free} DOFs, untangle", flush=True)
    Xu, md, outer, _ = cpp_untangle(ctx, bsc, g.X, session=s)
    assert md > 0
    # one sweep on a sample of every 40th DOF against the FP64 oracle
    nd = len(ctx.wire["groups"][0]["dof_idx"])
    sample = np.arange(0, nd, 40)
    patches = oracle.materialize_patches(ctx.wire, g.blocks, 3, dofs=sample)
    no_report = {"num_samples": 0, "gc": np.zeros(0, np.int32), "W_inv": np.eye(3),
                 **{f"gn{k}": np.zeros(0, np.int32) for k in range(3)}, **{f"s{k}": np.zeros(0) for k in range(3)}}
    Xo, _, _ = oracle.jacobi(3, "barrier", 0.0, 0.8, 1, Xu, patches, no_report)
    e1, m1 = s.run(1, omega=0.8)
    dofs = np.asarray(ctx.wire["groups"][0]["dof_idx"])[sample]
    Xg = s.get_X()
    err = np.abs(Xg[dofs] - Xo[dofs]).max() / _extent(Xo)
    print(f"sample of {len(sample)} DOFs: pos {err:.2e} of extent", flush=True)
    assert err < 1e-5
    ec, mc = oracle.energy_mindet(3, "barrier", 0.0, ctx.energy_stencil, Xg)
    assert abs(ec - e1[-1]) <= 1e-9 * ec and mc == pytest.approx(m1[-1], rel=1e-6)
    e, m = s.run_fas(2, nu_pre=4, nu_post=4, omega=0.8, max_levels=3)
    print(f"FAS V(4,4): E {e.tolist()}", flush=True)
    assert np.all(m > 0) and e[-1] < e[0] < e1[-1]
    # FAS from the untangled state to the pinned converged energy (the GPU
    # fixed point, tests/golden/estar_cubed_sphere65.json)
    f = json.loads((Path(__file__).parent / "golden" / "estar_cubed_sphere65.json").read_text())
    s.set_X(Xu)
    cycles = 0
    for _ in range(20):
        ef, mf = s.run_fas(50, nu_pre=4, nu_post=4, omega=0.8, max_levels=3)
        cycles += 50
        assert np.all(mf > 0)
        if abs(ef[-1] - f["E_star"]) <= f["convergence_rtol"] * f["E_star"]:
            break
    print(f"FAS V(4,4) to E*: {cycles} cycles, E {ef[-1]:.10g} vs {f['E_star']:.10g}", flush=True)
    assert abs(ef[-1] - f["E_star"]) <= f["convergence_rtol"] * f["E_star"] and cycles <= 1000
