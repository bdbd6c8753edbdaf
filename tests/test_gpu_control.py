"""The device control-net reduced Gauss-Newton smoother against the FP64
oracle loop (oracle/control.py) on the same block, net and options
(reference tests/smoothing/test_control_cpp.py:62-184): the evaluated net,
the error paths, the energy trace, iteration count and accepted state with a
tight PCG (the oracle solves densely), the default inexact-Newton forcing,
the shape_size phase, a 3D block, and the nodal smoother on the state the
control run leaves."""

from __future__ import annotations

import numpy as np
import pytest

from egg_amd import grids
from egg_amd.smoothing import build_structured_context_from_block_maps, build_sweep_context
from egg_amd.smoothing.control import (boolean_sum_offset, boundary_ctrl_mask, build_control_wire, control_map,
                                       fit_control_net)

pytestmark = pytest.mark.gpu


def _problem(g, ctrl):
    shape = tuple(np.shape(g.blocks[0]))
    ctx = build_sweep_context(g.blocks, g.free, g.d)
    cm = control_map(shape, ctrl)
    X0 = g.X.reshape(shape + (g.d,))
    C0 = fit_control_net(cm, X0)
    wire = build_control_wire(cm, C0, ~boundary_ctrl_mask(ctrl), boolean_sum_offset(cm, C0, X0))
    return ctx, cm, wire


def _session(g, ctx, wire):
    from egg_amd.smoothing import CppStructuredSweepSession

    return CppStructuredSweepSession(ctx, build_structured_context_from_block_maps(g.d, g.blocks), g.X, control=wire)


def test_set_C_evaluates_the_map():
    from oracle import control as OC

    g = grids.wavy_block_2d(9)
    ctx, cm, wire = _problem(g, (5, 5))
    s = _session(g, ctx, wire)
    C0 = np.asarray(wire["C0"]).reshape(-1, 2)
    M = OC.block_map(wire, 2)
    b = np.asarray(wire["b"]).reshape(-1, 2)
    # the session starts on the net
    np.testing.assert_allclose(s.get_X(), M @ C0 + b, atol=2e-6)
    C1 = C0 + 0.01 * np.random.default_rng(1).standard_normal(C0.shape)
    s.set_C(C1)
    np.testing.assert_allclose(s.get_X(), M @ C1 + b, atol=2e-6)
    np.testing.assert_array_equal(s.get_C(), C1)
    with pytest.raises(ValueError, match="expected"):
        s.set_C(C1[:-1])


def test_errors():
    from egg_amd.smoothing import CppStructuredSweepSession

    g = grids.wavy_block_2d(9)
    ctx, cm, wire = _problem(g, (5, 5))
    s = _session(g, ctx, wire)
    with pytest.raises(ValueError, match="untangle"):
        s.run_control(1, phase="untangle")
    plain = CppStructuredSweepSession(ctx, build_structured_context_from_block_maps(2, g.blocks), g.X)
    with pytest.raises(ValueError, match="control wire"):
        plain.run_control(1)
    bad = dict(wire)
    bad["off0"] = np.asarray(wire["off0"]) + 3
    with pytest.raises(ValueError, match="leaves the lattice"):
        _session(g, ctx, bad)
    with pytest.raises(ValueError, match="sliding wall"):
        _session(g, ctx, {**wire, "walls": [1]})


def _parity(g, ctrl, kind, oracle, **kw):
    from oracle import control as OC

    ctx, cm, wire = _problem(g, ctrl)
    print(f"control parity: dim {g.d} nodes {np.shape(g.blocks[0])} net {ctrl} {kind}", flush=True)
    Xo, Co, ro = OC.run_control(g.d, kind, g.X, ctx.energy_stencil, wire, g.blocks[0], n_outer=30)
    s = _session(g, ctx, wire)
    rd = s.run_control(30, phase=kind, pcg_max_iter=500, pcg_rtol=1e-12, pcg_forcing=False, **kw)
    e_o, e_d = np.asarray(ro["energies"]), np.asarray(rd["energies"])
    print("oracle", e_o, ro["iters"], "device", e_d, rd["iters"], "pcg", rd["pcg_iters"], flush=True)
    assert np.all(np.isfinite(e_d)) and np.all(np.asarray(rd["mindets"]) > 0)
    assert np.all(np.diff(e_d) <= 1e-6)
    assert abs(rd["iters"] - ro["iters"]) <= 2 and rd["converged"] == ro["converged"]
    n = min(len(e_o), len(e_d))
    np.testing.assert_allclose(e_d[:n], e_o[:n], rtol=1e-9, atol=1e-12 * e_o[0])
    np.testing.assert_allclose(e_d[-1], e_o[-1], rtol=1e-9, atol=1e-12 * e_o[0])
    extent = float((g.X.max(0) - g.X.min(0)).max())
    np.testing.assert_allclose(s.get_X(), Xo, atol=2e-7 * extent)
    np.testing.assert_allclose(s.get_C(), Co, atol=1e-8 * extent)
    # the loop's state is the net's FP64 positions; the store holds their
    # FP32 rounding, whose FP64 energy is the reported one to that rounding
    ec, mc = oracle.energy_mindet(g.d, kind, 0.0, ctx.energy_stencil, s.get_X())
    assert abs(ec - e_d[-1]) <= 1e-6 * abs(ec) and mc > 0
    return s, rd, ctx


def test_parity_2d(oracle):
    _parity(grids.wavy_block_2d(13), (6, 6), "barrier", oracle)


def test_parity_2d_shape_size(oracle):
    _parity(grids.wavy_block_2d(13, seed=3), (7, 6), "shape_size", oracle)


def test_parity_3d(oracle):
    # a perturbed unit cube under a smooth bulge of its upper part, so that
    # the optimum is not the uniform lattice
    g = grids.single_block_3d(9, seed=2, amp=0.25)
    x, y, z = g.X.T
    g.X = g.X + np.stack([0 * x, 0 * y, 0.3 * np.sin(np.pi * x) * np.sin(np.pi * y) * z], -1)
    _parity(g, (5, 5, 5), "barrier", oracle)


def test_default_forcing_and_nodal_interop(oracle):
    from oracle import control as OC

    g = grids.wavy_block_2d(13)
    ctx, cm, wire = _problem(g, (6, 6))
    Xo, Co, ro = OC.run_control(2, "barrier", g.X, ctx.energy_stencil, wire, g.blocks[0], n_outer=30)
    s = _session(g, ctx, wire)
    rd = s.run_control(30)
    assert rd["converged"] and np.all(np.diff(rd["energies"]) <= 1e-6)
    np.testing.assert_allclose(rd["energies"][-1], ro["energies"][-1], rtol=1e-9)
    # the nodal smoother continues from the control state
    e, m = s.run(10, omega=0.8)
    assert np.isfinite(e[-1]) and m[-1] > 0 and e[-1] <= rd["energies"][-1] + 1e-6
