"""The device multi-block control-net smoother (nets over one stacked
control vector, reduction dC = R dq, penalty rows) against the FP64 oracle
loop (oracle/control.py run_control_topo) on the same topologies and
options (reference tests/smoothing/test_control_topo_cpp.py:62-160: parity
with exact solves on a two-block seam, a junction topology, a penalty, a 3D
topology): the energy trace, iteration count, stacked controls and state
with a tight PCG, watertight seams after the run; the evaluated nets, the
set_C / set_control_b / set_control_reduction / set_control_penalty calls
and the error paths."""

from __future__ import annotations

import numpy as np
import pytest

from egg_amd import grids
from egg_amd.smoothing import build_structured_context_from_block_maps, build_sweep_context
from egg_amd.smoothing.control import build_control_topology, topology_mismatch
from tests.cases import block_grid, ogrid
from tests.test_control_topo import TEE, periodic_ring, wavy_pair

pytestmark = pytest.mark.gpu


def _session(g, ctx, wire):
    from egg_amd.smoothing import CppStructuredSweepSession

    return CppStructuredSweepSession(ctx, build_structured_context_from_block_maps(g.d, g.blocks), g.X, control=wire)


def _eval(wire, g):
    """The oracle's global positions of the wire's nets at C0."""
    from oracle import control as OC

    Mg, bg, owned = OC.topo_map(wire, g.blocks, len(g.X), g.d)
    X = g.X.astype(np.float64).copy()
    X[owned] = (Mg @ np.asarray(wire["C0"]).reshape(-1, g.d) + bg)[owned]
    return X, Mg, bg, owned


def _parity(g, shapes, kind, oracle, relax=0, **kw):
    """relax > 0: start from the oracle's Jacobi sweeps of g (omega 0.8),
    where the reduced Hessian is positive definite; far from the optimum it
    is indefinite, the PCG stops at negative curvature (as the reference's
    does) and the dense solve of the oracle takes another step."""
    from oracle import control as OC

    ctx = build_sweep_context(g.blocks, g.free, g.d)
    if relax:
        pats = oracle.materialize_patches(ctx.wire, g.blocks, g.d)
        X, _, _ = oracle.jacobi(g.d, "barrier", 0.0, 0.8, relax, g.X, pats, ctx.energy_stencil,
                                report_every=relax, threads=8)
        g = grids.Grid(g.d, g.blocks, X, g.free)
    # nodes no net owns keep the FP32 store's values on the device: start
    # both from FP32-representable positions
    g = grids.Grid(g.d, g.blocks, g.X.astype(np.float32).astype(np.float64), g.free)
    wire, _ = build_control_topology(g.blocks, g.X, shapes, **kw)
    print(f"control topo parity: dim {g.d} blocks {len(g.blocks)} nets {len(wire['blocks'])} nq {wire['nq']} "
          f"{kind} {kw}", flush=True)
    Xo, Co, ro = OC.run_control_topo(g.d, kind, g.X, ctx.energy_stencil, wire, g.blocks, n_outer=30)
    s = _session(g, ctx, wire)
    rd = s.run_control(30, phase=kind, pcg_max_iter=2000, pcg_rtol=1e-12, pcg_forcing=False)
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
    C = s.get_C().reshape(-1)
    np.testing.assert_allclose(C, Co, atol=1e-8 * extent)
    assert topology_mismatch(wire, g.blocks, C) < 1e-12 * extent
    # the store holds the FP32 rounding of the FP64 state
    ec, mc = oracle.energy_mindet(g.d, kind, 0.0, ctx.energy_stencil, s.get_X())
    pen = 0.0
    if "p_w" in wire:
        R, Pm, w = OC.topo_operators(wire, g.d)
        pen = 0.5 * float(np.sum(w * (Pm @ C) ** 2))
    assert abs(ec + pen - e_d[-1]) <= 1e-6 * abs(e_d[-1]) and mc > 0
    return s, rd, ctx, wire


def test_parity_two_block_seam(oracle):
    _parity(wavy_pair(9).grid, [(6, 6)] * 2, "barrier", oracle)


def test_parity_two_block_tangent_rows(oracle):
    _parity(wavy_pair(9).grid, [(6, 6)] * 2, "barrier", oracle, c1=True)


def test_parity_two_block_penalty(oracle):
    _parity(wavy_pair(9).grid, [(6, 6)] * 2, "barrier", oracle, seam_weight=100.0)


def test_parity_shape_size(oracle):
    _parity(wavy_pair(9, seed=5).grid, [(6, 5)] * 2, "shape_size", oracle)


def test_parity_junctions(oracle):
    # the O-grid: four rings round a disc, three blocks at each disc corner
    _parity(ogrid(17, 9).grid, [(6, 6)] + [(6, 4)] * 4, "barrier", oracle, relax=300)


def test_parity_blocks_without_nets(oracle):
    # block 0 carries no net: it owns the nodes of its face with block 1,
    # whose net then holds that face's controls fixed
    g = block_grid(TEE)
    x, y = g.X.T
    g.X = g.X + np.stack([0.08 * np.sin(np.pi * y / 3.0) * x, 0.0 * y], -1)
    _parity(g, [(5, 6), (5, 5)], "barrier", oracle, relax=300, net_blocks=[1, 2])


def test_parity_periodic_seam_in_one_block(oracle):
    # the seam's copies and owners sit in the same block
    _parity(periodic_ring(), [(9, 5)], "barrier", oracle, relax=300)


def test_parity_3d(oracle):
    _parity(grids.cubed_sphere_3d(5), [(4, 4, 4)] * 7, "barrier", oracle, relax=300)


def test_state_calls_and_errors():
    from oracle import control as OC

    g = wavy_pair(9).grid
    ctx = build_sweep_context(g.blocks, g.free, 2)
    wire, info = build_control_topology(g.blocks, g.X, [(6, 6)] * 2)
    s = _session(g, ctx, wire)
    X0, Mg, bg, owned = _eval(wire, g)
    np.testing.assert_allclose(s.get_X(), X0, atol=2e-6)
    C0 = np.asarray(wire["C0"]).reshape(-1, 2)
    C1 = C0 + 0.01 * np.random.default_rng(1).standard_normal(C0.shape)
    s.set_C(C1)
    Xe = g.X.astype(np.float64).copy()
    Xe[owned] = (Mg @ C1 + bg)[owned]
    np.testing.assert_allclose(s.get_X(), Xe, atol=2e-6)
    np.testing.assert_array_equal(s.get_C(), C1)
    with pytest.raises(ValueError, match="expected"):
        s.set_C(C1[:-1])
    # a new offset for net 1 moves the nodes that net owns
    s.set_C(C0)
    b1 = np.asarray(wire["blocks"][1]["b"]).copy() + 0.001
    s.set_control_b(1, b1)
    wire_b = {**wire, "blocks": [wire["blocks"][0], {**wire["blocks"][1], "b": b1}]}
    np.testing.assert_allclose(s.get_X(), _eval(wire_b, g)[0], atol=2e-6)
    with pytest.raises(ValueError, match="set_b"):
        s.set_control_b(2, b1)
    s.set_control_b(1, np.asarray(wire["blocks"][1]["b"]))
    # no reduced DOFs: the run reports the start and stops
    nsd = C0.size
    s.set_control_reduction(0, np.zeros(nsd + 1, np.int32), np.zeros(0, np.int32), np.zeros(0))
    r = s.run_control(5)
    assert len(r["energies"]) == 1 and r["iters"] == 0
    s.set_control_reduction(wire["nq"], wire["r_off"], wire["r_col"], wire["r_coef"])
    # a penalty joins the reported energy: 1/2 w (C_0 . C_0) on one row
    e_plain = s.run_control(0)["energies"][0]
    s.set_control_penalty(np.array([0, 1], np.int32), np.array([0], np.int32), np.array([1.0]), np.array([2.0]))
    e_pen = s.run_control(0)["energies"][0]
    assert e_pen == pytest.approx(e_plain + 0.5 * 2.0 * C0[0, 0] ** 2, rel=1e-12)
    s.set_control_penalty(np.zeros(0, np.int32), np.zeros(0, np.int32), np.zeros(0), np.zeros(0))
    assert s.run_control(0)["energies"][0] == pytest.approx(e_plain, rel=1e-15)
    with pytest.raises(ValueError, match="column out of range"):
        s.set_control_penalty(np.array([0, 1], np.int32), np.array([nsd], np.int32), np.array([1.0]),
                              np.array([1.0]))
    bad = {**wire, "r_col": np.asarray(wire["r_col"]).copy()}
    bad["r_col"][0] = wire["nq"]
    with pytest.raises(ValueError, match="column out of range"):
        _session(g, ctx, bad)
    with pytest.raises(ValueError, match="two nets"):
        _session(g, ctx, {**wire, "blocks": [wire["blocks"][0], wire["blocks"][0]],
                          "C0": np.concatenate([wire["C0"][:72], wire["C0"][:72]]),
                          "r_off": np.zeros(145, np.int32), "r_col": np.zeros(0, np.int32), "r_coef": np.zeros(0)})
    with pytest.raises(ValueError, match="sliding-wall"):
        _session(g, ctx, {**wire, "slide": [1]})
    with pytest.raises(ValueError, match="untangle"):
        s.run_control(1, phase="untangle")
    # a reduction with another column count (the tangent-row topology's R and
    # C0) replaces the first: the run is the oracle's on that wire
    from oracle import control as OC2

    w1, _ = build_control_topology(g.blocks, g.X, [(6, 6)] * 2, c1=True)
    assert w1["nq"] != wire["nq"]
    s.set_control_reduction(w1["nq"], w1["r_off"], w1["r_col"], w1["r_coef"])
    s.set_C(np.asarray(w1["C0"]).reshape(-1, 2))
    rq = s.run_control(30, pcg_max_iter=2000, pcg_rtol=1e-12, pcg_forcing=False)
    _, _, ro = OC2.run_control_topo(2, "barrier", g.X, ctx.energy_stencil, w1, g.blocks, n_outer=30)
    assert rq["iters"] == ro["iters"]
    np.testing.assert_allclose(rq["energies"], ro["energies"], rtol=1e-9)
    s.set_control_reduction(wire["nq"], wire["r_off"], wire["r_col"], wire["r_coef"])
    s.set_C(C0)
    # the nodal smoother continues from the control state
    rd = s.run_control(30)
    e, m = s.run(10, omega=0.8)
    assert np.isfinite(e[-1]) and m[-1] > 0 and e[-1] <= rd["energies"][-1] + 1e-6
