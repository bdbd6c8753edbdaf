"""The multi-block control front end (build_control_topology) and the
oracle's multi-block reduced Gauss-Newton loop (oracle/control.py
run_control_topo), pinned by what the mathematics fixes: seam controls
shared across blocks keep every shared node watertight for any reduced
step (reference tests/smoothing/test_control_topo_cpp.py:108-115 checks
watertight_mismatch after a run); the reduced gradient and Hessian are the
derivatives of E(q) = E_fine(X(C0 + R (q - q0))) + 1/2 sum w (P C)^2 by
finite differences (tangent rows and penalty rows included); the
assembled-Hessian path equals the per-sample contraction; a one-net
topology whose reduction selects the interior controls runs the
single-block loop step for step; and the loop meets the reference's
acceptance criteria (monotone, converged, valid, no worse than the
converged node smoother by more than 5 %)."""

from __future__ import annotations

import numpy as np
import pytest

from egg_amd import grids
from egg_amd.smoothing import build_sweep_context
from egg_amd.smoothing.control import build_control_topology, topology_mismatch
from tests.cases import block_grid, lr_pair, ogrid

TEE = [(0.0, 0.0, 9, 17, 0.25), (2.0, 0.0, 9, 17, 0.25), (2.0, 4.0, 9, 9, 0.25)]


def wavy_pair(n=9, wave=0.3, seed=3):
    """The two-block pair under a smooth vertical wave that bends its lower
    and upper boundaries, so that the optimum is not the uniform lattice."""
    c = lr_pair(n, seed=seed)
    x, y = c.grid.X.T
    c.grid.X = c.grid.X + np.stack([0.0 * x, wave * np.sin(0.5 * np.pi * x) * (0.5 + y)], -1)
    return c


def periodic_ring(nt=25, nr=9):
    """One block closing on itself: an annulus whose last angular row of
    nodes is its first (a periodic seam inside one block)."""
    ids = np.arange(nt * nr).reshape(nt, nr)
    ids[-1, :] = ids[0, :]
    used, ids = np.unique(ids, return_inverse=True)
    ids = ids.reshape(nt, nr).astype(np.int32)
    th = -2.0 * np.pi * np.arange(nt) / (nt - 1)
    T, Rr = np.meshgrid(th, np.linspace(0.5, 1.5, nr), indexing="ij")
    X = np.zeros((used.size, 2))
    X[ids.reshape(-1)] = np.stack([Rr * np.cos(T), Rr * np.sin(T)], -1).reshape(-1, 2)
    free = np.ones(used.size, bool)
    free[ids[:, 0]] = False
    free[ids[:, -1]] = False
    rng = np.random.default_rng(8)
    X[free] += rng.uniform(-0.02, 0.02, (int(free.sum()), 2))
    return grids.Grid(2, [ids], X, free)


CASES = {
    "periodic_ring": (periodic_ring, [(9, 5)], {}),
    "lr_pair": (lambda: lr_pair(9).grid, [(5, 5)] * 2, {}),
    "lr_pair_c1": (lambda: lr_pair(9).grid, [(5, 5)] * 2, {"c1": True}),
    "lr_pair_pen": (lambda: lr_pair(9).grid, [(5, 5)] * 2, {"seam_weight": 10.0}),
    "ogrid17": (lambda: ogrid(17, 9).grid, [(6, 6)] + [(6, 4)] * 4, {}),
    "tee_subset": (lambda: block_grid(TEE), [(5, 6), (5, 5)], {"net_blocks": [1, 2]}),
    "cubed5": (lambda: grids.cubed_sphere_3d(5), [(4, 4, 4)] * 7, {}),
}


def _R(wire, dim):
    from oracle import control as OC

    return OC.topo_operators(wire, dim)[0]


@pytest.mark.parametrize("name", list(CASES))
def test_builder_is_watertight_for_any_step(name):
    make, shapes, kw = CASES[name]
    g = make()
    wire, info = build_control_topology(g.blocks, g.X, shapes, **kw)
    R = _R(wire, g.d)
    C0 = np.asarray(wire["C0"])
    assert topology_mismatch(wire, g.blocks, C0) < 1e-12
    rng = np.random.default_rng(2)
    for _ in range(3):
        C = C0 + R @ (0.05 * rng.standard_normal(R.shape[1]))
        assert topology_mismatch(wire, g.blocks, C) < 1e-12
    # C0 = R q0 on every row R reaches; rows without entries are the fixed controls
    reached = np.abs(R).sum(1) > 0
    np.testing.assert_allclose((R @ info["q0"])[reached], C0[reached], atol=1e-14)
    fixed = np.repeat(info["root_of"] == -1, g.d)
    assert np.array_equal(~reached, fixed)
    # each net reproduces its block's nodes on faces no other net shares and
    # is pure spline (b = 0) on the seams
    from oracle import control as OC

    off = 0
    shared = {}
    for e in wire["blocks"]:
        for gid in np.asarray(g.blocks[e["block"]]).reshape(-1).tolist():
            shared[gid] = shared.get(gid, 0) + 1
    for e in wire["blocks"]:
        M = OC.block_map(e, g.d)
        ids = np.asarray(g.blocks[e["block"]])
        b = np.asarray(e["b"]).reshape(ids.shape + (g.d,))
        Xb = (M @ C0[off:off + M.shape[1] * g.d].reshape(-1, g.d)).reshape(ids.shape + (g.d,)) + b
        for k in range(g.d):
            for side in (0, -1):
                face = np.take(ids, side, axis=k)
                Xf, bf = np.take(Xb, side, axis=k), np.take(b, side, axis=k)
                seam = sum(frozenset(face.reshape(-1).tolist()) == frozenset(np.take(np.asarray(g.blocks[o["block"]]),
                           s2, axis=k2).reshape(-1).tolist()) for o in wire["blocks"] for k2 in range(g.d)
                           for s2 in (0, -1)) >= 2
                if seam:
                    on_outer = np.zeros(face.shape, bool)  # seam nodes on the domain boundary keep their mismatch
                    for k3 in range(g.d):
                        for s3 in (0, -1):
                            f3 = np.take(ids, s3, axis=k3)
                            if k3 != k and sum(frozenset(f3.reshape(-1).tolist()) == frozenset(
                                    np.take(np.asarray(g.blocks[o["block"]]), s4, axis=k4).reshape(-1).tolist())
                                    for o in wire["blocks"] for k4 in range(g.d) for s4 in (0, -1)) < 2:
                                on_outer |= np.isin(face, f3)
                    np.testing.assert_allclose(bf[~on_outer], 0.0, atol=1e-14)
                else:
                    np.testing.assert_allclose(Xf, g.X[face], atol=1e-12)
        off += M.shape[1] * g.d


def test_builder_rejects_unmatched_seams():
    g = lr_pair(9).grid
    with pytest.raises(ValueError, match="seam"):
        build_control_topology(g.blocks, g.X, [(5, 5), (5, 6)])


def test_assembled_hessian_equals_the_contraction():
    from oracle import control as OC

    g = lr_pair(9).grid
    es = build_sweep_context(g.blocks, g.free, 2).energy_stencil
    wire, _ = build_control_topology(g.blocks, g.X, [(5, 5)] * 2)
    Mg = OC.topo_map(wire, g.blocks, len(g.X), 2)[0]
    G1, A1, P1 = OC.reduced_system(2, "barrier", g.X, es, Mg)
    G2, A2, P2 = OC.reduced_system_sparse(2, "barrier", g.X, es, Mg)
    np.testing.assert_allclose(G2, G1, atol=1e-12)
    np.testing.assert_allclose(A2, A1, atol=1e-12 * np.abs(A1).max())
    np.testing.assert_allclose(P2, P1, atol=1e-12)


@pytest.mark.parametrize("kw", [{"c1": True}, {"seam_weight": 10.0}])
def test_reduced_gradient_and_hessian_are_derivatives_of_E_of_q(kw, oracle):
    from oracle import control as OC

    g = lr_pair(9).grid
    es = build_sweep_context(g.blocks, g.free, 2).energy_stencil
    wire, info = build_control_topology(g.blocks, g.X, [(5, 5)] * 2, **kw)
    Mg, bg, owned = OC.topo_map(wire, g.blocks, len(g.X), 2)
    R, Pm, w = OC.topo_operators(wire, 2)
    C0, q0 = np.asarray(wire["C0"]), info["q0"]

    def state(q):
        C = C0 + R @ (q - q0)
        X = g.X.astype(np.float64).copy()
        X[owned] = (Mg @ C.reshape(-1, 2) + bg)[owned]
        return X, C

    def E_of(q):
        X, C = state(q)
        r = Pm @ C
        return oracle.energy_mindet(2, "barrier", 0.0, es, X)[0] + 0.5 * float(np.sum(w * r * r))

    X, C = state(q0)
    Gq, Aq, _ = OC.topo_system(2, "barrier", X, es, Mg, R, Pm, w, C)
    eps = 1e-6
    rng = np.random.default_rng(7)
    for j in rng.choice(len(q0), 5, replace=False):
        qp, qm = q0.copy(), q0.copy()
        qp[j] += eps
        qm[j] -= eps
        assert Gq[j] == pytest.approx((E_of(qp) - E_of(qm)) / (2 * eps), rel=1e-6, abs=1e-9)
        Gp = OC.topo_system(2, "barrier", *state(qp)[:1], es, Mg, R, Pm, w, state(qp)[1])[0]
        Gm = OC.topo_system(2, "barrier", *state(qm)[:1], es, Mg, R, Pm, w, state(qm)[1])[0]
        np.testing.assert_allclose(Aq[:, j], (Gp - Gm) / (2 * eps), rtol=1e-5, atol=1e-7)
    np.testing.assert_allclose(Aq, Aq.T, atol=1e-12 * np.abs(Aq).max())


def test_one_net_topology_runs_the_single_block_loop():
    from oracle import control as OC
    from egg_amd.smoothing.control import (boolean_sum_offset, boundary_ctrl_mask, build_control_wire, control_map,
                                           fit_control_net)

    g = grids.wavy_block_2d(11)
    es = build_sweep_context(g.blocks, g.free, 2).energy_stencil
    cm = control_map((11, 11), (6, 6))
    X0 = g.X.reshape(11, 11, 2)
    C0 = fit_control_net(cm, X0)
    single = build_control_wire(cm, C0, ~boundary_ctrl_mask((6, 6)), boolean_sum_offset(cm, C0, X0))
    free = np.asarray(single["free_idx"])
    nq = len(free) * 2
    rows = np.full(36 * 2, -1)
    for j, f in enumerate(free):
        rows[2 * f:2 * f + 2] = [2 * j, 2 * j + 1]
    r_off = np.concatenate([[0], np.cumsum(rows >= 0)]).astype(np.int32)
    topo = {"blocks": [{k: single[k] for k in ("block", "ctrl_shape", "k_supp", "b", "off0", "wt0", "off1", "wt1")}],
            "C0": single["C0"], "nq": nq, "r_off": r_off, "r_col": rows[rows >= 0].astype(np.int32),
            "r_coef": np.ones(nq)}
    _, Cs, rs = OC.run_control(2, "barrier", g.X, es, single, g.blocks[0], n_outer=30)
    _, Ct, rt = OC.run_control_topo(2, "barrier", g.X, es, topo, g.blocks, n_outer=30)
    assert rt["iters"] == rs["iters"] and rt["converged"] == rs["converged"]
    np.testing.assert_allclose(rt["energies"], rs["energies"], rtol=1e-12)
    np.testing.assert_allclose(Ct.reshape(-1, 2), Cs, atol=1e-10)


def test_oracle_topology_loop_meets_the_acceptance_criteria(oracle):
    from oracle import control as OC

    c = wavy_pair(9)
    g, es = c.grid, c.ctx.energy_stencil
    wire, _ = build_control_topology(g.blocks, g.X, [(6, 6)] * 2)
    X, C, rep = OC.run_control_topo(2, "barrier", g.X, es, wire, g.blocks, n_outer=30)
    e = np.asarray(rep["energies"])
    assert rep["converged"] and np.all(np.diff(e) <= 1e-12) and np.all(np.asarray(rep["mindets"]) > 0)
    assert topology_mismatch(wire, g.blocks, C) < 1e-12
    Xn, _ = oracle.gauss_seidel(2, "barrier", g.X, c.patches(oracle))
    e_node = oracle.energy_mindet(2, "barrier", 0.0, es, Xn)[0]
    assert e[-1] <= 1.05 * e_node
