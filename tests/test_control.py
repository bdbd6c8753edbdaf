"""The control-net front end (egg_amd.smoothing.control) and the oracle's
reduced Gauss-Newton loop (oracle/control.py), pinned by what the
mathematics fixes: the basis against scipy's B-spline design matrix, the fit
and the Boolean-sum offset by exactness on their own spans (reference
tests/smoothing/test_control_ref.py:171-195), the reduced gradient and
Hessian by finite differences of E(C) through the map (:214-245), the
Jacobi blocks by the C oracle's golden-pinned per-DOF patch Hessians, and
the loop by the reference's acceptance criteria (:248-277): monotone,
within 1 % of its end in at most 30 iterations, no worse than 1.05 times
the converged node smoother, and a net that stays valid when resampled at
2x and 4x."""

from __future__ import annotations

import numpy as np
import pytest

from egg_amd import grids
from egg_amd.smoothing import build_sweep_context
from egg_amd.smoothing.control import (boolean_sum_offset, boundary_ctrl_mask, bspline_matrix, build_control_wire,
                                       clamped_knots, control_map, fit_control_net)


@pytest.mark.parametrize("n_node,n_ctrl,p", [(11, 6, 3), (13, 4, 3), (9, 5, 2), (17, 9, 3), (7, 7, 1)])
def test_basis_is_scipy_design_matrix(n_node, n_ctrl, p):
    from scipy.interpolate import BSpline

    u = np.linspace(0.0, 1.0, n_node)
    ref = BSpline.design_matrix(u, clamped_knots(n_ctrl, p), p).toarray()
    np.testing.assert_allclose(bspline_matrix(n_node, n_ctrl, p), ref, atol=1e-14)


def test_fit_recovers_a_net_in_span_and_offset_fixes_the_boundary():
    rng = np.random.default_rng(4)
    cm = control_map((11, 9), (6, 5))
    C = np.stack(np.meshgrid(np.linspace(0, 1, 6), np.linspace(0, 1, 5), indexing="ij"), -1)
    C += 0.05 * rng.standard_normal(C.shape)
    X = cm.prolong(C)
    np.testing.assert_allclose(cm.prolong(fit_control_net(cm, X)), X, atol=1e-12)
    for shape, cshape in (((9, 9), (5, 5)), ((7, 6, 5), (5, 4, 4))):
        cm = control_map(shape, cshape)
        X0 = rng.standard_normal(shape + (len(shape),))
        Cf = fit_control_net(cm, X0)
        Xb = cm.prolong(Cf) + boolean_sum_offset(cm, Cf, X0)
        bnd = boundary_ctrl_mask(shape)
        np.testing.assert_allclose(Xb[bnd], X0[bnd], atol=1e-12)
        # the offset is the transfinite fill of the boundary mismatch: zero
        # for data the net reproduces
        Xn = cm.prolong(Cf)
        np.testing.assert_allclose(boolean_sum_offset(cm, Cf, Xn), 0.0, atol=1e-12)


def _wavy(n=13, ctrl=(6, 6), seed=11):
    from oracle import control as OC

    g = grids.wavy_block_2d(n, seed=seed)
    ctx = build_sweep_context(g.blocks, g.free, 2)
    cm = control_map((n, n), ctrl)
    X0 = g.X.reshape(n, n, 2)
    C0 = fit_control_net(cm, X0)
    wire = build_control_wire(cm, C0, ~boundary_ctrl_mask(ctrl), boolean_sum_offset(cm, C0, X0))
    return g, ctx, cm, wire, OC


def test_wire_reproduces_the_map():
    g, ctx, cm, wire, OC = _wavy(9, (5, 5))
    M = OC.block_map(wire, 2)
    C0 = np.asarray(wire["C0"]).reshape(-1, 2)
    np.testing.assert_allclose(M @ C0, cm.prolong(C0.reshape(5, 5, 2)).reshape(-1, 2), atol=1e-13)


def test_reduced_gradient_and_hessian_are_derivatives_of_E_of_C(oracle):
    g, ctx, cm, wire, OC = _wavy(9, (5, 5))
    es = ctx.energy_stencil
    M = OC.block_map(wire, 2)
    b = np.asarray(wire["b"]).reshape(-1, 2)
    free = np.asarray(wire["free_idx"])
    C0 = np.asarray(wire["C0"]).reshape(-1, 2)

    def X_of(C):
        return M @ C + b

    def E_of(C):
        return oracle.energy_mindet(2, "barrier", 0.0, es, X_of(C))[0]

    G, A, P = OC.reduced_system(2, "barrier", X_of(C0), es, M[:, free])
    eps = 1e-6
    rng = np.random.default_rng(5)
    for f in rng.choice(len(free), 4, replace=False):
        for ax in range(2):
            Cp, Cm = C0.copy(), C0.copy()
            Cp[free[f], ax] += eps
            Cm[free[f], ax] -= eps
            assert G[f, ax] == pytest.approx((E_of(Cp) - E_of(Cm)) / (2 * eps), rel=1e-6, abs=1e-9)
            Gp = OC.reduced_system(2, "barrier", X_of(Cp), es, M[:, free])[0]
            Gm = OC.reduced_system(2, "barrier", X_of(Cm), es, M[:, free])[0]
            np.testing.assert_allclose(A[:, 2 * f + ax], ((Gp - Gm) / (2 * eps)).reshape(-1), rtol=1e-5, atol=1e-7)
    np.testing.assert_allclose(A, A.T, atol=1e-12)
    # Jacobi blocks: sum over nodes of M^2 times the per-DOF patch Hessian of
    # the golden-pinned oracle
    X = X_of(C0)
    pats = oracle.materialize_patches(ctx.wire, g.blocks, 2)
    dofs = np.asarray(ctx.wire["groups"][0]["dof_idx"])
    hn = np.zeros((len(X), 2, 2))
    for i, d in enumerate(pats.dof):
        sl = slice(pats.poff[i], pats.poff[i + 1])
        _, h, _, _ = oracle.patch_eval(2, "barrier", 0.0, pats.gc[sl], pats.gn[:, sl], pats.s[:, sl],
                                       pats.role[sl], np.eye(2).reshape(-1), X)
        hn[d] = h
    assert len(pats.dof) == len(dofs)
    np.testing.assert_allclose(P, np.einsum("nf,nij->fij", M[:, free] ** 2, hn), rtol=1e-12, atol=1e-12)


def _resampled_mindet(cm, C, factor):
    fine = control_map(tuple((n - 1) * factor + 1 for n in cm.node_shape), cm.ctrl_shape)
    X = fine.prolong(C)
    du = np.diff(X, axis=0)[:, :-1]
    dv = np.diff(X, axis=1)[:-1, :]
    return float((du[..., 0] * dv[..., 1] - du[..., 1] * dv[..., 0]).min())


def test_oracle_loop_meets_the_acceptance_criteria(oracle):
    g, ctx, cm, wire, OC = _wavy()
    es = ctx.energy_stencil
    X, C, rep = OC.run_control(2, "barrier", g.X, es, wire, g.blocks[0], n_outer=30)
    e = np.asarray(rep["energies"])
    assert np.all(np.isfinite(e)) and np.all(np.asarray(rep["mindets"]) > 0)
    assert np.all(np.diff(e) <= 1e-6) and rep["iters"] <= 30
    assert int(np.argmax(e <= e[-1] + 0.01 * abs(e[-1]))) <= 30
    # no worse than the converged node smoother (Gauss-Seidel to a fixed point)
    Xn, _ = oracle.gauss_seidel(2, "barrier", g.X, oracle.materialize_patches(ctx.wire, g.blocks, 2))
    e_node = oracle.energy_mindet(2, "barrier", 0.0, es, Xn)[0]
    assert e[-1] <= 1.05 * e_node
    # the net regrids validly
    Cn = C.reshape(cm.ctrl_shape + (2,))
    assert _resampled_mindet(cm, Cn, 2) > 0 and _resampled_mindet(cm, Cn, 4) > 0
    # the evaluated net is the returned state
    M = OC.block_map(wire, 2)
    np.testing.assert_allclose(M @ C + np.asarray(wire["b"]).reshape(-1, 2), X, atol=1e-12)
