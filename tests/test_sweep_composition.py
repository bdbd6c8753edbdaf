"""One sweep of a single free DOF, pinned by the mathematics rather than by
a restatement of the update: the Newton step comes from central
differences of the global FP64 energy (gradient and Hessian of E - tau.x),
the acceptance of each trial from the global energy and min det, and the
sweep must return pos + omega 2^-i delta for the first accepted trial i of
the ladder 1, 1/2, ..., 1/512, or pos when all ten are rejected (reference
src/sweep.hpp:1534-1650: kMaxBacktracks = 10, accept iff finite, e <= e0 +
tol and min det > 0, omega post-blend). The linear shift tau (the FAS
coarse objective) pushes the shifted minimiser past a cell inversion, so
the ladder runs to every depth. CPU: the oracle; GPU: the HIP sweep on the
same single-DOF grids (without tau, which only FAS coarse levels carry)."""

from __future__ import annotations

import numpy as np
import pytest

from egg_amd import grids
from tests.cases import Case


def _one_dof(dim, seed, disp, stored=False):
    g = grids.single_block_2d(7, seed=seed, amp=0.2) if dim == 2 else grids.single_block_3d(5, seed=seed, amp=0.15)
    ids = np.asarray(g.blocks[0])
    node = int(ids[(3, 3)] if dim == 2 else ids[(2, 2, 2)])
    free = np.zeros_like(g.free)
    free[node] = True
    g.free = free
    h = 1.0 / (ids.shape[0] - 1)
    v = np.random.default_rng(seed + 100).normal(size=dim)
    g.X = g.X.copy()
    g.X[node] += disp * h * v / np.linalg.norm(v)
    ns = (1 << dim) * (ids.shape[0] - 1) ** dim
    return Case(g, weight=np.ones(ns) if stored else None), node, h


def _predict(O, c, node, dim, h, tau_vec, omega):
    """FD Newton step of E - tau.x at the node and the ladder decision;
    None when a decision is within rounding of the acceptance threshold."""
    es = c.ctx.energy_stencil
    x0 = c.X[node].copy()

    def E(x):
        X = c.X.copy()
        X[node] = x
        e, m = O.energy_mindet(dim, "barrier", 0.0, es, X)
        return e - tau_vec @ x, m

    eps = 1e-3 * h
    g = np.zeros(dim)
    H = np.zeros((dim, dim))
    for i in range(dim):
        ei = np.eye(dim)[i] * eps
        g[i] = (E(x0 + ei)[0] - E(x0 - ei)[0]) / (2 * eps)
        for j in range(dim):
            ej = np.eye(dim)[j] * eps
            H[i, j] = (E(x0 + ei + ej)[0] - E(x0 + ei - ej)[0] - E(x0 - ei + ej)[0] + E(x0 - ei - ej)[0]) / (4 * eps**2)
    delta = -np.linalg.solve(H, g)
    e0 = E(x0)[0]
    margin = 1e-8 * max(1.0, abs(e0))
    for i in range(10):
        xt = x0 + 0.5**i * delta
        e, m = E(xt)
        if abs(e - e0) < margin or abs(m) < 1e-9 * h**dim:
            return None  # too close to call
        if e <= e0 and m > 0:
            return i, x0 + omega * 0.5**i * delta
    return 10, x0


def _cases():
    out = []
    for dim in (2, 3):
        for seed in range(6):
            out.append((dim, seed, 0.3, 0.0, False))
            for mag in (1, 3, 10, 30, 100):
                out.append((dim, seed, 0.3, mag, False))
        out.append((dim, 0, 1.3, 0.0, False))  # starts inverted: every trial rejected
    out += [(2, s, 0.3, mag, True) for s in range(3) for mag in (0, 3, 30)]  # stored-stencil DOF
    return out


def _tau(dim, seed, mag, c, node, h, O):
    """A linear shift of `mag` times the node's energy curvature times h, in a
    seeded direction."""
    if mag == 0.0:
        return np.zeros(dim)
    es = c.ctx.energy_stencil
    eps = 1e-4 * h
    curv = 0.0
    for i in range(dim):
        X = [c.X.copy() for _ in range(3)]
        X[0][node, i] -= eps
        X[2][node, i] += eps
        e = [O.energy_mindet(dim, "barrier", 0.0, es, x)[0] for x in X]
        curv += abs(e[0] - 2 * e[1] + e[2]) / eps**2 / dim
    v = np.random.default_rng(seed).normal(size=dim)
    return mag * curv * h * v / np.linalg.norm(v)


def test_oracle_sweep_matches_the_ladder_from_finite_differences(oracle):
    depths = set()
    for dim, seed, disp, mag, stored in _cases():
        c, node, h = _one_dof(dim, seed, disp, stored)
        tau_vec = _tau(dim, seed, mag, c, node, h, oracle)
        for omega in (1.0, 0.8):
            pred = _predict(oracle, c, node, dim, h, tau_vec, omega)
            if pred is None:
                continue
            depth, want = pred
            tau = np.zeros_like(c.X)
            tau[node] = tau_vec
            Xo, _, _ = oracle.jacobi(dim, "barrier", 0.0, omega, 1, c.X, c.patches(oracle), c.ctx.energy_stencil,
                                     tau=tau if mag else None)
            # finite differences give the step to ~1e-5 of its length; a wrong
            # halving factor, trial count or blend is off by a fraction of it
            step = np.linalg.norm(want - c.X[node])
            np.testing.assert_allclose(Xo[node], want, rtol=0, atol=1e-5 * (h + step),
                                       err_msg=f"dim {dim} seed {seed} disp {disp} tau {mag} omega {omega}")
            others = np.arange(len(c.X)) != node
            assert np.array_equal(Xo[others], c.X[others])
            depths.add(depth)
    # the cases reach the full step, the halvings and the all-rejected end
    assert 0 in depths and 10 in depths and len(depths & set(range(1, 10))) >= 5, depths


@pytest.mark.gpu
def test_device_sweep_matches_the_ladder(oracle):
    """The HIP sweep of the same one-DOF grids (no tau) against the
    FD-predicted update, synthesized and stored DOFs, omega 1 and 0.8; an
    inverted start leaves the node where it was, bit for bit."""
    from egg_amd.smoothing import CppStructuredSweepSession

    n = 0
    for dim, seed, disp, mag, stored in _cases():
        if mag:
            continue
        c, node, h = _one_dof(dim, seed, disp, stored)
        for omega in (1.0, 0.8):
            pred = _predict(oracle, c, node, dim, h, np.zeros(dim), omega)
            if pred is None:
                continue
            depth, want = pred
            print(f"one-DOF sweep: dim {dim} seed {seed} disp {disp} stored {stored} omega {omega} depth {depth}",
                  flush=True)
            s = CppStructuredSweepSession(c.ctx, c.bsc, c.X, device="gpu")
            s.run(1, omega=omega)
            got = s.get_X()[node]
            if depth == 10:
                assert np.array_equal(got.astype(np.float32), c.X[node].astype(np.float32))
            else:
                np.testing.assert_allclose(got, want, rtol=0, atol=2e-5 * h)
            n += 1
    assert n >= 20
