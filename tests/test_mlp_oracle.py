"""CPU checks of the learned-proposal oracle (oracle/mlp.py): the backward
pass against finite differences, the BF16 emulation against exact float64,
and the features' invariances."""

from __future__ import annotations

import numpy as np
import pytest

from oracle import mlp as M


def _params(dim, rng, scale=0.3):
    return (rng.standard_normal(M.n_params(dim)) * scale).astype(np.float64)


@pytest.mark.parametrize("dim", [2, 3])
def test_f64_gradient_matches_finite_differences(dim):
    rng = np.random.default_rng(1)
    p = _params(dim, rng)
    d = M.DIMS[dim]
    X = rng.standard_normal((8, d["n_in_padded"]))
    T = rng.standard_normal((8, d["n_out"]))
    _, g = M.loss_grad_f64(p, X, T, dim)
    idx = rng.choice(p.size, 40, replace=False)
    for i in idx:
        h = 1e-6
        pp = p.copy()
        pp[i] += h
        pm = p.copy()
        pm[i] -= h
        fd = (M.loss_grad_f64(pp, X, T, dim)[0] - M.loss_grad_f64(pm, X, T, dim)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


def test_bf16_rounding_rule():
    x = np.array([1.0, 1.0 + 2**-8, 1.0 + 3 * 2**-8, -2.5, 1e-3, 65504.0], np.float32)
    b = M.bf16(x)
    # ties to even: 1 + 2^-8 is a tie between 1 and 1 + 2^-7 -> 1; 1 + 3*2^-8 -> 1 + 2^-6
    assert b[0] == 1.0 and b[1] == 1.0 and b[2] == np.float32(1 + 2**-6) and b[3] == -2.5
    assert np.all(np.abs(b - x) <= np.abs(x) * 2**-8)


@pytest.mark.parametrize("dim", [2, 3])
def test_bf16_network_close_to_exact(dim):
    rng = np.random.default_rng(2)
    p = _params(dim, rng, 0.2)
    d = M.DIMS[dim]
    X = rng.standard_normal((256, d["n_in_padded"])).astype(np.float32)
    T = rng.standard_normal((256, d["n_out"])).astype(np.float32)
    l16, g16 = M.loss_grad(p, X, T, dim)
    l64, g64 = M.loss_grad_f64(p, X, T, dim)
    assert abs(l16 - l64) <= 0.05 * l64
    assert np.linalg.norm(g16 - g64) <= 0.05 * np.linalg.norm(g64)


@pytest.mark.parametrize("dim", [2, 3])
def test_features_are_frame_invariant(dim):
    rng = np.random.default_rng(3)
    offs = np.array(np.meshgrid(*[[-1, 0, 1]] * dim, indexing="ij")).reshape(dim, -1).T
    ring = offs * 0.1 + rng.uniform(-0.02, 0.02, offs.shape)
    dprev = rng.uniform(-0.01, 0.01, dim)
    E, h, f = M.frame_features(ring, dprev, dim)
    assert np.allclose(E @ E.T, np.eye(dim), atol=1e-12) and np.linalg.det(E) > 0
    # rigid motion + uniform scaling of the ring leaves the features unchanged
    q, _ = np.linalg.qr(rng.standard_normal((dim, dim)))
    if np.linalg.det(q) < 0:
        q[:, 0] = -q[:, 0]
    s = 3.7
    E2, h2, f2 = M.frame_features(s * ring @ q.T + 5.0, s * dprev @ q.T, dim)
    np.testing.assert_allclose(f2, f, atol=1e-12)
    assert h2 == pytest.approx(s * h)
    # the centre contributes zeros; the constants follow the ring
    C = (3**dim - 1) // 2
    assert np.all(f[C * dim:(C + 1) * dim] == 0.0)
    assert f[3**dim * dim] == 1.0
