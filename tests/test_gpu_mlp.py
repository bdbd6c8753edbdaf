"""The learned proposal on the MI355X: BF16 matrix-core forward and backward
passes against the NumPy oracle (oracle/mlp.py), training, harvested pairs
against features and displacements recomputed from the session's states,
and the accept-or-fallback sweep."""

from __future__ import annotations

import numpy as np
import pytest

from oracle import mlp as M
from tests.cases import Case, ogrid, single

pytestmark = pytest.mark.gpu


def _net(dim, capacity=0, seed=0):
    from egg_amd._cpp import cpp_core

    return cpp_core.ProposalMLP(dim, seed, capacity)


def _session(c):
    from egg_amd.smoothing import CppStructuredSweepSession

    return CppStructuredSweepSession(c.ctx, c.bsc, c.X)


def _data(dim, n, rng):
    d = M.DIMS[dim]
    X = np.zeros((n, d["n_in_padded"]), np.float32)
    X[:, :d["n_in"]] = rng.standard_normal((n, d["n_in"]))
    T = (0.5 * np.tanh(X[:, :d["n_out"]] - 0.3 * X[:, d["n_out"]:2 * d["n_out"]])).astype(np.float32)
    return X, T


@pytest.mark.parametrize("dim", [2, 3])
def test_forward_matches_oracle(dim):
    rng = np.random.default_rng(10 + dim)
    net = _net(dim)
    p = (rng.standard_normal(M.n_params(dim)) * 0.2).astype(np.float32)
    net.set_params(p)
    np.testing.assert_array_equal(net.get_params(), p)
    X, _ = _data(dim, 1000, rng)
    Y = net.forward(X)
    Yo = M.forward(p, X, dim)
    # same BF16 roundings; FP32 summation order and exp differ, which can flip
    # the rounding of an activation (2^-8 relative) now and then
    assert np.linalg.norm(Y - Yo) <= 2e-3 * np.linalg.norm(Yo)
    assert np.abs(Y - Yo).max() <= 2e-2 * np.abs(Yo).max()


@pytest.mark.parametrize("dim", [2, 3])
def test_gradient_matches_oracle(dim):
    rng = np.random.default_rng(20 + dim)
    net = _net(dim, capacity=1024)
    p = (rng.standard_normal(M.n_params(dim)) * 0.2).astype(np.float32)
    net.set_params(p)
    X, T = _data(dim, 1024, rng)
    net.data_add(X, T)
    Xs, Ts = net.data_get(0, 1024)
    np.testing.assert_array_equal(Xs, M.bf16(X))
    np.testing.assert_array_equal(Ts, T)
    rows = rng.permutation(1024)[:512].astype(np.int32)
    g, loss = net.gradient(rows)
    lo, go = M.loss_grad(p, X[rows], T[rows], dim)
    assert loss == pytest.approx(lo, rel=1e-3)
    assert np.linalg.norm(g - go) <= 1e-2 * np.linalg.norm(go)
    # and the exact float64 gradient is close (BF16 operands)
    _, g64 = M.loss_grad_f64(p, X[rows], T[rows], dim)
    assert np.linalg.norm(g - g64) <= 5e-2 * np.linalg.norm(g64)


@pytest.mark.parametrize("dim,n", [(2, 32768), (3, 16384)])
def test_large_batch_gradient_matches_oracle(dim, n):
    """Batches whose weight-gradient waves sum several 256-sample chunks."""
    rng = np.random.default_rng(40 + dim)
    net = _net(dim, capacity=n)
    p = (rng.standard_normal(M.n_params(dim)) * 0.2).astype(np.float32)
    net.set_params(p)
    X, T = _data(dim, n, rng)
    net.data_add(X, T)
    rows = rng.permutation(n).astype(np.int32)
    g, loss = net.gradient(rows)
    lo, go = M.loss_grad(p, X[rows], T[rows], dim)
    assert loss == pytest.approx(lo, rel=1e-3)
    assert np.linalg.norm(g - go) <= 1e-2 * np.linalg.norm(go)


@pytest.mark.parametrize("dim", [2, 3])
def test_training_refreshes_fragments(dim):
    """The Adam step rewrites the BF16 fragments of every parameter it moves:
    the trained network's forward equals the forward after repacking its
    FP32 parameters from scratch."""
    rng = np.random.default_rng(50 + dim)
    net = _net(dim, capacity=4096, seed=5)
    X, T = _data(dim, 4096, rng)
    net.data_add(X, T)
    net.train(2, batch=1024, lr=1e-2, seed=2)
    Y1 = net.forward(X[:2048])
    net.set_params(net.get_params())
    Y2 = net.forward(X[:2048])
    np.testing.assert_array_equal(Y1, Y2)


def test_training_reduces_loss():
    rng = np.random.default_rng(30)
    net = _net(2, capacity=16384, seed=3)
    X, T = _data(2, 16384, rng)
    net.data_add(X, T)
    _, l0 = net.gradient(np.arange(4096, dtype=np.int32))
    hist = net.train(30, batch=1024, lr=3e-3, seed=1)
    assert np.all(np.isfinite(hist))
    assert hist[-1] < 0.1 * l0, (l0, hist)
    # the trained network's own forward agrees with its training loss
    Y = net.forward(X[:4096])
    assert np.mean(np.sum((Y - T[:4096]) ** 2, 1)) < 0.15 * l0


def test_harvested_pairs_match_states():
    """One snapshot after 1 sweep, displacement over the next 2 sweeps: the
    stored features and targets equal the oracle's features of the states
    the same sweeps produce."""
    c = single(n=17, seed=4)
    s = _session(c)
    net = _net(2, capacity=4096)
    added = s.harvest(net, 1, every=1, horizon=2, omega=0.8)
    n = 15 * 15
    assert added == n
    r = _session(c)
    X0 = c.X
    r.run(1, omega=0.8)
    X1 = r.get_X().reshape(-1, 2)
    r.run(2, omega=0.8)
    X3 = r.get_X().reshape(-1, 2)
    Xf, Tf = net.data_get(0, n)
    ids = c.grid.blocks[0]
    q = 0
    for i in range(1, 16):
        for j in range(1, 16):
            ring = X1[ids[i - 1:i + 2, j - 1:j + 2].reshape(-1)]
            E, h, f = M.frame_features(ring, X1[ids[i, j]] - X0[ids[i, j]], 2)
            np.testing.assert_allclose(Xf[q], M.bf16(f.astype(np.float32)), rtol=2**-7, atol=1e-6)
            np.testing.assert_allclose(Tf[q], E @ (X3[ids[i, j]] - X1[ids[i, j]]) / h, rtol=1e-3, atol=1e-5)
            q += 1


def test_zero_network_falls_back_to_plain():
    """A network that proposes no motion is never accepted (the patch energy
    must drop by more than the FP32 slack), so the sweep is the plain method."""
    c = ogrid(33, 17)
    net = _net(2)
    net.set_params(np.zeros(M.n_params(2), np.float32))
    a = _session(c)
    b = _session(c)
    b.set_proposal(net)
    b.core.set_counters(True)
    ea, ma = a.run(20, omega=0.8, report_every=5)
    eb, mb = b.run(20, omega=0.8, report_every=5)
    # the fallback is the plain kernel's code in another instantiation:
    # equal up to the compiler's choice of fused multiply-adds
    np.testing.assert_allclose(eb, ea, rtol=1e-6)
    np.testing.assert_allclose(b.get_X(), a.get_X(), rtol=0, atol=1e-5)
    cnt = b.core.last_counters(1)[0]
    assert cnt[2] == 0 and cnt[1] == c.grid.n_free * 20


def test_proposal_in_fas_fine_smoothing():
    """With a proposal set, FAS's fine-level smoothing sweeps run the
    proposal kernel: a zero network is never accepted (the cycle equals the
    plain FAS cycle up to fused multiply-adds), and its counted updates are
    the fine sweeps'."""
    c = ogrid(33, 17)
    net = _net(2)
    net.set_params(np.zeros(M.n_params(2), np.float32))
    a = _session(c)
    b = _session(c)
    b.set_proposal(net)
    b.core.set_counters(True)
    ea, ma = a.run_fas(3, nu_pre=2, nu_post=2, omega=0.8, max_levels=3)
    eb, mb = b.run_fas(3, nu_pre=2, nu_post=2, omega=0.8, max_levels=3)
    np.testing.assert_allclose(eb, ea, rtol=1e-6)
    np.testing.assert_allclose(b.get_X(), a.get_X(), rtol=0, atol=1e-5)
    cnt = b.core.last_counters(3)
    assert cnt[0][2] == 0 and cnt[0][1] == c.grid.n_free * 3 * 4
    # a random network's proposals are accepted now and then, on the fine
    # level only, and the cycles stay valid
    rng = np.random.default_rng(7)
    net.set_params((rng.standard_normal(M.n_params(2)) * 0.05).astype(np.float32))
    b.set_X(c.X)
    e, m = b.run_fas(3, nu_pre=2, nu_post=2, omega=0.8, max_levels=3)
    cnt = b.core.last_counters(3)
    assert cnt[0][2] > 0 and cnt[1][2] == 0 and cnt[2][2] == 0
    assert np.all(np.isfinite(e)) and np.all(m > 0)


def test_wild_network_keeps_mesh_valid():
    """Large random proposals are mostly rejected; the accepted ones never
    fold a cell and never raise a DOF's patch energy, so the energy stays
    finite and every reported min det positive."""
    c = ogrid(33, 17)
    rng = np.random.default_rng(5)
    net = _net(2)
    net.set_params((rng.standard_normal(M.n_params(2)) * 0.5).astype(np.float32))
    s = _session(c)
    s.set_proposal(net)
    s.core.set_counters(True)
    e, m = s.run(30, omega=0.8, report_every=1)
    assert np.all(np.isfinite(e)) and np.all(m > 0)
    trials, updates, accepted = (int(v) for v in s.core.last_counters(1)[0])
    assert updates == c.grid.n_free * 30 and 0 <= accepted < updates


def test_3d_proposal_paths():
    """3D (hex) proposal kernels: a zero network is never accepted; a
    random one keeps every cell valid and every patch energy from rising."""
    from tests.cases import cube

    c = cube(n=9, seed=2)
    zero = _net(3)
    zero.set_params(np.zeros(M.n_params(3), np.float32))
    a = _session(c)
    b = _session(c)
    b.set_proposal(zero)
    b.core.set_counters(True)
    ea, _ = a.run(6, omega=0.8, report_every=3)
    eb, _ = b.run(6, omega=0.8, report_every=3)
    np.testing.assert_allclose(eb, ea, rtol=1e-6)
    assert b.core.last_counters(1)[0][2] == 0
    rng = np.random.default_rng(7)
    wild = _net(3)
    wild.set_params((rng.standard_normal(M.n_params(3)) * 0.3).astype(np.float32))
    s = _session(c)
    s.set_proposal(wild, mode="predict_correct", scale=0.5)
    e, m = s.run(10, omega=0.8, report_every=1)
    assert np.all(np.isfinite(e)) and np.all(m > 0)


def test_trained_proposal_gains():
    """What the trained network gains over the method without it: trained
    (on the GPU, by the kernels above) on pairs harvested from plain runs on
    held-out tangle seeds of the O-grid generator, the predict-correct
    proposal reaches a lower energy than the plain sweep at the same number
    of sweeps on an unseen seed, with every proposal it takes accepted only
    on an energy drop (min det stays > 0)."""
    from egg_amd import grids
    from egg_amd.smoothing import cpp_untangle

    def make(seed):
        c = Case(grids.tangle(grids.ogrid_2d(65, 33), frac=0.02, factor=1.5, seed=seed))
        s = _session(c)
        cpp_untangle(c.ctx, c.bsc, c.X, session=s)
        return c, s

    net = _net(2, capacity=2_000_000)
    for seed in (21, 22, 23):
        _, sh = make(seed)
        sh.harvest(net, 40, every=25, horizon=2, omega=0.8)
    hist = net.train(10, batch=4096, lr=1e-3, seed=1)
    assert hist[-1] < 0.2 * hist[0]
    c, plain = make(99)
    X0 = plain.get_X()
    ep, mp = plain.run(1000, omega=0.8, report_every=250)
    prop = _session(c)
    prop.set_X(X0)
    prop.set_proposal(net, mode="predict_correct", scale=0.5)
    prop.core.set_counters(True)
    eq, mq = prop.run(1000, omega=0.8, report_every=250)
    accepted = prop.core.last_counters(1)[0][2]
    print("plain", ep, "proposal", eq, "accepted", accepted)
    assert np.all(mq > 0) and accepted > 0
    assert np.all(eq < ep)


@pytest.mark.parametrize("variant", [0, 8])
@pytest.mark.parametrize("graphs", [0, 6])
def test_proposal_cadence_composes(variant, graphs):
    """With every = 3 a run proposes in sweeps 0, 3, 6, ... and runs the
    plain kernel of the session's tile variant in between: the same bits as
    one-sweep runs that set and unset the proposal by hand. Variant 8 has no
    proposal kernel of its own; its proposal sweeps run variant 3's kernel,
    which has the same tile table."""
    c = ogrid(33, 17)
    rng = np.random.default_rng(3)
    net = _net(2)
    net.set_params((rng.standard_normal(M.n_params(2)) * 0.05).astype(np.float32))
    a, b = _session(c), _session(c)
    for s in (a, b):
        s.core.set_tile_variant(variant)
        s.core.set_graph_chunk(graphs)
    a.set_proposal(net, mode="replace", every=3)
    a.core.set_counters(True)
    ea, ma = a.run(12, omega=0.8, report_every=12)
    accepted = a.core.last_counters(1)[0][2]
    for i in range(12):
        b.set_proposal(net if i % 3 == 0 else None, mode="replace")
        eb, mb = b.run(1, omega=0.8)
    print(f"cadence: variant {variant} graphs {graphs} accepted {accepted}", flush=True)
    assert accepted > 0
    np.testing.assert_array_equal(a.get_X(), b.get_X())
    np.testing.assert_array_equal(ea[-1:], eb)


def test_proposal_cadence_in_fas():
    """With every > 1 a FAS fine smoothing call proposes in its first sweep
    only: one-sweep smoothing then matches every = 1 bit for bit, longer
    smoothing does not."""
    c = ogrid(33, 17)
    rng = np.random.default_rng(4)
    net = _net(2)
    net.set_params((rng.standard_normal(M.n_params(2)) * 0.05).astype(np.float32))
    runs = {}
    for every in (1, 2):
        for nu in (1, 2):
            s = _session(c)
            s.set_proposal(net, mode="replace", every=every)
            e, m = s.run_fas(3, nu_pre=nu, nu_post=nu, omega=0.8, max_levels=3)
            assert np.all(m > 0)
            runs[every, nu] = (e, s.get_X())
    np.testing.assert_array_equal(runs[1, 1][1], runs[2, 1][1])
    assert not np.array_equal(runs[1, 2][1], runs[2, 2][1])


def test_trained_proposal_gains_3d():
    """What the trained 3D network gains over the method without it: trained
    on the GPU on pairs harvested from plain runs on held-out tangle seeds of
    the cubed-sphere generator (7 blocks of 33^3), the replace-mode proposal
    in every 4th sweep reaches a lower energy than the plain sweep at the
    same number of sweeps on an unseen seed, with every accepted proposal
    an energy drop (min det stays > 0)."""
    from egg_amd import grids
    from egg_amd.smoothing import cpp_untangle

    def make(seed):
        c = Case(grids.tangle(grids.cubed_sphere_3d(33), frac=0.01, factor=1.5, seed=seed))
        s = _session(c)
        cpp_untangle(c.ctx, c.bsc, c.X, session=s)
        return c, s

    net = _net(3, capacity=8_500_000)
    for seed in (11, 12):
        _, sh = make(seed)
        sh.harvest(net, 20, every=25, horizon=2, omega=0.8)
        del sh
    hist = net.train(10, batch=4096, lr=1e-3, seed=1)
    assert hist[-1] < 0.2 * hist[0]
    c, plain = make(99)
    X0 = plain.get_X()
    ep, mp = plain.run(400, omega=0.8, report_every=100)
    prop = _session(c)
    prop.set_X(X0)
    prop.set_proposal(net, mode="replace", every=4)
    prop.core.set_counters(True)
    eq, mq = prop.run(400, omega=0.8, report_every=100)
    accepted = prop.core.last_counters(1)[0][2]
    print("3D plain", ep, "proposal", eq, "accepted", accepted)
    assert np.all(mq > 0) and accepted > 0
    assert np.all(eq < ep)
