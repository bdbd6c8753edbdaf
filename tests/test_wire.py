"""The flat wire builders (egg_amd.smoothing.flat_context, directional,
wire_tables) pinned by what the wire means (reference
egg/smoothing/flat_context.py:47-346 conventions): which DOFs are
synthesized, which samples a stored DOF ships and in which roles, the
orientation of every sample, and the per-DOF window tables."""

from __future__ import annotations

import numpy as np
import pytest

from egg_amd import grids
from egg_amd.smoothing import build_flat_context
from tests.cases import star


def _cells_of(blocks, d):
    """Per node: the (block, cell corner tuple) cells it is a corner of."""
    out = {}
    for b, a in enumerate(blocks):
        a = np.asarray(a)
        for c in np.ndindex(*[s - 1 for s in a.shape]):
            for off in np.ndindex(*([2] * d)):
                out.setdefault(int(a[tuple(np.add(c, off))]), []).append((b, c))
    return out


@pytest.mark.parametrize("name", ["single2d", "ogrid", "star5", "cube", "cubed_sphere"])
def test_synthesized_and_stored_dofs(name):
    g = {"single2d": grids.single_block_2d(9, seed=1), "ogrid": grids.ogrid_2d(17, 9), "star5": star(k=5).grid,
         "cube": grids.single_block_3d(6, seed=1), "cubed_sphere": grids.cubed_sphere_3d(7)}[name]
    d = g.d
    w = build_flat_context(g.blocks, g.free, d)
    grp, es = w["groups"][0], w["energy_stencil"]
    dofs = np.asarray(grp["dof_idx"])
    assert np.array_equal(dofs, np.flatnonzero(g.free))
    # synthesized exactly when strictly inside a block: logical in [1, n-2] on every axis
    inside = {}
    for b, a in enumerate(g.blocks):
        a = np.asarray(a)
        if min(a.shape) >= 3:
            for lg in np.ndindex(*[s - 2 for s in a.shape]):
                inside[int(a[tuple(np.add(lg, 1))])] = (b, tuple(np.add(lg, 1)))
    ib = np.asarray(grp["interior_block"])
    il = np.asarray(grp["interior_logical"]).reshape(-1, d)
    for i, n in enumerate(dofs.tolist()):
        if n in inside:
            assert (ib[i], tuple(il[i])) == inside[n]
        else:
            assert ib[i] == -1
    # a stored DOF ships 2^d samples per incident cell, each naming it in its role
    cells = _cells_of(g.blocks, d)
    P = np.asarray(grp["P_of"])
    off = np.concatenate([[0], np.cumsum(P)])
    gc = np.asarray(grp["gc"])
    gn = [np.asarray(grp[f"gn{k}"]) for k in range(d)]
    role = np.asarray(grp["role"])
    for i, n in enumerate(dofs.tolist()):
        if ib[i] >= 0:
            assert P[i] == 0
            continue
        assert P[i] == (1 << d) * len(cells[n])
        for q in range(off[i], off[i + 1]):
            r = role[q]
            if r == 0:
                assert gc[q] == n
            elif r > 0:
                assert gn[r - 1][q] == n and gc[q] != n
            else:
                assert gc[q] != n and all(gn[k][q] != n for k in range(d))
    # every energy sample of an undeformed grid has a positive det
    X = g.X
    A = np.stack([np.asarray(es[f"s{k}"])[:, None] * (X[np.asarray(es[f"gn{k}"])] - X[np.asarray(es["gc"])])
                  for k in range(d)], axis=2)
    assert es["num_samples"] == sum((1 << d) * int(np.prod([s - 1 for s in np.shape(a)])) for a in g.blocks)
    assert np.linalg.det(A).min() > 0


def test_unit_block_sample_geometry():
    """On the unit lattice every sample's columns are the unit axes: corner
    b's axis-k neighbour flips bit k (axis 0 the most significant), with sign
    -1 exactly where that bit is set, so A = I at every sample."""
    n = 4
    ids = np.arange(n**3).reshape(n, n, n)
    X = np.indices((n, n, n)).reshape(3, -1).T.astype(float)
    es = build_flat_context([ids], np.zeros(n**3, bool), 3)["energy_stencil"]
    A = np.stack([np.asarray(es[f"s{k}"])[:, None] * (X[np.asarray(es[f"gn{k}"])] - X[np.asarray(es["gc"])])
                  for k in range(3)], axis=2)
    np.testing.assert_array_equal(A, np.broadcast_to(np.eye(3), A.shape))
    gc = np.asarray(es["gc"]).reshape(-1, 8)
    np.testing.assert_array_equal(gc[0], [0, 1, 4, 5, 16, 17, 20, 21])  # first cell's corners, bit order


def test_targets_weights_and_entities_disable_synthesis_where_they_must():
    g = grids.single_block_2d(7, seed=1)
    ns = 4 * 36
    base = build_flat_context(g.blocks, g.free, 2)["groups"][0]
    assert np.all(np.asarray(base["interior_block"]) == 0)
    for kw in (dict(weight=np.ones(ns)), dict(w_inv=np.diag([1.0, 2.0])),
               dict(w_inv=np.concatenate([np.eye(2)[None]] * (ns - 1) + [2 * np.eye(2)[None]]))):
        grp = build_flat_context(g.blocks, g.free, 2, **kw)["groups"][0]
        assert np.all(np.asarray(grp["interior_block"]) == -1), kw
    grp = build_flat_context(g.blocks, g.free, 2, w_inv=3.0 * np.eye(2))["groups"][0]
    assert np.all(np.asarray(grp["interior_block"]) == 0)  # an isotropic target still synthesizes
    ann, ents = grids.annulus_2d(17, 9)
    grp = build_flat_context(ann.blocks, ann.free, 2, entities=ents)["groups"][0]
    dofs = np.asarray(grp["dof_idx"])
    for kind, part in grp["entities"].items():
        local = np.asarray(part["dof_local"])
        assert np.all(np.diff(local) > 0)
        for j, i in enumerate(local.tolist()):
            n = int(dofs[i])
            want = ents.get(n, ("free", ()))
            assert want[0] == kind
            np.testing.assert_array_equal(np.asarray(part["records"]).reshape(len(local), -1)[j]
                                          if part["kFields"] else [], want[1])
            assert grp["interior_block"][i] == -1 or kind == "free"
    with pytest.raises(ValueError, match="not supported"):
        build_flat_context(ann.blocks, ann.free, 2, entities={int(dofs[0]): ("torus", (1.0,))})
    with pytest.raises(ValueError, match="holds 3 values"):
        build_flat_context(ann.blocks, ann.free, 2, entities={int(dofs[0]): ("circle", (1.0,))})


@pytest.mark.parametrize("dim", [2, 3])
def test_directional_and_curvature_tables(dim):
    """Every (sample, slot) whose node moves is listed once, under that DOF,
    in sample then slot order; padding is slot -1; reserved kinds list no
    node. Curvature windows likewise (2D)."""
    from egg_amd.smoothing.directional import DirectionalSamples, kind_arity

    g = grids.single_block_2d(9, seed=4) if dim == 2 else grids.single_block_3d(6, seed=2)
    s = grids.wall_directional_samples(g)
    kind = s.kind.copy()
    kind[1] = 4  # a reserved kind: no slot listed
    s = DirectionalSamples(s.nodes, kind, s.ref, s.axis, s.weight, s.eps)
    grp = build_flat_context(g.blocks, g.free, dim, directional=s)["groups"][0]
    K = grp["dir_k"]
    dofs = np.asarray(grp["dof_idx"])
    slot = np.asarray(grp["dir_slot"]).reshape(-1, K)
    nodes = np.asarray(grp["dir_nodes"]).reshape(-1, K, 4)
    listed = []
    for i, n in enumerate(dofs.tolist()):
        for j in range(K):
            if slot[i, j] < 0:
                assert np.all(nodes[i, j] == -1) and grp["dir_weight"][i * K + j] == 0
                continue
            assert nodes[i, j, slot[i, j]] == n
            sample = int(np.flatnonzero((s.nodes == nodes[i, j]).all(axis=1))[0])
            listed.append((sample, int(slot[i, j])))
    expect = [(si, sl) for si in range(s.n) for sl in range(kind_arity(int(s.kind[si])))
              if s.nodes[si, sl] >= 0 and g.free[s.nodes[si, sl]]]
    assert sorted(listed) == sorted(expect) and len(listed) == len(set(listed))
    if dim == 2:
        o = grids.ogrid_2d(17, 9)
        cn, cw = grids.interface_c2_windows(o)
        grp = build_flat_context(o.blocks, o.free, 2, curvature=(cn, cw))["groups"][0]
        K = grp["curv_k"]
        place = np.asarray(grp["curv_slot"]).reshape(-1, K)
        win = np.asarray(grp["curv_nodes"]).reshape(-1, K, 4)
        got = 0
        for i, n in enumerate(np.asarray(grp["dof_idx"]).tolist()):
            for j in range(K):
                if place[i, j] >= 0:
                    assert win[i, j, place[i, j]] == n
                    got += 1
        assert got == int(np.sum(o.free[np.asarray(cn)]))
