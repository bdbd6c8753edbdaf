"""The halo-table builders of egg_amd.smoothing.structured_tables, pinned by
what the tables mean rather than by how they are built: every ghost entry
names a grid-line neighbour of its face node lying across the face, the
interface builder covers every face node of every conforming interface, the
singular nodes are the ones whose valence the grid fixes, and a session
accepts the singular hosts (tables only on the CPU; parity on the GPU)."""

from __future__ import annotations

import numpy as np
import pytest

from egg_amd import grids
from egg_amd.smoothing import structured_arrays
from egg_amd.smoothing import structured_tables as T
from tests.cases import Case, star


def _grids():
    return {
        "ogrid": grids.ogrid_2d(33, 17),
        "star5": star(k=5).grid,
        "star6": star(k=6).grid,
        "cubed_sphere": grids.cubed_sphere_3d(9),
        "star5_3d": star(k=5, layers=5).grid,
        "strip": grids.strip_2d([9, 17, 5], 9),
    }


def _edges(g):
    out = set()
    for a in g.blocks:
        a = np.asarray(a)
        for axis in range(g.d):
            u = np.take(a, np.arange(a.shape[axis] - 1), axis=axis).reshape(-1)
            v = np.take(a, np.arange(1, a.shape[axis]), axis=axis).reshape(-1)
            out.update(zip(np.minimum(u, v).tolist(), np.maximum(u, v).tolist()))
    return out


def _check_entries(g, bsc):
    """Each entry: the source is a grid-line neighbour of the face node the
    ghost sits behind, in a block other than the ghost's, and never on the
    block's side of the face: across it, checked at free face nodes. (At a
    fixed node where a seam meets the physical boundary "across" has no
    meaning: the id-array builder fills the ghost behind the boundary face
    there with the other block's boundary neighbour, reference
    cpp_backend.py:163-186, and no patch reads it.) Returns the number of
    entries checked."""
    edges = _edges(g)
    d = g.d
    blocks = [np.asarray(b) for b in g.blocks]
    for sb, sp, db, dp in zip(bsc.halo_src_block, bsc.halo_src_padded, bsc.halo_dst_block, bsc.halo_dst_padded):
        shape = blocks[db].shape
        src = int(blocks[sb][tuple(np.asarray(sp) - 1)])
        axes = [k for k in range(d) if dp[k] in (0, shape[k] + 1)]
        assert len(axes) == 1, "a face ghost sits behind exactly one face"
        k = axes[0]
        face = np.asarray(dp) - 1
        face[k] = 0 if dp[k] == 0 else shape[k] - 1
        inner = face.copy()
        inner[k] += 1 if dp[k] == 0 else -1
        F = int(blocks[db][tuple(face)])
        if shape[k] > 1:
            I = int(blocks[db][tuple(inner)])
            if g.free[F]:
                assert np.dot(g.X[src] - g.X[F], g.X[F] - g.X[I]) > 0, "the source lies across the face"
        assert (min(F, src), max(F, src)) in edges, "the source is a grid-line neighbour of the face node"
        assert src not in set(blocks[db].reshape(-1).tolist()), "the source is not in the ghost's block"
    return len(bsc.halo_src_block)


@pytest.mark.parametrize("name", list(_grids()))
def test_twin_tables_name_neighbours_across_the_face(name):
    g = _grids()[name]
    bsc = T.build_structured_context_from_block_maps(g.d, g.blocks)
    n = _check_entries(g, bsc)
    assert n > 0
    # a source is always its node's owner copy: the first copy in block order
    first = {}
    for b, a in enumerate(g.blocks):
        for idx in zip(*np.unravel_index(np.arange(np.asarray(a).size), np.shape(a))):
            first.setdefault(int(np.asarray(a)[idx]), (b, idx))
    for sb, sp in zip(bsc.halo_src_block, bsc.halo_src_padded):
        gid = int(np.asarray(g.blocks[sb])[tuple(np.asarray(sp) - 1)])
        assert first[gid] == (int(sb), tuple(int(x) - 1 for x in sp))
    assert bsc.sing_block.size == 0


@pytest.mark.parametrize("name", list(_grids()))
def test_interface_tables_cover_every_face_node(name):
    g = _grids()[name]
    conns = T.face_connections(g.d, g.blocks)
    bsc = T.build_block_structured_context(g.d, g.blocks, g.free)
    n = _check_entries(g, bsc)
    expected = sum(2 * np.asarray(g.blocks[fa[0]]).size // np.shape(g.blocks[fa[0]])[fa[1]] for fa, _ in conns)
    assert n == expected
    dst = set(zip(bsc.halo_dst_block.tolist(), map(tuple, bsc.halo_dst_padded.tolist())))
    assert len(dst) == n, "no ghost is filled twice"


def test_face_connections_and_singular_nodes():
    """O-grid: four interfaces of the disc, four seams; its four disc
    corners are valence-3 singular nodes. A k-block star's centre has
    valence k (singular for k != 4). In the 7-block cubed sphere of side n
    the central cube's 8 corners have valence 4; the n - 2 inner nodes of its
    12 edges and of the 8 radial seams where three shell blocks meet have
    valence 5 (the seams' outer ends are on the fixed boundary)."""
    g = grids.ogrid_2d(33, 17)
    assert len(T.face_connections(2, g.blocks)) == 8
    s = T.singular_nodes(2, g.blocks, g.free)
    assert [(b, lg, v) for b, lg, v in s] == [(0, (0, 0), 3), (0, (0, 32), 3), (0, (32, 0), 3), (0, (32, 32), 3)]
    for k in (3, 5, 6):
        gs = star(k=k).grid
        sk = T.singular_nodes(2, gs.blocks, gs.free)
        assert len(sk) == 1 and sk[0][2] == k
    gs = star(k=4).grid
    assert T.singular_nodes(2, gs.blocks, gs.free) == []
    c = grids.cubed_sphere_3d(9)
    sc = T.singular_nodes(3, c.blocks, c.free)
    val = [v for _, _, v in sc]
    assert val.count(4) == 8 and val.count(5) == (12 + 8) * 7 and len(sc) == 8 + 20 * 7


def test_singular_hosts_follow_the_spare_ghost_room():
    """With room to spare every node keeps its first copy's block; a host
    short of room hands the node to the roomiest fan block, which then holds
    that much less."""
    g = grids.ogrid_2d(33, 17)
    bsc = T.build_block_structured_context(2, g.blocks, g.free)
    assert bsc.sing_block.tolist() == [0, 0, 0, 0]
    sings = T.singular_nodes(2, g.blocks, g.free)
    big = [(b, lg, 10**6) for b, lg, _ in sings]  # no block has that much room
    hosts = T.singular_hosts(2, g.blocks, big, bsc.halo_dst_block, bsc.halo_dst_padded)
    for (b0, lg0, _), (hb, hl) in zip(big, hosts):
        assert hb != 0 and int(np.asarray(g.blocks[hb])[hl]) == int(np.asarray(g.blocks[b0])[lg0])


def _forced_hosts(g, bsc):
    """The same tables with every singular node hosted by the highest-numbered
    block of its fan (a non-default owner)."""
    fan = {}
    for b, a in enumerate(g.blocks):
        a = np.asarray(a)
        for corner in np.ndindex(*([2] * g.d)):
            lg = tuple(0 if c == 0 else s - 1 for c, s in zip(corner, a.shape))
            fan.setdefault(int(a[lg]), []).append((b, lg))
    sb, sl = [], []
    for b, lg in zip(bsc.sing_block, bsc.sing_logical):
        gid = int(np.asarray(g.blocks[b])[tuple(lg)])
        hb, hl = max(fan[gid], key=lambda c: c[0])
        sb.append(hb)
        sl.append(hl)
    bsc.sing_block = np.asarray(sb, np.int32)
    bsc.sing_logical = np.asarray(sl, np.int32).reshape(-1, g.d)
    return bsc


def test_session_takes_singular_hosts():
    """A tables-only session accepts the interface tables with their
    singular hosts, default or forced to another fan block; the owner slot
    of a singular node is then in its host block."""
    from egg_amd._cpp import cpp_core

    g = grids.ogrid_2d(33, 17)
    c = Case(g)
    bsc = _forced_hosts(g, T.build_block_structured_context(2, g.blocks, g.free))
    s = cpp_core.CppStructuredSweepSession(c.ctx.wire, structured_arrays(bsc), g.X.ravel(), device="tables", dim=2)
    info, tab = s.info(), s.tables()
    base = np.asarray(tab["block_base"])
    for b, lg in zip(bsc.sing_block, bsc.sing_logical):
        gid = int(np.asarray(g.blocks[b])[tuple(lg)])
        slot = int(tab["g2s"][gid])
        assert base[b] <= slot < (base[b + 1] if b + 1 < len(base) else info["n_slots"])


@pytest.mark.gpu
@pytest.mark.parametrize("forced", [False, True])
def test_singular_hosts_parity(oracle, forced):
    """The O-grid on the interface tables with singular hosts (the default,
    or every singular node moved to another fan block): after an exchange
    every copy equals its owner, and 40 sweeps match the FP64 oracle."""
    from egg_amd.smoothing import CppStructuredSweepSession

    g = grids.ogrid_2d(33, 17)
    c = Case(g)
    bsc = T.build_block_structured_context(2, g.blocks, g.free)
    if forced:
        bsc = _forced_hosts(g, bsc)
    print(f"singular hosts: forced={forced} hosts={bsc.sing_block.tolist()}", flush=True)
    s = CppStructuredSweepSession(c.ctx, bsc, g.X, device="gpu")
    core = s.core
    core.halo_exchange()
    P = core.packed()
    tab = core.tables()
    assert np.array_equal(P[:, tab["copy_dst"]].view(np.uint32), P[:, tab["copy_src"]].view(np.uint32))
    e, m = s.run(40, omega=0.8, report_every=10)
    Xo, eo, mo = oracle.jacobi(2, "barrier", 0.0, 0.8, 40, g.X, c.patches(oracle), c.ctx.energy_stencil,
                               report_every=10)
    assert np.abs(s.get_X() - Xo).max() / 4.0 < 1e-5
    np.testing.assert_allclose(e, eo, rtol=1e-5)


def _level_maps(shapes_by_level, fine_shapes, d):
    """Per level: slot -> (block, padded index), and the per-block factor to
    the parent level (2 where a count halved)."""
    out = []
    parent = fine_shapes
    for shapes in shapes_by_level:
        base, loc = 0, {}
        for b, shp in enumerate(shapes):
            pad = [s + 2 for s in shp]
            for i, p in enumerate(np.ndindex(*pad)):
                loc[base + i] = (b, p)
            base += int(np.prod(pad))
        factor = [[2 if c != f else 1 for c, f in zip(cs, fs)] for cs, fs in zip(shapes, parent)]
        out.append((loc, factor, shapes))
        parent = shapes
    return out


@pytest.mark.parametrize("name,builder", [("ogrid", "twins"), ("ogrid", "faces"), ("cubed_sphere", "twins"),
                                          ("star5", "faces")])
def test_coarse_levels_fill_ghosts_from_coarse_neighbours(name, builder):
    """On every FAS level a ghost fill names the node one coarse step past
    the interface: its fine image lies in another block, across the face
    (at relaxed face nodes); every relaxed coarse DOF's fine image is a
    relaxed fine DOF, and every interior coarse DOF is strictly inside its
    block."""
    from egg_amd._cpp import cpp_core

    g = {"ogrid": grids.ogrid_2d(33, 17), "cubed_sphere": grids.cubed_sphere_3d(9), "star5": star(k=5, n=17).grid}[name]
    d = g.d
    c = Case(g)
    bsc = (T.build_structured_context_from_block_maps(d, g.blocks) if builder == "twins"
           else T.build_block_structured_context(d, g.blocks, g.free))
    s = cpp_core.CppStructuredSweepSession(c.ctx.wire, structured_arrays(bsc), g.X.ravel(), device="tables", dim=d)
    shapes = [[tuple(b) for b in lv] for lv in s.mg_levels()]
    assert shapes, "the grid coarsens"
    maps = _level_maps(shapes, [np.shape(b) for b in g.blocks], d)
    blocks = [np.asarray(b) for b in g.blocks]

    def fine_gid(level, b, logical):
        lg = np.asarray(logical)
        for lv in range(level, 0, -1):
            lg = lg * np.asarray(maps[lv - 1][1][b])
        return int(blocks[b][tuple(lg)])

    for level in range(1, len(shapes) + 1):
        t = s.mg_level_tables(level)
        loc, _, shp = maps[level - 1]
        checked = 0
        for e in range(t["n_halo"]):
            sb, sp = loc[int(t["copy_src"][e])]
            db, dp = loc[int(t["copy_dst"][e])]
            axes = [k for k in range(d) if dp[k] in (0, shp[db][k] + 1)]
            assert len(axes) == 1
            k = axes[0]
            face = np.asarray(dp) - 1
            face[k] = 0 if dp[k] == 0 else shp[db][k] - 1
            inner = face.copy()
            inner[k] += 1 if dp[k] == 0 else -1
            S = fine_gid(level, sb, np.asarray(sp) - 1)
            F, I = fine_gid(level, db, face), fine_gid(level, db, inner)
            assert S not in set(blocks[db].reshape(-1).tolist())
            if g.free[F]:
                assert np.dot(g.X[S] - g.X[F], g.X[F] - g.X[I]) > 0
                checked += 1
        assert checked > 0
        free, inter = np.asarray(t["free"]), np.asarray(t["free_interior"])
        for slot in np.flatnonzero(free):
            b, p = loc[int(slot)]
            lg = np.asarray(p) - 1
            assert g.free[fine_gid(level, b, lg)]
            if inter[slot]:
                assert np.all(lg >= 1) and np.all(lg <= np.asarray(shp[b]) - 2)
