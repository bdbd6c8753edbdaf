"""Halo tables of a conforming multi-block grid, built from its per-block
global node-id arrays: the ``structured`` dict of the compute core
(reference egg/smoothing/cpp_backend.py:56-117 ``BlockStructuredContext``,
:679-691 ``structured_arrays``).

A face-ghost entry names a ghost slot just outside a block face (padded
index, ``0`` or ``n + 1`` along the face normal) and the node that fills it,
the neighbour one step past the interface in the adjacent block. Two
builders, with the reference's semantics and their own construction (every
table is assembled with array operations over all node copies at once):

* :func:`build_structured_context_from_block_maps` (reference
  cpp_backend.py:120-250): contacts found from the id arrays alone. A face
  node's ghost is filled only where the other copies of the node agree on a
  single node "behind" the contact; the filling node is read at its owner
  copy. No singular-node table.
* :func:`build_block_structured_context` (reference cpp_backend.py:318-455):
  contacts are whole faces with equal node sets (the conforming interfaces
  a block topology declares), every face node of both faces gets its ghost,
  and the singular nodes (:func:`singular_nodes`) get a host block chosen
  for the room its spare ghost slots leave (:func:`singular_hosts`).
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

__all__ = [
    "BlockStructuredContext",
    "build_block_structured_context",
    "build_structured_context_from_block_maps",
    "face_connections",
    "singular_hosts",
    "singular_nodes",
    "structured_arrays",
]


@dataclass
class BlockStructuredContext:
    """The compute core's structured tables; field names are the wire keys."""

    d: int
    interior_shapes: list
    block_global_dof: list
    halo_src_block: np.ndarray
    halo_src_padded: np.ndarray
    halo_dst_block: np.ndarray
    halo_dst_padded: np.ndarray
    sing_block: np.ndarray
    sing_logical: np.ndarray

    @property
    def num_blocks(self) -> int:
        return len(self.interior_shapes)

    @property
    def num_halo_entries(self) -> int:
        return int(self.halo_dst_block.shape[0])


def structured_arrays(bsc: BlockStructuredContext) -> dict:
    """The ``structured`` dict of the core binding (reference :679-691)."""
    return {
        "interior_shapes": [np.asarray(s, np.int32) for s in bsc.interior_shapes],
        "block_global_dof": list(bsc.block_global_dof),
        "halo_src_block": bsc.halo_src_block,
        "halo_src_padded": bsc.halo_src_padded,
        "halo_dst_block": bsc.halo_dst_block,
        "halo_dst_padded": bsc.halo_dst_padded,
        "sing_block": bsc.sing_block,
        "sing_logical": bsc.sing_logical,
    }


class _Copies:
    """Every (block, node) copy of the grid, blocks in order and each block
    row-major: its global id, block, logical coordinates and the index of the
    same block's copy one step along an axis."""

    def __init__(self, d: int, blocks: list):
        self.d = d
        self.arrays = [np.asarray(b) for b in blocks]
        self.shapes = [tuple(int(s) for s in a.shape) for a in self.arrays]
        sizes = np.array([a.size for a in self.arrays], np.int64)
        self.base = np.concatenate([[0], np.cumsum(sizes)])
        self.gid = (np.concatenate([a.reshape(-1) for a in self.arrays]).astype(np.int64)
                    if self.arrays else np.zeros(0, np.int64))
        self.blk = np.repeat(np.arange(len(self.arrays)), sizes)
        self.coord = (np.concatenate([np.indices(s).reshape(d, -1).T for s in self.shapes])
                      if self.arrays else np.zeros((0, d), np.int64))
        self.extent = (np.array(self.shapes, np.int64)[self.blk] if self.arrays else np.zeros((0, d), np.int64))
        strides = np.array([[int(np.prod(s[k + 1:])) for k in range(d)] for s in self.shapes], np.int64)
        self.stride = strides[self.blk] if self.arrays else np.zeros((0, d), np.int64)
        self.n_ids = int(self.gid.max()) + 1 if self.gid.size else 0
        self.count = np.bincount(self.gid, minlength=self.n_ids)
        # the owner copy of a node: its first copy in (block, row-major) order
        self.owner = np.full(self.n_ids, -1, np.int64)
        uniq, first = np.unique(self.gid, return_index=True)
        self.owner[uniq] = first
        self._keys = np.sort(self.blk * max(self.n_ids, 1) + self.gid)

    def in_block(self, blk: np.ndarray, gid: np.ndarray) -> np.ndarray:
        """Whether block ``blk`` holds node ``gid`` (elementwise)."""
        key = np.asarray(blk, np.int64) * max(self.n_ids, 1) + np.asarray(gid, np.int64)
        pos = np.searchsorted(self._keys, key)
        hit = pos < self._keys.size
        hit[hit] = self._keys[pos[hit]] == key[hit]
        return hit

    def padded(self, copy: np.ndarray) -> np.ndarray:
        return self.coord[copy] + 1

    def face(self, b: int, axis: int, side: int) -> np.ndarray:
        """Copies on face (axis, side) of block b, in row-major order."""
        lo, hi = self.base[b], self.base[b + 1]
        fixed = 0 if side == 0 else self.shapes[b][axis] - 1
        return lo + np.flatnonzero(self.coord[lo:hi, axis] == fixed)

    def ghost(self, copy: np.ndarray, axis: int, side: int) -> np.ndarray:
        """Padded index of the ghost just outside face (axis, side) next to copy."""
        g = self.coord[copy] + 1
        g[:, axis] = 0 if side == 0 else self.extent[copy, axis] + 1
        return g


def _tables(d, cp: _Copies, src_copy, dst_blk, dst_padded, sing_block=None, sing_logical=None):
    rows = lambda a: np.asarray(a, np.int32).reshape(-1, d)  # noqa: E731
    return BlockStructuredContext(
        d=d,
        interior_shapes=list(cp.shapes),
        block_global_dof=[np.ascontiguousarray(a.reshape(-1), dtype=np.int32) for a in cp.arrays],
        halo_src_block=np.asarray(cp.blk[src_copy], np.int32),
        halo_src_padded=rows(cp.padded(src_copy)),
        halo_dst_block=np.asarray(dst_blk, np.int32),
        halo_dst_padded=rows(dst_padded),
        sing_block=np.asarray(sing_block if sing_block is not None else [], np.int32),
        sing_logical=rows(sing_logical if sing_logical is not None else np.zeros((0, d))),
    )


def build_structured_context_from_block_maps(d: int, block_global_dof: list) -> BlockStructuredContext:
    """Face-ghost tables found from the id arrays alone (reference
    cpp_backend.py:120-250).

    For a shared node on a face of block ``b``, every copy of it in another
    block that lies on one of its own block's faces proposes the node one
    step inward from that face, unless ``b`` holds that node too (then the
    step runs along the contact, not across it). The ghost just outside
    ``b``'s face is filled when exactly one distinct node is proposed, from
    that node's owner copy; otherwise (no proposal, or a self-contact /
    singular neighbourhood proposing several) it is left to the core's
    spare-ghost mirrors. Entries run block by block, axis by axis, low side
    then high side, face nodes row-major; no singular-node table."""
    cp = _Copies(d, block_global_dof)
    shared = cp.count > 1
    # proposals: (node, proposing block, node one step inward) from every
    # shared copy on a face; the low face wins when a block is one node wide
    prop_gid, prop_blk, prop_in = [], [], []
    on_shared = np.flatnonzero(shared[cp.gid]) if cp.gid.size else np.zeros(0, np.int64)
    for axis in range(d):
        c = cp.coord[on_shared, axis]
        n = cp.extent[on_shared, axis]
        step = np.where(c == 0, 1, np.where(c == n - 1, -1, 0))
        ok = (step != 0) & (c + step >= 0) & (c + step < n)
        src = on_shared[ok]
        inner = src + step[ok] * cp.stride[src, axis]
        prop_gid.append(cp.gid[src])
        prop_blk.append(cp.blk[src])
        prop_in.append(cp.gid[inner])
    prop_gid = np.concatenate(prop_gid) if prop_gid else np.zeros(0, np.int64)
    prop_blk = np.concatenate(prop_blk) if prop_blk else np.zeros(0, np.int64)
    prop_in = np.concatenate(prop_in) if prop_in else np.zeros(0, np.int64)
    order = np.argsort(prop_gid, kind="stable")
    prop_gid, prop_blk, prop_in = prop_gid[order], prop_blk[order], prop_in[order]

    # face occurrences in emission order: (face copy, axis, side)
    occ, occ_axis, occ_side = [], [], []
    for b in range(len(cp.shapes)):
        for axis in range(d):
            for side in (0, 1):
                f = cp.face(b, axis, side)
                f = f[shared[cp.gid[f]]]
                occ.append(f)
                occ_axis.append(np.full(f.size, axis))
                occ_side.append(np.full(f.size, side))
    occ = np.concatenate(occ) if occ else np.zeros(0, np.int64)
    occ_axis = np.concatenate(occ_axis) if occ_axis else np.zeros(0, np.int64)
    occ_side = np.concatenate(occ_side) if occ_side else np.zeros(0, np.int64)

    # every proposal for every occurrence's node, then the valid ones
    g = cp.gid[occ]
    lo = np.searchsorted(prop_gid, g, "left")
    cnt = np.searchsorted(prop_gid, g, "right") - lo
    who = np.repeat(np.arange(occ.size), cnt)
    row = np.repeat(lo - np.concatenate([[0], np.cumsum(cnt)[:-1]]), cnt) + np.arange(cnt.sum())
    dst_b = cp.blk[occ][who]
    valid = (prop_blk[row] != dst_b) & ~cp.in_block(dst_b, prop_in[row])
    who, cand = who[valid], prop_in[row][valid]
    nval = np.bincount(who, minlength=occ.size)
    cmin = np.full(occ.size, np.iinfo(np.int64).max)
    cmax = np.full(occ.size, -1)
    np.minimum.at(cmin, who, cand)
    np.maximum.at(cmax, who, cand)
    hit = (nval > 0) & (cmin == cmax)

    occ, occ_axis, occ_side, behind = occ[hit], occ_axis[hit], occ_side[hit], cmin[hit]
    dst_padded = np.zeros((occ.size, d), np.int64)
    for axis in range(d):
        for side in (0, 1):
            m = (occ_axis == axis) & (occ_side == side)
            dst_padded[m] = cp.ghost(occ[m], axis, side)
    return _tables(d, cp, cp.owner[behind], cp.blk[occ], dst_padded)


def face_connections(d: int, block_global_dof: list) -> list:
    """Conforming interfaces: pairs of block faces ((b, axis, side), (b', axis',
    side')) holding the same node set, ordered by the first face. Faces of a
    block one node wide along their normal are never part of one."""
    cp = _Copies(d, block_global_dof)
    by_set: dict = {}
    order = []
    for b in range(len(cp.shapes)):
        for axis in range(d):
            if cp.shapes[b][axis] < 2:
                continue
            for side in (0, 1):
                key = np.sort(cp.gid[cp.face(b, axis, side)]).tobytes()
                if key not in by_set:
                    order.append(key)
                by_set.setdefault(key, []).append((b, axis, side))
    return [tuple(by_set[k]) for k in order if len(by_set[k]) == 2]


def singular_nodes(d: int, block_global_dof: list, free_mask: np.ndarray) -> list:
    """Free nodes whose number of distinct grid-line neighbours (over every
    block holding them) is above 2d, or below 2d away from the domain
    boundary (reference egg/topology/block_topology.py:587-743). A node is on
    the boundary when some block puts it on a face that is not part of a
    conforming interface. Returns ``[(block, logical, valence), ...]`` at each
    node's first copy, in (block, row-major) order."""
    cp = _Copies(d, block_global_dof)
    free = np.asarray(free_mask, bool)
    ends = []
    for a in cp.arrays:
        for axis in range(d):
            if a.shape[axis] < 2:
                continue
            u = np.take(a, np.arange(a.shape[axis] - 1), axis=axis).reshape(-1)
            v = np.take(a, np.arange(1, a.shape[axis]), axis=axis).reshape(-1)
            ends.append(np.stack([np.minimum(u, v), np.maximum(u, v)], 1))
    edges = np.unique(np.concatenate(ends), axis=0) if ends else np.zeros((0, 2), np.int64)
    valence = np.bincount(edges.reshape(-1), minlength=cp.n_ids)
    interface = {f for conn in face_connections(d, block_global_dof) for f in conn}
    boundary = np.zeros(cp.n_ids, bool)
    for b in range(len(cp.shapes)):
        for axis in range(d):
            for side in (0, 1):
                if (b, axis, side) not in interface:
                    boundary[cp.gid[cp.face(b, axis, side)]] = True
    nodes = np.flatnonzero(free[: cp.n_ids] & (cp.owner >= 0))
    vn = valence[nodes]
    sing = nodes[(vn > 2 * d) | ((vn < 2 * d) & ~boundary[nodes])]
    first = np.sort(cp.owner[sing])
    return [(int(cp.blk[c]), tuple(int(x) for x in cp.coord[c]), int(valence[cp.gid[c]])) for c in first]


def singular_hosts(d: int, block_global_dof: list, sings: list, dst_block, dst_padded) -> list:
    """Host block of every singular node (reference cpp_backend.py:253-315):
    its first copy's block (the core's default owner) unless the fan blocks'
    spare ghost-ring slots say otherwise. A block's spare slots are its ghost
    ring minus the distinct face-ghost destinations of the halo table; a node
    of valence v may need v of them. Nodes are placed in decreasing valence
    (ties in table order); a node keeps the lowest-numbered block of its fan
    if that block has v spare slots, else moves to the fan block with the
    most spare slots, which then holds v fewer. The fan of a node is every
    block having it as a block corner."""
    cp = _Copies(d, block_global_dof)
    spare = []
    dst_block = np.asarray(dst_block)
    dst_padded = np.asarray(dst_padded).reshape(-1, d)
    for b, shp in enumerate(cp.shapes):
        ring = int(np.prod([s + 2 for s in shp]) - np.prod(shp))
        used = np.unique(dst_padded[dst_block == b], axis=0).shape[0] if np.any(dst_block == b) else 0
        spare.append(ring - used)
    corner_copy = np.all((cp.coord == 0) | (cp.coord == cp.extent - 1), axis=1)
    fan: dict = {}
    for c in np.flatnonzero(corner_copy):  # (block, row-major): the corners of each block in order
        fan.setdefault(int(cp.gid[c]), []).append((int(cp.blk[c]), tuple(int(x) for x in cp.coord[c])))
    hosts = [None] * len(sings)
    for i in sorted(range(len(sings)), key=lambda k: -sings[k][2]):
        b0, lg0, val = sings[i]
        gid = int(cp.arrays[b0][lg0])
        cands = fan.get(gid, [(b0, lg0)])
        lowest = min(cands, key=lambda c: c[0])
        if spare[lowest[0]] >= val:
            hosts[i] = lowest
        else:
            roomiest = max(cands, key=lambda c: spare[c[0]])
            hosts[i] = roomiest
            spare[roomiest[0]] -= val
    return hosts


def build_block_structured_context(d: int, block_global_dof: list, free_mask: np.ndarray) -> BlockStructuredContext:
    """Tables of a grid whose interfaces are whole conforming faces
    (reference cpp_backend.py:318-455, which reads them from a block
    topology; here they are the :func:`face_connections`): for every face
    node of both faces of an interface, the ghost outside its face is filled
    from the node one step inward from the twin's face in the other block,
    one face's ghosts then the other's per interface; plus the singular
    nodes and their hosts."""
    cp = _Copies(d, block_global_dof)
    src_copy, dst_blk, dst_pad = [], [], []
    for fa, fb in face_connections(d, block_global_dof):
        ca, cb = cp.face(*fa), cp.face(*fb)
        # twin of every face-a copy on face b, by node id
        ob = np.argsort(cp.gid[cb], kind="stable")
        twin = cb[ob[np.searchsorted(cp.gid[cb][ob], cp.gid[ca])]]
        for here, there, (b, axis, side), (_, taxis, tside) in ((ca, twin, fa, fb), (twin, ca, fb, fa)):
            step = 1 if tside == 0 else -1
            src_copy.append(there + step * cp.stride[there, taxis])
            dst_blk.append(np.full(here.size, b))
            dst_pad.append(cp.ghost(here, axis, side))
    src_copy = np.concatenate(src_copy) if src_copy else np.zeros(0, np.int64)
    dst_blk = np.concatenate(dst_blk) if dst_blk else np.zeros(0, np.int64)
    dst_pad = np.concatenate(dst_pad) if dst_pad else np.zeros((0, d), np.int64)
    sings = singular_nodes(d, block_global_dof, free_mask)
    hosts = singular_hosts(d, block_global_dof, sings, dst_blk, dst_pad)
    return _tables(d, cp, src_copy, dst_blk, dst_pad, [h[0] for h in hosts],
                   np.asarray([h[1] for h in hosts], np.int64).reshape(-1, d))
