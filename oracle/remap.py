"""Restatement of the reference's structured remap (TEST INFRASTRUCTURE ONLY).

Follows ``apply_structured_remap`` (reference src/bindings.cpp:758-1101) and
``BlockTopologyDevice`` (src/structured_halo.hpp:70-124) in plain Python, in
node units (the reference's real offsets divided by dim):

* blocks padded by one ghost layer and packed back to back, row-major;
* per block one dict ``block_map[b]: global id -> slot``: interior slots
  first (pass 1), then halo destinations (which override), then spare-ghost
  or overflow mirrors allocated on demand by ``lookup``;
* owner = lowest block containing the node, overridden by the singular
  host table; share = every non-owner interior copy, in block / row-major
  order; fan mirrors sourced from the owner slot (g2s);
* the device copy table is halo entries, then fan entries, then shares.

Used by tests only, as the reference the core's index maps are compared with.
"""

from __future__ import annotations

from itertools import product

import numpy as np


def remap(dim, shapes, bgd, halo_src_block, halo_src_padded, halo_dst_block, halo_dst_padded,
          sing_block=(), sing_logical=(), groups=(), n_nodes=None):
    shapes = [tuple(int(v) for v in s) for s in shapes]
    bgd = [np.asarray(b).reshape(-1) for b in bgd]
    base = []
    acc = 0
    for s in shapes:
        base.append(acc)
        acc += int(np.prod([v + 2 for v in s]))
    padded_total = acc

    def strides(b):
        st = [0] * dim
        a = 1
        for k in reversed(range(dim)):
            st[k] = a
            a *= shapes[b][k] + 2
        return st

    def padded_slot(b, p):
        st = strides(b)
        return base[b] + sum(int(p[k]) * st[k] for k in range(dim))

    def interior_slot(b, l):
        return padded_slot(b, [int(v) + 1 for v in l])

    M = n_nodes if n_nodes is not None else max(int(b.max()) for b in bgd) + 1
    block_map = [dict() for _ in shapes]
    owner_block = [-1] * M
    owner_slot = [-1] * M
    for b, s in enumerate(shapes):
        for f, l in enumerate(product(*[range(v) for v in s])):
            gid = int(bgd[b][f])
            sl = interior_slot(b, l)
            block_map[b][gid] = sl
            if owner_block[gid] < 0:
                owner_block[gid] = b
                owner_slot[gid] = sl
    for sb, sl in zip(sing_block, sing_logical):
        sb = int(sb)
        f = int(np.ravel_multi_index(tuple(int(v) for v in sl), shapes[sb]))
        gid = int(bgd[sb][f])
        owner_block[gid] = sb
        owner_slot[gid] = interior_slot(sb, sl)
    share_src, share_dst = [], []
    for b, s in enumerate(shapes):
        for f, l in enumerate(product(*[range(v) for v in s])):
            gid = int(bgd[b][f])
            if owner_block[gid] == b:
                continue
            share_src.append(owner_slot[gid])
            share_dst.append(interior_slot(b, l))
    halo_src, halo_dst = [], []
    used = [set() for _ in shapes]
    for e in range(len(halo_src_block)):
        sb, db = int(halo_src_block[e]), int(halo_dst_block[e])
        sp = [int(v) for v in halo_src_padded[e]]
        dp = [int(v) for v in halo_dst_padded[e]]
        f = int(np.ravel_multi_index(tuple(v - 1 for v in sp), shapes[sb]))
        gid = int(bgd[sb][f])
        ds = padded_slot(db, dp)
        block_map[db][gid] = ds
        used[db].add(ds)
        halo_src.append(padded_slot(sb, sp))
        halo_dst.append(ds)
    g2s = [-1] * M
    for g in range(M):
        ob = owner_block[g]
        if ob >= 0:
            g2s[g] = block_map[ob][g]
    free = []
    for b, s in enumerate(shapes):
        ring = []
        for p in product(*[range(v + 2) for v in s]):
            if any(p[k] == 0 or p[k] == s[k] + 1 for k in range(dim)):
                sl = padded_slot(b, p)
                if sl not in used[b]:
                    ring.append(sl)
        free.append(ring)
    cursor = [0] * len(shapes)
    fan_src, fan_dst = [], []
    n_over = [0]

    def lookup(b, gid):
        if gid in block_map[b]:
            return block_map[b][gid]
        src = g2s[gid]
        if src < 0:
            raise ValueError("absent from every block's interior map")
        if cursor[b] < len(free[b]):
            sl = free[b][cursor[b]]
            cursor[b] += 1
        else:
            sl = padded_total + n_over[0]
            n_over[0] += 1
        block_map[b][gid] = sl
        fan_src.append(src)
        fan_dst.append(sl)
        return sl

    stored = []  # per stored DOF: (dof slot, [(gc, gn...), ...])
    for g in groups:
        P_of = np.asarray(g["P_of"])
        off = np.concatenate([[0], np.cumsum(P_of)])
        for d_i, gid in enumerate(np.asarray(g["dof_idx"])):
            ob = owner_block[int(gid)]
            occ = []
            for p in range(int(off[d_i]), int(off[d_i + 1])):
                row = [lookup(ob, int(g["gc"][p]))]
                for k in range(dim):
                    row.append(lookup(ob, int(g[f"gn{k}"][p])))
                occ.append(tuple(row))
            ds = lookup(ob, int(gid))
            if int(np.asarray(g["interior_block"])[d_i]) < 0:
                stored.append((ds, occ))
    return {
        "copy_src": np.asarray(halo_src + fan_src + share_src, np.int64),
        "copy_dst": np.asarray(halo_dst + fan_dst + share_dst, np.int64),
        "n_halo": len(halo_src), "n_fan": len(fan_src), "n_share": len(share_src),
        "g2s": np.asarray(g2s, np.int64), "n_overflow": n_over[0], "base": np.asarray(base, np.int64),
        "n_slots": padded_total + n_over[0], "stored": stored,
    }
