"""Array building blocks of the flat wire: the (cell, corner) sample table of
structured blocks and the per-DOF padded incidence tables of window terms.

Sample table (reference egg/smoothing/flat_context.py:64-133 fixes the
order, which the wire format depends on): blocks in order, each block's
cells row-major, each cell's 2^d corners with axis 0 as the most
significant bit; sample id = 2^d cell + corner. The sample at corner b is b
with its d axis neighbours b ^ bit(k), signs +1 where bit k of b is 0 and
-1 where it is 1. A node of cell c at corner a meets the cell's sample b in
role 0 (a = b), 1 + k (a = b ^ bit(k)) or -1 (otherwise).
"""

from __future__ import annotations

import numpy as np
from numpy.lib.stride_tricks import sliding_window_view

__all__ = ["CellSamples", "per_dof_slots"]


class CellSamples:
    """The (cell, corner) samples of a list of blocks of global node ids."""

    def __init__(self, blocks: list, d: int):
        self.d = d
        self.nc = 1 << d
        # cells[c, a]: node at corner a of cell c; a 2^d window of a block,
        # flattened row-major, lists the corners in that bit order already
        parts = [sliding_window_view(np.asarray(a), (2,) * d).reshape(-1, self.nc)
                 for a in blocks if min(np.shape(a)) >= 2]
        self.cells = np.concatenate(parts).astype(np.int64) if parts else np.zeros((0, self.nc), np.int64)
        corner = np.arange(self.nc)
        self.bits = [1 << (d - 1 - k) for k in range(d)]
        self.flip = [corner ^ b for b in self.bits]
        self.sign = [np.where(corner & b, -1.0, 1.0) for b in self.bits]
        x = corner[:, None] ^ corner[None, :]
        role = np.where(x == 0, 0, -1)
        for k, b in enumerate(self.bits):
            role = np.where(x == b, 1 + k, role)
        self.role = role.astype(np.int32)  # role[a, b]

    @property
    def n_cells(self) -> int:
        return self.cells.shape[0]

    @property
    def n_samples(self) -> int:
        return self.n_cells * self.nc

    def fields(self, sid: np.ndarray) -> dict:
        """gc, gn{k}, s{k} of the samples ``sid``."""
        c, b = np.divmod(np.asarray(sid, np.int64), self.nc)
        out = {"gc": self.cells[c, b].astype(np.int32)}
        for k in range(self.d):
            out[f"gn{k}"] = self.cells[c, self.flip[k][b]].astype(np.int32)
            out[f"s{k}"] = self.sign[k][b]
        return out

    def incidence(self, select: np.ndarray):
        """Every (node, sample, role) with the node selected (a boolean mask
        over node ids), grouped by node in ascending id and, per node, in
        ascending sample id."""
        cell, a = np.nonzero(np.asarray(select, bool)[self.cells])
        node = np.repeat(self.cells[cell, a], self.nc)
        sid = (cell[:, None] * self.nc + np.arange(self.nc)[None, :]).reshape(-1)
        role = self.role[a].reshape(-1)
        order = np.lexsort((sid, node))
        return node[order], sid[order], role[order]


def per_dof_slots(free_mask, nodes: np.ndarray, arity: np.ndarray):
    """Per moving DOF (ascending id), the windows it fills a slot of: window
    w holds ``nodes[w, :arity[w]]`` (-1 = none). Returns (K, window, slot),
    both ``(ndof, K)``, in window order then slot order per DOF, padded with
    -1 to the widest DOF's count K (0 when no DOF takes part)."""
    free = np.asarray(free_mask, bool)
    dofs = np.flatnonzero(free)
    nodes = np.asarray(nodes, np.int64)
    used = (np.arange(nodes.shape[1])[None, :] < np.asarray(arity)[:, None]) & (nodes >= 0)
    win, slot = np.nonzero(used)
    node = nodes[win, slot]
    keep = free[node]
    win, slot, row = win[keep], slot[keep], np.searchsorted(dofs, node[keep])
    order = np.argsort(row, kind="stable")
    win, slot, row = win[order], slot[order], row[order]
    count = np.bincount(row, minlength=dofs.size)
    K = int(count.max()) if count.size else 0
    rank = np.arange(row.size) - np.repeat(np.cumsum(count) - count, count)
    tw = np.full((dofs.size, K), -1, np.int64)
    ts = np.full((dofs.size, K), -1, np.int64)
    tw[row, rank] = win
    ts[row, rank] = slot
    return K, tw, ts
