"""Directional soft-energy terms: the sample table and its two wire layouts
(reference egg/smoothing/directional.py:63-93, :372-458).

Kinds: ``KIND_PARALLEL`` (nodes i, j: the segment X[j] - X[i] is kept
perpendicular to the frozen normal ``ref``), ``KIND_LINE`` (nodes a, c, b:
the legs from c are kept opposed), ``KIND_STEM`` (nodes n1, c: the leg
X[n1] - X[c] is kept perpendicular to the frozen through axis ``ref``).
Kinds 3-5 are reserved for 3D-only terms and add nothing in the compute core.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

KIND_PARALLEL = 0
KIND_LINE = 1
KIND_STEM = 2
KIND_SHEET_NORMAL = 3
KIND_EDGE_LINE = 4
KIND_EDGE_STEM = 5
MAX_ARITY = 4

__all__ = ["DirectionalSamples", "KIND_PARALLEL", "KIND_LINE", "KIND_STEM", "kind_arity",
           "directional_dof_table", "directional_energy_table"]


@dataclass
class DirectionalSamples:
    """SoA table of directional samples, one layout for every kind."""

    nodes: np.ndarray  # (n, MAX_ARITY) int32, -1 padded
    kind: np.ndarray  # (n,) int32
    ref: np.ndarray  # (n, d) frozen reference vector
    axis: np.ndarray  # (n, d) frozen projector direction (3D-only kinds; zeros)
    weight: np.ndarray  # (n,)
    eps: float = 1e-12

    @property
    def n(self) -> int:
        return int(self.kind.shape[0])

    @property
    def d(self) -> int:
        return int(self.ref.shape[1])


def kind_arity(kind: int) -> int:
    """Occupied node slots of a kind (0 for the reserved kinds)."""
    if kind == KIND_LINE:
        return 3
    if kind in (KIND_PARALLEL, KIND_STEM):
        return 2
    return 0


def directional_dof_table(free_mask, samples: DirectionalSamples | None) -> dict:
    """Per-moving-DOF table in the group's DOF order (``np.flatnonzero(free_mask)``):
    each sample a DOF takes part in, with the DOF's slot, padded to a common
    width ``dir_k`` (slot -1 = padding). Node ids are global. Empty dict when
    there is nothing to pack."""
    free = np.asarray(free_mask, bool)
    dofs = np.flatnonzero(free)
    if samples is None or samples.n == 0 or dofs.size == 0:
        return {}
    pos = np.full(free.shape[0], -1, np.int64)
    pos[dofs] = np.arange(dofs.size)
    lists: list[list[tuple[int, int]]] = [[] for _ in range(dofs.size)]
    for si in range(samples.n):
        for slot in range(kind_arity(int(samples.kind[si]))):
            node = int(samples.nodes[si, slot])
            if node >= 0 and free[node]:
                lists[int(pos[node])].append((si, slot))
    K = max((len(x) for x in lists), default=0)
    if K == 0:
        return {}
    d = samples.d
    nodes = np.full((dofs.size, K, MAX_ARITY), -1, np.int32)
    slot = np.full((dofs.size, K), -1, np.int32)
    kind = np.zeros((dofs.size, K), np.int32)
    ref = np.zeros((dofs.size, K, d))
    axis = np.zeros((dofs.size, K, d))
    weight = np.zeros((dofs.size, K))
    for i, entry in enumerate(lists):
        for j, (si, sl) in enumerate(entry):
            nodes[i, j], slot[i, j], kind[i, j] = samples.nodes[si], sl, samples.kind[si]
            ref[i, j], axis[i, j], weight[i, j] = samples.ref[si], samples.axis[si], samples.weight[si]
    return {"dir_k": int(K), "dir_eps": float(samples.eps), "dir_nodes": nodes.reshape(-1),
            "dir_slot": slot.reshape(-1), "dir_kind": kind.reshape(-1), "dir_ref": ref.reshape(-1),
            "dir_axis": axis.reshape(-1), "dir_weight": weight.reshape(-1)}


def directional_energy_table(samples: DirectionalSamples | None) -> dict:
    """The whole-sample table the energy reports add (``dir_num`` samples)."""
    if samples is None or samples.n == 0:
        return {}
    return {"dir_num": int(samples.n), "dir_eps": float(samples.eps),
            "dir_nodes": np.ascontiguousarray(samples.nodes, np.int32).reshape(-1),
            "dir_kind": np.ascontiguousarray(samples.kind, np.int32),
            "dir_ref": np.ascontiguousarray(samples.ref, np.float64).reshape(-1),
            "dir_axis": np.ascontiguousarray(samples.axis, np.float64).reshape(-1),
            "dir_weight": np.ascontiguousarray(samples.weight, np.float64)}
