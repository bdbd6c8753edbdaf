"""Directional soft-energy terms on the wire (reference
egg/smoothing/directional.py:63-93 sample table, :372-458 wire layouts).

A sample ties up to four nodes (``-1`` pads) under one kind:
``KIND_PARALLEL`` (i, j): the segment X[j] - X[i] kept perpendicular to the
frozen normal ``ref``; ``KIND_LINE`` (a, c, b): the legs from c kept opposed;
``KIND_STEM`` (n1, c): the leg X[n1] - X[c] kept perpendicular to the frozen
through axis ``ref``. Kinds 3-5 are reserved for 3D-only terms; they occupy
no node slot and add nothing in the compute core.

The wire carries the table twice: whole, on the energy stencil (every
report and line search adds its total), and per moving DOF, on the group
(each sample the DOF takes part in, with the DOF's slot in it, padded to a
common width).
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .wire_tables import per_dof_slots

KIND_PARALLEL = 0
KIND_LINE = 1
KIND_STEM = 2
KIND_SHEET_NORMAL = 3
KIND_EDGE_LINE = 4
KIND_EDGE_STEM = 5
MAX_ARITY = 4

# occupied node slots per kind; the reserved kinds occupy none
_ARITY = np.array([2, 3, 2, 0, 0, 0], np.int32)

__all__ = ["DirectionalSamples", "KIND_LINE", "KIND_PARALLEL", "KIND_STEM", "directional_dof_table",
           "directional_energy_table", "kind_arity"]


def kind_arity(kind: int) -> int:
    """Occupied node slots of a sample kind (0 for the reserved kinds)."""
    return int(_ARITY[kind]) if 0 <= kind < _ARITY.size else 0


@dataclass
class DirectionalSamples:
    """The sample table, one row per sample (field names are the wire's)."""

    nodes: np.ndarray  # (n, MAX_ARITY) global node ids, -1 pads
    kind: np.ndarray   # (n,)
    ref: np.ndarray    # (n, d) frozen reference vector
    axis: np.ndarray   # (n, d) projector direction of the 3D-only kinds (zeros otherwise)
    weight: np.ndarray  # (n,)
    eps: float = 1e-12

    def __post_init__(self):
        self.nodes = np.asarray(self.nodes, np.int32).reshape(-1, MAX_ARITY)
        self.kind = np.asarray(self.kind, np.int32).reshape(-1)
        self.ref = np.asarray(self.ref, np.float64)
        self.axis = np.asarray(self.axis, np.float64).reshape(self.ref.shape)
        self.weight = np.asarray(self.weight, np.float64).reshape(-1)

    @property
    def n(self) -> int:
        return int(self.kind.shape[0])

    @property
    def d(self) -> int:
        return int(self.ref.shape[1])

    def arity(self) -> np.ndarray:
        """Occupied node slots of every sample."""
        k = self.kind
        return np.where((k >= 0) & (k < _ARITY.size), _ARITY[np.clip(k, 0, _ARITY.size - 1)], 0)


def directional_dof_table(free_mask, samples: DirectionalSamples | None) -> dict:
    """The per-DOF layout, in the group's DOF order (the moving nodes in
    ascending id): for each DOF the samples it occupies a slot of, in sample
    order, padded to the widest DOF's count ``dir_k`` (slot -1, nodes -1,
    kind / ref / axis / weight 0 on padding). Empty when nothing is packed."""
    if samples is None or samples.n == 0:
        return {}
    K, sample, slot = per_dof_slots(free_mask, samples.nodes, samples.arity())
    if K == 0:
        return {}
    pad = sample < 0
    s = np.where(pad, 0, sample)

    def take(a, fill):
        out = a[s].copy()
        out[pad] = fill
        return out

    return {"dir_k": K, "dir_eps": float(samples.eps),
            "dir_nodes": take(samples.nodes, -1).reshape(-1),
            "dir_slot": slot.astype(np.int32).reshape(-1),
            "dir_kind": take(samples.kind, 0).astype(np.int32).reshape(-1),
            "dir_ref": take(samples.ref, 0.0).reshape(-1),
            "dir_axis": take(samples.axis, 0.0).reshape(-1),
            "dir_weight": take(samples.weight, 0.0).reshape(-1)}


def directional_energy_table(samples: DirectionalSamples | None) -> dict:
    """The whole table the energy reductions add (``dir_num`` samples)."""
    if samples is None or samples.n == 0:
        return {}
    return {"dir_num": samples.n, "dir_eps": float(samples.eps),
            "dir_nodes": np.ascontiguousarray(samples.nodes).reshape(-1),
            "dir_kind": np.ascontiguousarray(samples.kind),
            "dir_ref": np.ascontiguousarray(samples.ref).reshape(-1),
            "dir_axis": np.ascontiguousarray(samples.axis).reshape(-1),
            "dir_weight": np.ascontiguousarray(samples.weight)}
