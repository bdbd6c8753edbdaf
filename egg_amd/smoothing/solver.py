"""Sweep context for the structured path (reference egg/smoothing/solver.py:51-373).

The reference derives its context from a ``MultiBlockGrid``; here it is built
directly from per-block global-id arrays and a free mask, which is what the
wire needs (reference ``build_flat_context`` inputs).
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .flat_context import build_flat_context


@dataclass
class SweepContext:
    d: int
    wire: dict
    block_global_dof: list = field(default_factory=list)
    free_mask: np.ndarray | None = None

    @property
    def energy_stencil(self) -> dict:
        return self.wire["energy_stencil"]


def build_sweep_context(blocks: list, free_mask: np.ndarray, d: int, *, w_inv=None, weight=None,
                        entities=None, curvature=None, directional=None) -> SweepContext:
    """Context over ``blocks`` (per-block global-id arrays) with moving DOFs
    ``free_mask``; ``w_inv`` / ``weight`` / ``entities`` / ``curvature`` / ``directional`` as in
    :func:`build_flat_context`."""
    blocks = [np.asarray(b) for b in blocks]
    wire = build_flat_context(blocks, np.asarray(free_mask, bool), d, w_inv=w_inv, weight=weight,
                              entities=entities, curvature=curvature, directional=directional)
    return SweepContext(d=d, wire=wire, block_global_dof=blocks, free_mask=np.asarray(free_mask, bool))
