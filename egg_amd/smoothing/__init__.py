"""Structured TMOP smoothing on the MI355X: wire builders and the core bridge."""

from .cpp_backend import (
    BlockStructuredContext,
    CppStructuredSweepSession,
    build_block_structured_context,
    build_structured_context_from_block_maps,
    cpp_structured_sweep,
    cpp_untangle,
    structured_arrays,
)
from .flat_context import build_flat_context, cell_stencil
from .solver import SweepContext, build_sweep_context

__all__ = [
    "BlockStructuredContext",
    "CppStructuredSweepSession",
    "SweepContext",
    "build_block_structured_context",
    "build_flat_context",
    "build_structured_context_from_block_maps",
    "build_sweep_context",
    "cell_stencil",
    "cpp_structured_sweep",
    "cpp_untangle",
    "structured_arrays",
]
