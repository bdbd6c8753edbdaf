"""Bridge to the MI355X compute core for the structured block-Jacobi sweep.

Mirrors the reference's ``egg/smoothing/cpp_backend.py`` surface for the
structured path: :class:`CppStructuredSweepSession` (:456-677),
:func:`cpp_structured_sweep` (:694-770) and :func:`cpp_untangle` (:773-848);
the halo-table builders (:56-455, :679-691) live in
:mod:`egg_amd.smoothing.structured_tables` and are re-exported here. The compute core is
``egg_amd._cpp.cpp_core`` (HIP kernels behind the eggmi C ABI); there is no
CPU fallback — importing the session without the built extension raises.
"""

from __future__ import annotations

from typing import Any

import numpy as np

from .structured_tables import (
    BlockStructuredContext,
    build_block_structured_context,
    build_structured_context_from_block_maps,
    structured_arrays,
)

__all__ = [
    "BlockStructuredContext",
    "CppStructuredSweepSession",
    "build_block_structured_context",
    "build_structured_context_from_block_maps",
    "cpp_structured_sweep",
    "cpp_untangle",
    "structured_arrays",
]


def _core():
    from egg_amd._cpp import cpp_core  # raises ImportError when the extension is not built

    return cpp_core


def _grid_mindet(X: np.ndarray, es: dict, d: int = 2) -> float:
    """Raw min det A over the energy stencil (host NumPy; reference :44-52)."""
    X = np.asarray(X, np.float64).reshape(-1, d)
    gc = np.asarray(es["gc"])
    cols = [np.asarray(es[f"s{k}"])[:, None] * (X[np.asarray(es[f"gn{k}"])] - X[gc]) for k in range(d)]
    A = np.stack(cols, axis=2)  # (ns, d, d), A[:, :, k] = column k
    return float(np.linalg.det(A).min()) if len(gc) else float("inf")


def _as_bsc(grid: Any, d: int) -> BlockStructuredContext:
    if isinstance(grid, BlockStructuredContext):
        return grid
    maps = getattr(grid, "block_dof_maps", grid)
    return build_structured_context_from_block_maps(d, list(maps))


class CppStructuredSweepSession:
    """Persistent device-resident structured smoothing session (reference :456-677).

    ``ctx`` carries ``.wire`` (the flat dict), ``.d`` and ``.energy_stencil``;
    ``bsc`` the halo tables; ``X`` (N, d) positions in global order.
    """

    def __init__(self, ctx, bsc: BlockStructuredContext, X: np.ndarray, *, device: str = "auto",
                 control: dict | None = None) -> None:
        cpp_core = _core()
        self._shape = np.shape(X)
        self._d = ctx.d
        self._session = cpp_core.CppStructuredSweepSession(
            ctx.wire, structured_arrays(bsc), np.ascontiguousarray(X, np.float64).ravel(),
            device=device, dim=ctx.d, control=control,
        )

    def run(self, n_sweeps: int, *, phase: str = "barrier", delta: float = 0.0, omega: float = 0.8,
            report_every: int = 0):
        """``n_sweeps`` of block-Jacobi; returns ``(energies, mindets)`` with one
        report per ``report_every`` sweeps plus the final one (final only when 0)."""
        return self._session.run(n_sweeps, phase=phase, delta=delta, omega=omega, report_every=report_every)

    def run_fas(self, n_cycles: int, *, nu_pre: int = 2, nu_post: int = 2, nu_coarse: int = 32,
                omega: float = 0.8, max_levels: int = 32, phase: str = "barrier"):
        """``n_cycles`` FAS V-cycles; returns per-cycle ``(energies, mindets)``."""
        return self._session.run_fas(n_cycles, nu_pre=nu_pre, nu_post=nu_post, nu_coarse=nu_coarse,
                                     omega=omega, max_levels=max_levels, phase=phase)

    def mg_levels(self) -> list:
        return self._session.mg_levels()

    def get_X(self) -> np.ndarray:
        return self._session.get_X().reshape(self._shape)

    def set_X(self, X: np.ndarray) -> None:
        self._session.set_X(np.ascontiguousarray(X, np.float64).ravel())

    def run_control(self, n_outer: int, *, phase: str = "barrier", grad_tol: float = 1e-8, alpha_min: float = 1e-4,
                    lm0: float = 0.0, pcg_max_iter: int = 200, pcg_rtol: float = 1e-10, pcg_forcing: bool = True,
                    model_db: bool = True) -> dict:
        """Up to ``n_outer`` control-net reduced Gauss-Newton iterations on
        the resident net (the session needs ``control=``; reference
        cpp_backend.py:591-641): the fine gradient pulled back to the free
        controls, ``(A + lam I) dC = -G`` by Jacobi-preconditioned
        matrix-free FP64 PCG with the exact per-sample Hessian action, a line
        search of ``M dC`` on the global fine energy (accept: finite,
        monotone to 1e-6, min det > 0), the damping raised on rejection.
        ``pcg_forcing`` solves early steps only to max(pcg_rtol, min(0.5,
        sqrt(max|G| / max|G|_0))). ``model_db`` concerns sliding wall
        controls, which are not supported, and is ignored. Returns
        ``{energies, mindets, alphas, frame_jumps, iters, converged,
        pcg_iters, pcg_relres}``."""
        return self._session.run_control(n_outer, phase=phase, grad_tol=grad_tol, alpha_min=alpha_min, lm0=lm0,
                                         pcg_max_iter=pcg_max_iter, pcg_rtol=pcg_rtol, pcg_forcing=pcg_forcing,
                                         model_db=model_db)

    def get_C(self) -> np.ndarray:
        """The full control lattice, shape ``(n_ctrl, d)`` (the nets' stacked
        lattices under a multi-block wire)."""
        return self._session.get_C().reshape(-1, self._d)

    def set_C(self, C: np.ndarray) -> None:
        """Upload a full control lattice ``(n_ctrl, d)`` and re-evaluate the
        net (+ b) into the resident state."""
        self._session.set_C(np.ascontiguousarray(C, np.float64).ravel())

    def set_control_reduction(self, nq: int, off, col, coef) -> None:
        """Replace the multi-block reduction dC = R dq (CSR over the stacked
        components, ``nq`` columns; reference bindings.cpp:1475-1494)."""
        self._session.set_control_reduction(int(nq), np.ascontiguousarray(off, np.int32),
                                            np.ascontiguousarray(col, np.int32), np.ascontiguousarray(coef, np.float64))

    def set_control_penalty(self, off, col, coef, w) -> None:
        """Replace the multi-block penalty rows, energy 1/2 sum w (P C)^2
        (reference bindings.cpp:1496-1516); empty ``off`` removes them."""
        self._session.set_control_penalty(np.ascontiguousarray(off, np.int32), np.ascontiguousarray(col, np.int32),
                                          np.ascontiguousarray(coef, np.float64), np.ascontiguousarray(w, np.float64))

    def set_control_b(self, block_index: int, b) -> None:
        """Replace net ``block_index``'s Boolean-sum offset (its block's node
        order, d values per node) and re-evaluate the nets into the state."""
        self._session.set_control_b(int(block_index), np.ascontiguousarray(b, np.float64).ravel())

    def set_proposal(self, mlp, *, mode: str = "replace", scale: float = 1.0, every: int = 1) -> None:
        """Use a trained ``cpp_core.ProposalMLP`` (or None) as the learned
        Newton-step proposal of subsequent :meth:`run` calls and of the
        fine-level smoothing sweeps of :meth:`run_fas`. ``mode`` "replace":
        an accepted proposal is the update; "predict_correct": the Newton
        step continues from it. ``every``: :meth:`run` proposes in sweeps 0,
        every, 2 every, ... of each call; with every > 1 a FAS fine
        smoothing call proposes in its first sweep only."""
        self._session.set_proposal(mlp, mode=mode, scale=scale, every=every)

    def harvest(self, mlp, n_snap: int, *, every: int = 8, horizon: int = 8, omega: float = 0.8) -> int:
        """Append (features, displacement over ``horizon`` plain sweeps) pairs
        of every interior DOF, at ``n_snap`` snapshots ``every`` sweeps apart,
        to the network's training set."""
        return self._session.harvest(mlp, n_snap, every=every, horizon=horizon, omega=omega)

    @property
    def core(self):
        """The underlying ``cpp_core.CppStructuredSweepSession``."""
        return self._session


def cpp_structured_sweep(ctx, grid, X: np.ndarray, n_sweeps: int, *, device: str = "auto",
                         phase: str = "barrier", delta: float = 0.0, omega: float = 0.8, report_every: int = 0):
    """One-shot block-Jacobi run (reference :694-770). ``grid`` is a
    :class:`BlockStructuredContext`, an object with ``block_dof_maps`` or a
    list of per-block global-id arrays. Returns ``(X_out, energies, mindets)``."""
    cpp_core = _core()
    bsc = _as_bsc(grid, ctx.d)
    X_out, e, m = cpp_core.cpp_structured_sweep(
        ctx.wire, structured_arrays(bsc), np.ascontiguousarray(X, np.float64).ravel(), n_sweeps,
        device=device, phase=phase, delta=delta, dim=ctx.d, omega=omega, report_every=report_every,
    )
    return X_out.reshape(np.shape(X)), e, m


def cpp_untangle(ctx, grid, X: np.ndarray, *, device: str = "auto", sweeps_per_delta: int = 20,
                 delta0_factor: float = 2.0, shrink: float = 0.8, max_outer: int = 60, margin: float = 1e-9,
                 omega: float = 0.5, session: CppStructuredSweepSession | None = None):
    """delta-continuation untangle (reference :773-848): ``sweeps_per_delta``
    untangle sweeps per step on ``delta_k = delta_0 * shrink**k`` from
    ``delta0_factor * |min det A|`` until ``min det A > margin``.
    Returns ``(X_out, mindet, outer_iters, delta_final)``; pass ``session``
    (whose resident state is X) to keep working on a resident store; its
    min det is then reduced on the device instead of the host."""
    if session is None:
        md = _grid_mindet(X, ctx.energy_stencil, ctx.d)
        if md > margin:
            return X, md, 0, 0.0
        session = CppStructuredSweepSession(ctx, _as_bsc(grid, ctx.d), X, device=device)
    else:
        # the session holds X: reduce its min det on the device
        md = float(session.core.energy(phase="barrier")[1])
        if md > margin:
            return X, md, 0, 0.0
    delta = delta0_factor * max(abs(md), 1e-12)
    outer = 0
    for _ in range(max_outer):
        _e, mds = session.run(sweeps_per_delta, phase="untangle", delta=delta, omega=omega)
        outer += 1
        md = float(np.asarray(mds)[-1])
        if md > margin:
            break
        delta *= shrink
    return session.get_X(), md, outer, delta
