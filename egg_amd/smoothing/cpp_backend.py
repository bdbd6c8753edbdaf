"""Bridge to the MI355X compute core for the structured block-Jacobi sweep.

Mirrors the reference's ``egg/smoothing/cpp_backend.py`` surface for the
structured path: :func:`build_structured_context_from_block_maps`
(reference :120-250), :func:`structured_arrays` (:679-691),
:class:`CppStructuredSweepSession` (:456-677), :func:`cpp_structured_sweep`
(:694-770) and :func:`cpp_untangle` (:773-848). The compute core is
``egg_amd._cpp.cpp_core`` (HIP kernels behind the eggmi C ABI); there is no
CPU fallback — importing the session without the built extension raises.
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import Any

import numpy as np

__all__ = [
    "BlockStructuredContext",
    "CppStructuredSweepSession",
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


@dataclass
class BlockStructuredContext:
    """Halo-padded structured tables (reference cpp_backend.py:56-117)."""

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


def build_structured_context_from_block_maps(d: int, block_global_dof: list) -> BlockStructuredContext:
    """Face-ghost halo table by shared-global-id twin matching.

    For every face node of block ``bi`` that also lives in another block, the
    ghost just outside that face is filled from the node one step behind the
    twin, i.e. one step into the other block along the axis on which the twin
    sits on a face and whose inward step leaves ``bi``. A node behind the
    contact that is itself shared is sourced from its lowest (block, logical)
    copy — the owner — so a halo source is never a broadcast destination.
    Unresolvable or ambiguous contacts are left to the core's spare-ghost
    mirrors. Entries are emitted block by block, axis by axis, low side then
    high side, face nodes in row-major order (reference :120-250).
    """
    shapes = [tuple(int(s) for s in np.shape(b)) for b in block_global_dof]
    arrays = [np.asarray(b) for b in block_global_dof]
    flat = [np.ascontiguousarray(a.reshape(-1), dtype=np.int32) for a in arrays]
    n_ids = (max(int(f.max()) for f in flat if f.size) + 1) if flat else 0
    count = np.zeros(n_ids, np.int64)
    for f in flat:
        count += np.bincount(f, minlength=n_ids)
    shared = count > 1
    copies: dict[int, list] = {}
    for b, a in enumerate(arrays):
        for idx in zip(*np.nonzero(shared[a])):
            loc = tuple(int(i) for i in idx)
            copies.setdefault(int(a[loc]), []).append((b, loc))
    members = [set(f.tolist()) for f in flat]

    def source_of(dst: int, gid: int):
        found = None
        where = None
        for b, loc in copies.get(gid, ()):
            if b == dst:
                continue
            for ax in range(d):
                if loc[ax] == 0:
                    step = 1
                elif loc[ax] == shapes[b][ax] - 1:
                    step = -1
                else:
                    continue
                inner = list(loc)
                inner[ax] += step
                g_in = int(arrays[b][tuple(inner)])
                if g_in in members[dst]:
                    continue  # a seam direction, not the contact normal
                if found is not None and found != g_in:
                    return None
                found = g_in
                where = (b, tuple(inner))
        if found is None:
            return None
        if shared[found]:
            return min(copies[found])
        return where

    sb, sp, db, dp = [], [], [], []
    for b, a in enumerate(arrays):
        shp = shapes[b]
        for ax in range(d):
            for side in (0, 1):
                fixed = 0 if side == 0 else shp[ax] - 1
                face = np.take(a, fixed, axis=ax)
                for fidx in np.ndindex(face.shape):
                    gid = int(face[fidx])
                    if not shared[gid]:
                        continue
                    hit = source_of(b, gid)
                    if hit is None:
                        continue
                    logical = list(fidx[:ax]) + [fixed] + list(fidx[ax:])
                    ghost = [c + 1 for c in logical]
                    ghost[ax] = 0 if side == 0 else shp[ax] + 1
                    sb.append(hit[0])
                    sp.append(tuple(c + 1 for c in hit[1]))
                    db.append(b)
                    dp.append(tuple(ghost))
    rows = lambda v: np.asarray(v, np.int32).reshape(-1, d) if v else np.zeros((0, d), np.int32)  # noqa: E731
    return BlockStructuredContext(
        d=d,
        interior_shapes=shapes,
        block_global_dof=flat,
        halo_src_block=np.asarray(sb, np.int32),
        halo_src_padded=rows(sp),
        halo_dst_block=np.asarray(db, np.int32),
        halo_dst_padded=rows(dp),
        sing_block=np.zeros((0,), np.int32),
        sing_logical=np.zeros((0, d), np.int32),
    )


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

    def set_proposal(self, mlp, *, mode: str = "replace", scale: float = 1.0) -> None:
        """Use a trained ``cpp_core.ProposalMLP`` (or None) as the learned
        Newton-step proposal of subsequent :meth:`run` calls and of the
        fine-level smoothing sweeps of :meth:`run_fas`. ``mode`` "replace":
        an accepted proposal is the update; "predict_correct": the Newton
        step continues from it."""
        self._session.set_proposal(mlp, mode=mode, scale=scale)

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
