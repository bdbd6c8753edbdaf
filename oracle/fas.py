"""FAS V-cycle oracle for one structured block (TEST INFRASTRUCTURE ONLY).

Restates the reference's multigrid for a single block in global-id space:
hierarchy shapes and coarse DOF selection (src/mg_level.hpp:130-230: an
axis halves iff its node count is odd and >= 5; a coarse node relaxes iff
its fine image relaxes as an interior node and every node of its
full-weighting stencil relaxes), transfers (src/mg_transfer.hpp:116-341:
injection, (1/4, 1/2, 1/4) full weighting of the fine gradient,
multilinear prolongation), the recursive coarse leg and the safeguarded
fine line-search ladder of run_fas (src/fas.hpp:251-602). Coarse levels of
one block have no interfaces, so their DOFs are interior nodes and the
sweeps are the oracle's tau-shifted block-Jacobi.
"""

from __future__ import annotations

from itertools import product

import numpy as np

from . import oracle as O


class _Level:
    def __init__(self, ids, free, free_int, factor=None):
        self.ids = ids  # node ids of this level's block (ndarray of shape)
        self.free = free  # bool per node
        self.free_int = free_int
        self.factor = factor  # per axis, vs the parent
        self.patches = None


def _interior_patches(dim, ids, dofs_logical):
    dof, poff, gc, gn, s, role = [], [0], [], [[] for _ in range(dim)], [[] for _ in range(dim)], []
    for lg in dofs_logical:
        g, n, sg, r = O.interior_occurrences(dim, lg, lambda l: int(ids[l]))
        dof.append(int(ids[tuple(lg)]))
        gc += g
        for k in range(dim):
            gn[k] += n[k]
            s[k] += sg[k]
        role += r
        poff.append(poff[-1] + len(g))
    return O.Patches(dim, dof, poff, gc, gn, s, role, np.eye(dim).ravel())


def build_levels(dim, ids, free, max_levels=32, synth=None):
    """Level 0 = the fine block; returns the list of levels. ``synth`` (bool
    per node) overrides which fine nodes count as synthesized interior DOFs
    (a DOF carrying curvature or directional windows ships a stored patch,
    so its coarse image is no DOF)."""
    shape = ids.shape
    free_int = np.zeros(ids.size, bool)
    inner = ids[tuple(slice(1, n - 1) for n in shape)].reshape(-1)
    free_int[inner] = free[inner]
    if synth is not None:
        free_int &= np.asarray(synth, bool)
    levels = [_Level(ids, free.copy(), free_int)]
    for _ in range(1, max_levels):
        p = levels[-1]
        pshape = p.ids.shape
        fac = [2 if (n % 2 == 1 and n >= 5) else 1 for n in pshape]
        if all(f == 1 for f in fac):
            break
        cshape = tuple((n + 1) // 2 if f == 2 else n for n, f in zip(pshape, fac))
        cids = np.arange(int(np.prod(cshape))).reshape(cshape)
        cfree = np.zeros(cids.size, bool)
        for I in product(*[range(n) for n in cshape]):
            F = tuple(i * f for i, f in zip(I, fac))
            if not p.free_int[p.ids[F]]:
                continue
            ok = True
            for o in product(*[(-1, 0, 1)] * dim):
                if any(fk == 1 and ok_ != 0 for fk, ok_ in zip(fac, o)):
                    continue
                P = tuple(Fk + ok_ for Fk, ok_ in zip(F, o))
                ok = ok and bool(p.free[p.ids[P]])
            if ok:
                cfree[cids[I]] = True
        if not cfree.any():
            break
        levels.append(_Level(cids, cfree, cfree.copy(), fac))
    for lv in levels:
        logical = [tuple(int(v) for v in np.unravel_index(i, lv.ids.shape)) for i in np.flatnonzero(lv.free)]
        lv.patches = _interior_patches(dim, lv.ids, logical)
    return levels


def _inject(p, c, xp):
    xc = np.zeros((c.ids.size, xp.shape[1]))
    for I in product(*[range(n) for n in c.ids.shape]):
        F = tuple(i * f for i, f in zip(I, c.factor))
        xc[c.ids[I]] = xp[p.ids[F]]
    return xc


def _full_weight(p, c, gp):
    out = np.zeros((c.ids.size, gp.shape[1]))
    dim = gp.shape[1]
    for I in product(*[range(n) for n in c.ids.shape]):
        if not c.free[c.ids[I]]:
            continue
        F = [i * f for i, f in zip(I, c.factor)]
        acc = np.zeros(dim)
        for o in product(*[(-1, 0, 1)] * dim):
            if any(f == 1 and ok != 0 for f, ok in zip(c.factor, o)):
                continue
            w = np.prod([(0.5 if ok == 0 else 0.25) for f, ok in zip(c.factor, o) if f == 2])
            acc += w * gp[p.ids[tuple(Fk + ok for Fk, ok in zip(F, o))]]
        out[c.ids[I]] = acc
    return out


def _prolong(p, c, cc):
    out = np.zeros((p.ids.size, cc.shape[1]))
    for J in product(*[range(n) for n in p.ids.shape]):
        base = [j // f for j, f in zip(J, c.factor)]
        odd = [j % f for j, f in zip(J, c.factor)]
        m = sum(1 for r in odd if r)
        w0 = 0.5**m
        acc = 0.0
        for e in product(*[(0, 1) if r else (0,) for r in odd]):
            acc = acc + w0 * cc[c.ids[tuple(b + ee for b, ee in zip(base, e))]]
        out[p.ids[J]] = acc
    return out


def _stencil(dim, ids):
    from egg_amd.smoothing import build_flat_context

    w = build_flat_context([ids], np.ones(ids.size, bool), dim)
    return w["energy_stencil"]


def run_fas(dim, kind, ids, free, X, *, n_cycles, nu_pre=2, nu_post=2, nu_coarse=32, omega=1.0, max_levels=32,
            prec="f64", fine_patches=None, fine_stencil=None, synth=None):
    """Returns (X_out, energies, mindets) per cycle. ``fine_patches`` /
    ``fine_stencil`` replace the fine level's synthesized patches and plain
    cell stencil (a wire with soft terms: its windows enter the fine sweeps,
    the fine gradient that is restricted, and the line search); ``synth``
    as in :func:`build_levels`."""
    tol = 1e-12 if prec == "f64" else 1e-6
    levels = build_levels(dim, ids, free, synth=synth)
    if fine_patches is not None:
        levels[0].patches = fine_patches
    n_coarse = min(len(levels) - 1, max_levels - 1) if max_levels > 1 else 0
    es = _stencil(dim, ids) if fine_stencil is None else fine_stencil
    X = np.asarray(X, np.float64).copy()
    if n_coarse == 0:
        per = nu_pre + nu_post
        return O.jacobi(dim, kind, 0.0, omega, n_cycles * per, X, levels[0].patches, es, report_every=per, prec=prec)
    levels = levels[: n_coarse + 1]
    ext = max(levels[-1].ids.shape)
    nu_c = min(nu_coarse, 2 * ext)
    stencils = [None] + [_stencil(dim, lv.ids) for lv in levels[1:]]

    def smooth(lv, x, n, tau, es_):
        if n <= 0:
            return x
        x2, _, _ = O.jacobi(dim, kind, 0.0, omega, n, x, lv.patches, es_, prec=prec, tau=tau)
        return np.asarray(x2, np.float64)

    def coarse(li, x, tau):
        lv = levels[li]
        if li + 1 == len(levels):
            return smooth(lv, x, nu_c, tau, stencils[li])
        nx = levels[li + 1]
        x = smooth(lv, x, nu_pre, tau, stencils[li])
        g = O.patch_grads(dim, kind, x, lv.patches, tau=tau, prec=prec)
        xc = _inject(lv, nx, x)
        tau_r = _full_weight(lv, nx, g)
        gc = O.patch_grads(dim, kind, xc, nx.patches, prec=prec)
        xs = coarse(li + 1, xc.copy(), gc - tau_r)
        corr = _prolong(lv, nx, xs - xc)
        x = x + np.where(lv.free[:, None], corr, 0.0)
        return smooth(lv, x, nu_post, tau, stencils[li])

    E, M = [], []
    fine = levels[0]
    for _ in range(n_cycles):
        X = smooth(fine, X, nu_pre, None, es)
        g = O.patch_grads(dim, kind, X, fine.patches, prec=prec)
        c1 = levels[1]
        xc = _inject(fine, c1, X)
        tau_r = _full_weight(fine, c1, g)
        gc = O.patch_grads(dim, kind, xc, c1.patches, prec=prec)
        xs = coarse(1, xc.copy(), gc - tau_r)
        corr = np.where(fine.free[:, None], _prolong(fine, c1, xs - xc), 0.0)

        def en(a):
            return O.energy_mindet(dim, kind, 0.0, es, X + a * corr, prec=prec)

        def ok(e_pre, e, m):
            return np.isfinite(e) and e <= e_pre + tol and m > 0

        e_pre = en(0.0)[0]
        alpha = 0.0
        if ok(e_pre, *en(1.0)):
            alpha = 1.0
        a = 0.5
        it = 1
        while alpha == 0.0 and it < 10:
            if ok(e_pre, *en(a)):
                alpha = a
            elif it + 1 < 10 and ok(e_pre, *en(a * 0.5)):
                alpha = a * 0.5
            it += 2
            a *= 0.25
        X = X + alpha * corr
        X = smooth(fine, X, nu_post, None, es)
        e, m = O.energy_mindet(dim, kind, 0.0, es, X, prec=prec)
        E.append(e)
        M.append(m)
    return X, np.asarray(E), np.asarray(M)
