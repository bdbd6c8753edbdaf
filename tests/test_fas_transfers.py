"""The FAS transfer operators pinned by the identities the reference's own
tests hold them to (tests/cpp/test_multigrid_device.cpp:76-220, 389-470):
injection copies the parent node bitwise; full weighting of a field linear
in the logical index returns the centre value; prolongation of a linear
coarse field is linear in the fractional coarse index; injection after
prolongation gives the coarse field back; and (device) the tau-shifted
coarse gradient at the injected state equals the restricted fine gradient.
The CPU tests hold oracle/fas.py to them, the GPU tests the HIP kernels
(through eggmi_mg_probe) on a single block and on the O-grid's multi-block
levels with their interface DOFs."""

from __future__ import annotations

import numpy as np
import pytest

from egg_amd import grids
from tests.cases import Case, fas_single


def _oracle_levels(n=9, d=2):
    from oracle import fas

    ids = np.arange(n**d).reshape((n,) * d)
    free = np.zeros(ids.size, bool)
    free[ids[(slice(1, -1),) * d].reshape(-1)] = True
    return fas.build_levels(d, ids, free, max_levels=2)


@pytest.mark.parametrize("d", [2, 3])
def test_oracle_transfer_identities(d):
    from oracle import fas

    p, c = _oracle_levels(9, d)
    rng = np.random.default_rng(d)
    xp = rng.normal(size=(p.ids.size, d))
    xc = fas._inject(p, c, xp)
    for I in np.ndindex(*c.ids.shape):
        assert np.array_equal(xc[c.ids[I]], xp[p.ids[tuple(2 * np.asarray(I))]])
    # index-linear field: full weighting returns the centre value at free coarse nodes, 0 elsewhere
    coef = np.array([3.0, -2.0, 0.5])[:d]
    lin = np.zeros((p.ids.size, d))
    for J in np.ndindex(*p.ids.shape):
        lin[p.ids[J]] = [coef @ np.asarray(J) + k for k in range(d)]
    fw = fas._full_weight(p, c, lin)
    for I in np.ndindex(*c.ids.shape):
        want = lin[p.ids[tuple(2 * np.asarray(I))]] if c.free[c.ids[I]] else np.zeros(d)
        np.testing.assert_allclose(fw[c.ids[I]], want, rtol=0, atol=1e-12)
    # prolongation of a linear coarse field is linear in the fractional index
    cc = np.zeros((c.ids.size, d))
    for I in np.ndindex(*c.ids.shape):
        cc[c.ids[I]] = [np.array([1.5, 0.5, -1.0])[:d] @ np.asarray(I) + k for k in range(d)]
    pr = fas._prolong(p, c, cc)
    for J in np.ndindex(*p.ids.shape):
        want = [np.array([1.5, 0.5, -1.0])[:d] @ (np.asarray(J) / 2.0) + k for k in range(d)]
        np.testing.assert_allclose(pr[p.ids[J]], want, rtol=0, atol=1e-12)
    # injection after prolongation returns the coarse field bitwise
    pat = rng.normal(size=(c.ids.size, d))
    assert np.array_equal(fas._inject(p, c, fas._prolong(p, c, pat)), pat)


def test_oracle_full_weighting_catches_a_wrong_weight():
    """A full weighting with (1/3, 1/3, 1/3) weights still preserves
    constants but not the linear field: the identity above is what pins the
    weights."""
    from oracle import fas

    p, c = _oracle_levels(9, 2)
    lin = np.zeros((p.ids.size, 2))
    for J in np.ndindex(*p.ids.shape):
        lin[p.ids[J]] = [J[0] ** 2, J[1]]
    fw = fas._full_weight(p, c, lin)
    I = (2, 2)
    # the (1/4, 1/2, 1/4) weights of a quadratic in i give i^2 + 1/2 at the centre, not i^2 + 2/3
    assert fw[c.ids[I]][0] == pytest.approx(16 + 0.5)


# ---------------------------------------------------------------- device

def _layout(shapes):
    """slot -> (block, padded index) of a level's padded store."""
    loc, base = [], 0
    for b, shp in enumerate(shapes):
        pad = [s + 2 for s in shp]
        for p in np.ndindex(*pad):
            loc.append((b, p))
        base += int(np.prod(pad))
    return loc


def _fields(session, level):
    """(parent slots, child slots) as (block, padded) lists plus the shapes."""
    core = session.core
    fine_shapes = [tuple(np.shape(b)) for b in session._grid_blocks]
    shapes = [[tuple(s) for s in lv] for lv in core.mg_levels()]
    ps = fine_shapes if level == 1 else shapes[level - 2]
    cs = shapes[level - 1]
    return _layout(ps), _layout(cs), ps, cs


def _session(case):
    from egg_amd.smoothing import CppStructuredSweepSession

    s = CppStructuredSweepSession(case.ctx, case.bsc, case.X, device="gpu")
    s._grid_blocks = case.grid.blocks
    return s


def _cases():
    return {"single": Case(fas_single().grid), "ogrid": Case(grids.ogrid_2d(33, 17)),
            "cube": Case(grids.single_block_3d(9, seed=1, amp=0.1))}


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["single", "ogrid", "cube"])
def test_device_transfer_identities(name):
    c = _cases()[name]
    s = _session(c)
    core = s.core
    d = c.d
    n_levels = len(core.mg_levels())
    assert n_levels >= 1
    for level in range(1, min(n_levels, 2) + 1):
        print(f"transfer identities: {name} level {level}", flush=True)
        ploc, cloc, ps, cs = _fields(s, level)
        tab = core.mg_level_tables(level)
        npar = int(core.mg_level_tables(level - 1)["n_slots"])
        nch = int(tab["n_slots"])
        cfree = np.asarray(tab["free"], bool)
        pfree = np.asarray(core.mg_level_tables(level - 1)["free"], bool)
        fac = [[2 if cn != pn else 1 for cn, pn in zip(cshp, pshp)] for cshp, pshp in zip(cs, ps)]
        rng = np.random.default_rng(level)
        # injection: bitwise at every interior child node
        f = rng.normal(size=(d, npar)).astype(np.float32).astype(np.float64)
        out = core.mg_probe(level, 0, f.reshape(-1)).reshape(d, nch)
        pindex = {bp: i for i, bp in enumerate(ploc)}
        for q, (b, p) in enumerate(cloc):
            lg = np.asarray(p) - 1
            if np.any(lg < 0) or np.any(lg >= np.asarray(cs[b])):
                continue
            src = pindex[(b, tuple(int(v) for v in lg * np.asarray(fac[b]) + 1))]
            assert np.array_equal(out[:, q].astype(np.float32), f[:, src].astype(np.float32))
        # full weighting of a field linear in the (padded) index: the centre value at relaxed DOFs
        coef = np.array([3.0, -2.0, 0.5])[:d]
        lin = np.zeros((d, npar))
        for i, (b, p) in enumerate(ploc):
            lin[:, i] = [coef @ (np.asarray(p) - 1) + k for k in range(d)]
        fw = core.mg_probe(level, 1, lin.reshape(-1)).reshape(d, nch)
        prolong_in = np.zeros((d, nch))
        for q, (b, p) in enumerate(cloc):
            lg = np.asarray(p) - 1
            if np.any(lg < 0) or np.any(lg >= np.asarray(cs[b])):
                continue
            centre = [coef @ (lg * np.asarray(fac[b])) + k for k in range(d)]
            want = centre if cfree[q] else np.zeros(d)
            np.testing.assert_allclose(fw[:, q], want, rtol=0, atol=2e-5 * (1 + np.abs(centre).max()))
            prolong_in[:, q] = [np.array([1.5, 0.5, -1.0])[:d] @ lg + k for k in range(d)]
        # prolongation of a linear child field: linear at relaxed parent nodes, zero at the other interior ones
        pr = core.mg_probe(level, 2, prolong_in.reshape(-1)).reshape(d, npar)
        for i, (b, p) in enumerate(ploc):
            lg = np.asarray(p) - 1
            if np.any(lg < 0) or np.any(lg >= np.asarray(ps[b])):
                continue
            frac = lg / np.asarray(fac[b], float)
            want = [np.array([1.5, 0.5, -1.0])[:d] @ frac + k for k in range(d)] if pfree[i] else np.zeros(d)
            np.testing.assert_allclose(pr[:, i], want, rtol=0, atol=2e-5 * (1 + np.abs(want).max()))
        # injection after prolongation: the child field back where the parent image relaxes
        back = core.mg_probe(level, 0, pr.reshape(-1)).reshape(d, nch)
        for q, (b, p) in enumerate(cloc):
            lg = np.asarray(p) - 1
            if np.any(lg < 0) or np.any(lg >= np.asarray(cs[b])):
                continue
            src = pindex[(b, tuple(int(v) for v in lg * np.asarray(fac[b]) + 1))]
            if pfree[src]:
                assert np.array_equal(back[:, q].astype(np.float32), prolong_in[:, q].astype(np.float32))


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["single", "ogrid", "cube"])
def test_device_tau_coherence(name):
    """grad(E_c - tau . x)(Rx) = R grad E_f(x) at every relaxed coarse DOF,
    for the tau the restriction forms (the sign and placement of the
    shift; reference tests/cpp/test_multigrid_device.cpp:389-470)."""
    c = _cases()[name]
    s = _session(c)
    s.run(3, omega=0.8)  # a state off the initial grid
    core = s.core
    tab = core.mg_level_tables(1)
    n = int(tab["n_slots"])
    out = core.mg_probe(1, 3, np.zeros(0)).reshape(2, c.d, n)
    free = np.asarray(tab["free"], bool)
    shifted, rg = out[0][:, free], out[1][:, free]
    scale = np.abs(rg).max()
    assert scale > 0
    np.testing.assert_allclose(shifted, rg, rtol=0, atol=1e-4 * scale)
