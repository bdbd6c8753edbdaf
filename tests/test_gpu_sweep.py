"""HIP sweep parity on the MI355X through the C ABI (pybind cpp_core ->
libeggmi.so), checked against the CPU oracle on the same inputs.

Bar (FP32 device arithmetic, FP64 reductions): node positions within 1e-5
of the domain extent of the FP64 oracle on valid meshes; reported energies
and min det equal the FP64 reduction of the returned positions to 1e-9
relative; halo copies and index maps bit-exact. Energies along the
trajectory are compared to ENERGY_RTOL: the U(+-0.3h) starts hold cells with
det A ~ 1e-3 h^2, whose mu ~ 1/det turns the FP32 position resolution
(~1e-7 of the extent) into ~1e-3 relative energy differences on the first
sweeps and ~1e-4 after a few dozen.
"""

from __future__ import annotations

import numpy as np
import pytest

from tests.cases import Case, cube, lr_pair, ogrid, single, star

pytestmark = pytest.mark.gpu

POS_TOL = 1e-5  # x domain extent
ENERGY_RTOL = 2e-4  # final report
TRAJ_RTOL = 2e-3  # every report


def _session(case, **kw):
    from egg_amd.smoothing import CppStructuredSweepSession

    return CppStructuredSweepSession(case.ctx, case.bsc, case.X, **kw)


def _extent(case):
    X = case.X
    return float((X.max(0) - X.min(0)).max())


def _run_both(oracle, case, n, *, phase="barrier", delta=0.0, omega=0.8, report_every=1, prec="f64",
              variant=None):
    s = _session(case)
    if variant is not None:
        s.core.set_tile_variant(variant)
    e, m = s.run(n, phase=phase, delta=delta, omega=omega, report_every=report_every)
    Xg = s.get_X()
    Xo, eo, mo = oracle.jacobi(case.d, phase, delta, omega, n, case.X, case.patches(oracle),
                               case.ctx.energy_stencil, report_every=report_every, prec=prec)
    return s, Xg, np.asarray(e), np.asarray(m), Xo, eo, mo


@pytest.mark.parametrize("n,sweeps", [(33, 40), (129, 40)])
def test_single_block_parity(oracle, n, sweeps):
    """Config-1 shape (one block, U(+-0.3h) interior, omega 0.8): positions
    per the bar against the FP64 oracle, energies per report."""
    c = single(n=n, seed=1)
    _, Xg, e, m, Xo, eo, mo = _run_both(oracle, c, sweeps)
    err = np.abs(Xg - Xo).max() / _extent(c)
    print(f"n={n} max|dX|/extent={err:.2e} E_gpu={e[-1]:.9g} E_f64={eo[-1]:.9g}")
    assert err < POS_TOL
    np.testing.assert_allclose(e, eo, rtol=TRAJ_RTOL)
    np.testing.assert_allclose(e[-1], eo[-1], rtol=ENERGY_RTOL)
    np.testing.assert_allclose(m[-1], mo[-1], rtol=1e-3)


def test_single_block_parity_fp32_oracle(oracle):
    """Same algorithm, same precision: the FP32 oracle agrees even closer."""
    c = single(n=33, seed=2)
    _, Xg, e, m, Xo, eo, mo = _run_both(oracle, c, 40, prec="f32")
    assert np.abs(Xg - Xo).max() / _extent(c) < POS_TOL
    np.testing.assert_allclose(e[-1], eo[-1], rtol=ENERGY_RTOL)


def test_reports_equal_fp64_reduction_of_returned_x(oracle):
    """Reported (E, min det) are the FP64 reduction of the returned FP32 state."""
    c = ogrid(33, 17)
    s = _session(c)
    e, m = s.run(7, omega=0.8, report_every=3)
    assert len(e) == 3  # sweeps 3, 6 and the last one
    X = s.get_X().astype(np.float32)
    E, M = oracle.energy_mindet(2, "barrier", 0.0, c.ctx.energy_stencil, X, prec="f32")
    assert e[-1] == pytest.approx(E, rel=1e-9) and m[-1] == pytest.approx(M, rel=1e-9)


@pytest.mark.parametrize("make", [lambda: ogrid(33, 17), lambda: lr_pair(17)])
def test_multiblock_parity(oracle, make):
    """Multi-block (halo ghosts, owner broadcast, stored-stencil interface and
    valence-3 DOFs) against the FP64 oracle in global id space."""
    c = make()
    _, Xg, e, m, Xo, eo, mo = _run_both(oracle, c, 30)
    err = np.abs(Xg - Xo).max() / _extent(c)
    print(f"multi-block max|dX|/extent={err:.2e}")
    assert err < POS_TOL
    np.testing.assert_allclose(e, eo, rtol=TRAJ_RTOL)
    np.testing.assert_allclose(e[-1], eo[-1], rtol=ENERGY_RTOL)


@pytest.mark.parametrize("small", [False, True])
@pytest.mark.parametrize("kw,phase", [(dict(k=5, n=9), "barrier"), (dict(k=6, n=9), "barrier"),
                                      (dict(k=5, n=9), "untangle"), (dict(k=5, n=7, layers=6), "barrier")])
def test_many_occurrence_stored_dofs(oracle, kw, phase, small):
    """Singular DOFs whose patches take more than one round of their lane
    group: the centre of a valence-5 / 6 star (20 / 24 occurrences for 16
    lanes) and the centre line of the extruded star (80 occurrences for 64
    lanes), against the FP64 oracle; `small`: with the tile variant of the
    fewest lanes per stored DOF (2D: 4, 3D: 16), so every stored DOF takes
    several rounds."""
    c = star(**kw)
    info = _session(c).core.info()
    assert info["n_stored"] > 0
    delta = 1e-3 if phase == "untangle" else 0.0
    variant = (12 if c.d == 2 else 4) if small else None
    _, Xg, e, m, Xo, eo, mo = _run_both(oracle, c, 20, phase=phase, delta=delta, variant=variant)
    err = np.abs(Xg - Xo).max() / _extent(c)
    print(f"star {kw} {phase}: max|dX|/extent={err:.2e}")
    assert err < POS_TOL
    np.testing.assert_allclose(e, eo, rtol=TRAJ_RTOL)
    np.testing.assert_allclose(e[-1], eo[-1], rtol=ENERGY_RTOL)
    np.testing.assert_allclose(m, mo, rtol=1e-3)


def test_stored_path_and_sample_reduction(oracle):
    """A non-isotropic uniform target ships every DOF with a stored patch and
    forces the explicit-sample energy reduction."""
    from egg_amd import grids

    W = np.array([[1.0, 0.15], [0.0, 0.9]])
    c = Case(grids.single_block_2d(17, seed=5, amp=0.2), w_inv=W)
    s = _session(c)
    info = s.core.info()
    assert info["n_synth"] == 0 and info["n_stored"] == c.grid.n_free and info["structured_energy"] == 0
    e, m = s.run(20, omega=0.8, report_every=1)
    Xo, eo, mo = oracle.jacobi(2, "barrier", 0.0, 0.8, 20, c.X, c.patches(oracle), c.ctx.energy_stencil,
                               report_every=1)
    assert np.abs(s.get_X() - Xo).max() / _extent(c) < POS_TOL
    np.testing.assert_allclose(e, eo, rtol=TRAJ_RTOL)


def test_halo_pair_known_answer():
    """The reference's device exchange test (tests/cpp/test_halo_device.cpp:
    82-136) on the MI355X: after one exchange every face ghost holds the
    extrapolated coordinate exactly, and the shared face nodes are unchanged."""
    from tests.cases import halo_pair, halo_pair_expected

    c = halo_pair()
    s = _session(c)
    core = s.core
    core.halo_exchange()
    P = core.packed()
    base = core.tables()["block_base"]
    stride = 3 + 2
    for (b, (p0, p1)), want in halo_pair_expected().items():
        slot = int(base[b]) + p0 * stride + p1
        assert (float(P[0, slot]), float(P[1, slot])) == want, (b, p0, p1)
    for b, i in ((0, 3), (1, 0)):
        for j in range(3):
            assert float(P[0, int(base[b]) + (i + 1) * stride + (j + 1)]) == 3.0


def test_halo_exchange_bit_exact():
    """After the fused copy every ghost / copy / mirror slot holds its source's
    coordinates bit for bit, and every interior slot its node's FP32 value."""
    c = ogrid(33, 17)
    s = _session(c)
    core = s.core
    core.halo_exchange()
    P = core.packed()
    tab = core.tables()
    src, dst = tab["copy_src"], tab["copy_dst"]
    assert np.array_equal(P[:, dst].view(np.uint32), P[:, src].view(np.uint32))
    g2s = tab["g2s"]
    np.testing.assert_array_equal(P[:, g2s].T, c.X.astype(np.float32))
    # after sweeps the invariant still holds (the copy runs after every sweep)
    s.run(5, omega=0.8)
    P = core.packed()
    assert np.array_equal(P[:, dst].view(np.uint32), P[:, src].view(np.uint32))


def test_tile_variants_agree(oracle):
    """Every tile shape of one lane layout computes the same nodes the same
    way (bitwise); the one-lane-per-cell (0-3, and 6-7 without LDS staging)
    layouts are bitwise equal, and so are the two-cells-per-lane ones (8, 10,
    11); the several-cells-per-lane kernels (two, 9: four) fold a patch in
    the one-lane-per-cell order but fuse multiply-adds
    differently, and the one-lane-per-sample (4-5) layout sums it in another
    order; 12 and 13 are variant 0 with 4 / 8 lanes per stored DOF (its
    occurrences summed in another order); all match the oracle."""
    c = ogrid(33, 17)
    out = []
    nv = _session(c).core.set_tile_variant(-1)
    assert nv == 14
    for v in range(nv):
        s = _session(c)
        s.core.set_tile_variant(v)
        s.run(5, omega=0.8)
        out.append(s.get_X())
    for v in (1, 2, 3, 6, 7):  # one lane per cell, with or without LDS staging
        np.testing.assert_array_equal(out[v], out[0])
    np.testing.assert_array_equal(out[5], out[4])
    Xo, _, _ = oracle.jacobi(2, "barrier", 0.0, 0.8, 5, c.X, c.patches(oracle), c.ctx.energy_stencil)
    for v in (10, 11):
        np.testing.assert_array_equal(out[v], out[8])
    for X in (out[0], out[4], out[8], out[9], out[12], out[13]):
        assert np.abs(X - Xo).max() / _extent(c) < POS_TOL


@pytest.mark.parametrize("v", [8, 9, 10])
def test_cells_per_lane_phases_and_fas(oracle, v):
    """The several-cells-per-lane kernels in the untangle and shape_size
    phases against the FP64 oracle, and in FAS (gradient-only and
    tau-shifted sweeps) against the one-lane-per-cell kernels."""
    c = ogrid(33, 17, tangle_seed=2)
    s = _session(c)
    s.core.set_tile_variant(v)
    e1, _ = s.run(10, phase="untangle", delta=0.01, omega=0.5, report_every=5)
    Xo, eo, _ = oracle.jacobi(2, "untangle", 0.01, 0.5, 10, c.X, c.patches(oracle), c.ctx.energy_stencil,
                              report_every=5)
    assert np.abs(s.get_X() - Xo).max() / _extent(c) < POS_TOL
    np.testing.assert_allclose(e1, eo, rtol=TRAJ_RTOL)
    c2 = ogrid(33, 17)
    s = _session(c2)
    s.core.set_tile_variant(v)
    e2, _ = s.run(5, phase="shape_size", omega=0.8, report_every=5)
    Xo, eo, _ = oracle.jacobi(2, "shape_size", 0.0, 0.8, 5, c2.X, c2.patches(oracle), c2.ctx.energy_stencil,
                              report_every=5)
    assert np.abs(s.get_X() - Xo).max() / _extent(c2) < POS_TOL
    np.testing.assert_allclose(e2, eo, rtol=TRAJ_RTOL)
    fas = []
    for var in (0, v):
        s = _session(c2)
        s.core.set_tile_variant(var)
        e3, m3 = s.run_fas(3, nu_pre=2, nu_post=2, omega=0.8, max_levels=3)
        fas.append((s.get_X(), e3, m3))
    assert np.abs(fas[0][0] - fas[1][0]).max() / _extent(c2) < POS_TOL
    np.testing.assert_allclose(fas[1][1], fas[0][1], rtol=1e-5)
    assert np.all(fas[1][2] > 0)


def test_sample_variant_fas_and_untangle(oracle):
    """The one-lane-per-sample kernels also run the untangle phase and the
    FAS gradient / tau-shifted sweeps."""
    c = ogrid(33, 17, tangle_seed=2)
    s = _session(c)
    s.core.set_tile_variant(4)
    delta = 0.01
    e, m = s.run(10, phase="untangle", delta=delta, omega=0.5, report_every=5)
    Xo, eo, _ = oracle.jacobi(2, "untangle", delta, 0.5, 10, c.X, c.patches(oracle), c.ctx.energy_stencil,
                              report_every=5)
    assert np.abs(s.get_X() - Xo).max() / _extent(c) < POS_TOL
    np.testing.assert_allclose(e, eo, rtol=TRAJ_RTOL)
    u = ogrid(33, 17)
    a = _session(u)
    b = _session(u)
    b.core.set_tile_variant(5)
    ea, _ = a.run_fas(4, nu_pre=2, nu_post=2, omega=0.8)
    eb, _ = b.run_fas(4, nu_pre=2, nu_post=2, omega=0.8)
    np.testing.assert_allclose(eb, ea, rtol=1e-5)


def test_analytic_minimiser():
    """Perturbed Cartesian block -> the uniform grid (reference
    tests/smoothing/test_structured_sweep.py:159-174), at FP32 resolution."""
    c = single(n=9, seed=11, amp=0.25)
    uniform = np.stack(np.meshgrid(np.arange(9), np.arange(9), indexing="ij"), -1).reshape(-1, 2) / 8.0
    s = _session(c)
    _, m = s.run(3000, omega=1.0)
    assert m[-1] > 0
    np.testing.assert_allclose(s.get_X(), uniform, atol=2e-5)


def test_untangle_ogrid(oracle):
    """delta-continuation untangle clears a 2%-tangled O-grid; the reported
    min det is the FP64 reduction of the returned state."""
    from egg_amd.smoothing import cpp_untangle

    c = ogrid(33, 17, tangle_seed=4)
    md0 = oracle.energy_mindet(2, "barrier", 0.0, c.ctx.energy_stencil, c.X)[1]
    assert md0 < 0
    X, md, outer, delta = cpp_untangle(c.ctx, c.bsc, c.X)
    assert md > 1e-9 and outer >= 1
    E, M = oracle.energy_mindet(2, "barrier", 0.0, c.ctx.energy_stencil, X.astype(np.float32), prec="f32")
    assert M == pytest.approx(md, rel=1e-9)


def test_untangle_sweep_parity(oracle):
    """Untangle-phase sweeps on a mildly folded block track the FP64 oracle."""
    from egg_amd import grids

    g = grids.single_block_2d(17, seed=6, amp=0.2)
    g = grids.tangle(g, frac=0.02, factor=1.2, seed=2)
    c = Case(g)
    md = grids.min_det(g)
    _, Xg, e, m, Xo, eo, mo = _run_both(oracle, c, 10, phase="untangle", delta=2 * abs(md), omega=0.5)
    assert np.abs(Xg - Xo).max() / _extent(c) < 1e-4
    np.testing.assert_allclose(e, eo, rtol=1e-4)


def test_shape_size_parity(oracle):
    c = single(n=17, seed=8, amp=0.2)
    _, Xg, e, m, Xo, eo, mo = _run_both(oracle, c, 10, phase="shape_size")
    assert np.abs(Xg - Xo).max() / _extent(c) < POS_TOL
    np.testing.assert_allclose(e, eo, rtol=TRAJ_RTOL)


def test_3d_block_parity(oracle):
    """3D hex block (64-occurrence patches, mu_cond3) against the FP64 oracle."""
    c = cube(n=9, seed=1)
    _, Xg, e, m, Xo, eo, mo = _run_both(oracle, c, 10)
    err = np.abs(Xg - Xo).max() / _extent(c)
    print(f"3d max|dX|/extent={err:.2e}")
    assert err < POS_TOL
    np.testing.assert_allclose(e, eo, rtol=TRAJ_RTOL)


@pytest.mark.parametrize("phase,delta", [("shape_size", 0.0), ("untangle", 0.05)])
def test_3d_phases_parity(oracle, phase, delta):
    """3D shape_size (cond3 + size term) and untangle (delta surrogate)
    sweeps against the FP64 oracle, including stored DOFs of the cubed
    sphere for the untangle phase."""
    from egg_amd import grids

    c = cube(n=9, seed=3) if phase == "shape_size" else Case(grids.tangle(grids.cubed_sphere_3d(9), frac=0.01,
                                                                           factor=1.5, seed=5))
    _, Xg, e, m, Xo, eo, mo = _run_both(oracle, c, 6, phase=phase, delta=delta,
                                        omega=0.8 if phase == "shape_size" else 0.5)
    assert np.abs(Xg - Xo).max() / _extent(c) < POS_TOL
    np.testing.assert_allclose(e, eo, rtol=TRAJ_RTOL)


def test_set_get_x_roundtrip():
    c = ogrid(17, 9)
    s = _session(c)
    X2 = c.X + 1e-3
    s.set_X(X2)
    np.testing.assert_allclose(s.get_X(), X2.astype(np.float32), rtol=0, atol=0)


def test_3d_tile_variants_agree():
    """3D tile shapes (8x4x2, 4x4x4, 8x4x2 without LDS staging) are bitwise
    identical on the cubed sphere (stored DOFs included); two cells per lane
    (variant 3) folds in the same order with other multiply-add fusion, and
    variant 4 (16 lanes per stored DOF) sums stored patches in another
    order."""
    from egg_amd import grids

    c = Case(grids.cubed_sphere_3d(9))
    nv = _session(c).core.set_tile_variant(-1)
    assert nv == 5
    out = []
    for v in range(nv):
        s = _session(c)
        s.core.set_tile_variant(v)
        s.run(3, omega=0.8)
        out.append(s.get_X())
    for X in out[1:3]:
        np.testing.assert_array_equal(X, out[0])
    for X in out[3:]:
        assert np.abs(X - out[0]).max() / _extent(c) < POS_TOL


def test_graph_chunks_match_launches():
    """Sweeps replayed as captured hipGraph chunks give bitwise the same
    states and reports as single launches, for report cadences that do and
    do not divide into chunks, and after a tile-variant change (which
    re-captures)."""
    c = ogrid(33, 17)
    a = _session(c)
    b = _session(c)
    b.core.set_graph_chunk(8)
    for n, rep in ((50, 7), (64, 0), (13, 5), (40, 16)):
        ea, ma = a.run(n, omega=0.8, report_every=rep)
        eb, mb = b.run(n, omega=0.8, report_every=rep)
        np.testing.assert_array_equal(ea, eb)
        np.testing.assert_array_equal(ma, mb)
        np.testing.assert_array_equal(a.get_X(), b.get_X())
    a.core.set_tile_variant(2)
    b.core.set_tile_variant(2)
    ea, _ = a.run(33, omega=0.8, report_every=11)
    eb, _ = b.run(33, omega=0.8, report_every=11)
    np.testing.assert_array_equal(ea, eb)
    np.testing.assert_array_equal(a.get_X(), b.get_X())
    with pytest.raises(ValueError):
        b.core.set_graph_chunk(7)


@pytest.mark.gpu
def test_raw_ctypes_session_flow(oracle):
    """INTEGRATION's C-ABI flow through plain ctypes: eggmi_session_create ->
    eggmi_run -> eggmi_get_X -> eggmi_session_destroy gives the pybind
    session's bits and the oracle's positions on the O-grid."""
    from egg_amd.smoothing import CppStructuredSweepSession, structured_arrays
    from tests import capi_ctypes as CA
    from tests.cases import ogrid

    c = ogrid(33, 17, tangle_seed=None)
    print("ctypes flow: O-grid 33/17, 40 sweeps", flush=True)
    lib = CA.load()
    s = CA.Session(lib, c.ctx.wire, structured_arrays(c.bsc), c.X, 2, device=1)
    E, M = s.run(40, omega=0.8, report_every=10)
    Xc = s.get_X()
    s.destroy()
    p = CppStructuredSweepSession(c.ctx, c.bsc, c.X, device="gpu")
    e, m = p.run(40, omega=0.8, report_every=10)
    np.testing.assert_array_equal(Xc, p.get_X())
    np.testing.assert_array_equal(E, e)
    np.testing.assert_array_equal(M, m)
    Xo, eo, mo = oracle.jacobi(2, "barrier", 0.0, 0.8, 40, c.X, c.patches(oracle), c.ctx.energy_stencil,
                               report_every=10)
    assert np.abs(Xc - Xo).max() / 4.0 < 1e-5
    assert abs(E[-1] - eo[-1]) / abs(eo[-1]) < 1e-5
