"""Host-side logic without a GPU: the C ABI library, wire builders, the
structured remap's index maps and validation errors."""

from __future__ import annotations

import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

from tests.cases import halo_pair, halo_pair_expected, lr_pair, ogrid, single, star

ROOT = Path(__file__).resolve().parent.parent


def _core():
    from egg_amd._cpp import cpp_core

    return cpp_core


def test_capi_exports_every_declared_symbol():
    """libeggmi.so loads and exports every function include/eggmi.h declares."""
    lib = ctypes.CDLL(str(ROOT / "egg_amd" / "_lib" / "libeggmi.so"))
    decl = (ROOT / "include" / "eggmi.h").read_text()
    names = set(re.findall(r"\b(eggmi_[A-Za-z_]+)\s*\(", decl))
    assert {"eggmi_session_create", "eggmi_run", "eggmi_run_fas", "eggmi_get_X"} <= names
    for n in sorted(names):
        assert hasattr(lib, n), n
    lib.eggmi_ping.restype = ctypes.c_int
    assert lib.eggmi_ping() == 0


def test_flat_context_counts():
    c = single(n=9)
    w = c.ctx.wire
    g = w["groups"][0]
    assert w["energy_stencil"]["num_samples"] == 4 * 8 * 8
    assert g["D"] == 49 and int(np.sum(g["P_of"])) == len(g["gc"])
    assert np.all(g["interior_block"] == 0)  # every free DOF of one block is interior
    o = ogrid(17, 9)
    g = o.ctx.wire["groups"][0]
    stored = np.asarray(g["interior_block"]) < 0
    P = np.asarray(g["P_of"])[stored]
    # interface DOFs see 4 cells (16 samples), the 4 valence-3 corners 3 cells (12)
    assert np.count_nonzero(P == 12) == 4
    assert np.all((P == 12) | (P == 16))


def _tables(case, device="tables"):
    core = _core()
    from egg_amd.smoothing import structured_arrays

    s = core.CppStructuredSweepSession(case.ctx.wire, structured_arrays(case.bsc),
                                       np.ascontiguousarray(case.X).ravel(), device=device, dim=case.d)
    return s.info(), s.tables()


@pytest.mark.parametrize("make", [lambda: single(n=9), lambda: lr_pair(9), lambda: ogrid(17, 9),
                                  lambda: star(k=5, n=7), lambda: star(k=5, n=5, layers=4)])
def test_remap_tables_match_oracle(make):
    """Copy tables (halo | fan | share), g2s and block bases equal the
    restatement of the reference's apply_structured_remap, entry by entry."""
    from oracle.remap import remap

    c = make()
    info, tab = _tables(c)
    b = c.bsc
    ref = remap(c.d, b.interior_shapes, b.block_global_dof, b.halo_src_block, b.halo_src_padded,
                b.halo_dst_block, b.halo_dst_padded, b.sing_block, b.sing_logical,
                groups=c.ctx.wire["groups"], n_nodes=c.grid.n_nodes)
    assert info["n_halo"] == ref["n_halo"] and info["n_fan"] == ref["n_fan"] and info["n_share"] == ref["n_share"]
    assert info["n_overflow"] == ref["n_overflow"] and info["n_slots"] == ref["n_slots"]
    np.testing.assert_array_equal(tab["copy_src"], ref["copy_src"])
    np.testing.assert_array_equal(tab["copy_dst"], ref["copy_dst"])
    np.testing.assert_array_equal(tab["g2s"], ref["g2s"])
    np.testing.assert_array_equal(tab["block_base"], ref["base"])
    assert info["n_stored"] == len(ref["stored"])


def test_halo_pair_known_answer_tables():
    """The reference's two-block exchange (tests/cpp/test_halo_device.cpp:
    82-136): each face ghost is sourced from the neighbour block's first
    interior layer, whose global coordinate is the block's own field
    extrapolated one node past its face."""
    c = halo_pair()
    info, tab = _tables(c)
    n1 = 3
    # the 2 x n1 face ghosts, plus the twin-matching builder's entries for the
    # face corners seen from the side faces (cpp_backend.py:200-236 resolves
    # them through the same twin; no patch reads them)
    assert info["n_halo"] == 2 * n1 + 4
    stride = n1 + 2
    slot2node = {}
    for b, ids in enumerate(c.grid.blocks):
        for (i, j), g in np.ndenumerate(ids):
            slot2node[int(tab["block_base"][b]) + (i + 1) * stride + (j + 1)] = int(g)
    src_of = dict(zip(tab["copy_dst"][: info["n_halo"]].tolist(), tab["copy_src"][: info["n_halo"]].tolist()))
    for (b, (p0, p1)), want in halo_pair_expected().items():
        dst = int(tab["block_base"][b]) + p0 * stride + p1
        assert dst in src_of, (b, p0, p1)
        np.testing.assert_array_equal(c.X[slot2node[src_of[dst]]], want)


def test_ogrid_halo_tables_shape():
    """O-grid: shared face nodes get ghost sources and non-owner copies a
    broadcast; the valence-3 corner patches resolve through face ghosts (the
    radial node behind a corner is the twin of both ring copies), so no
    spare-ghost mirrors are needed."""
    c = ogrid(17, 9)
    info, _ = _tables(c)
    assert info["n_halo"] == c.bsc.num_halo_entries > 0
    # ring copies of the circle (4 x 17) and of the four connector lines (4 x 8)
    assert info["n_share"] == 4 * 17 + 4 * 8
    assert info["n_fan"] == 0 and info["n_overflow"] == 0
    assert info["structured_energy"] == 1


@pytest.mark.parametrize(
    "mutate,msg",
    [
        (lambda g, e: g.__setitem__("dof_idx", np.full_like(g["dof_idx"], 10**7)), "dof_idx index"),
        (lambda g, e: g.__setitem__("role", np.full_like(g["role"], 7)), "role 7 out of range"),
        (lambda g, e: g.__setitem__("P_of", np.asarray(g["P_of"]) + 1), "sum(P_of)"),
        (lambda g, e: e.__setitem__("gc", np.full_like(e["gc"], -1)), "energy_stencil gc index"),
        (lambda g, e: g.__setitem__("s0", np.asarray(g["s0"]) * 0.5), "must be +-1"),
    ],
)
def test_validation_errors(mutate, msg):
    """Malformed wires raise ValueError before any device work
    (reference tests/test_cpp_validation.py:78-143)."""
    import copy

    from egg_amd.smoothing import structured_arrays

    c = ogrid(9, 5)
    w = copy.deepcopy(c.ctx.wire)
    mutate(w["groups"][0], w["energy_stencil"])
    with pytest.raises(ValueError, match=re.escape(msg)):
        _core().CppStructuredSweepSession(w, structured_arrays(c.bsc), c.X.ravel(), device="tables")


def test_x_shape_errors():
    from egg_amd.smoothing import structured_arrays

    c = single(n=5)
    core = _core()
    with pytest.raises(ValueError, match="flat 1-D"):
        core.CppStructuredSweepSession(c.ctx.wire, structured_arrays(c.bsc), c.X, device="tables")
    with pytest.raises(ValueError, match="multiple of d"):
        core.CppStructuredSweepSession(c.ctx.wire, structured_arrays(c.bsc), c.X.ravel()[:-1], device="tables")
    with pytest.raises(ValueError, match="dim must be 2 or 3"):
        core.CppStructuredSweepSession(c.ctx.wire, structured_arrays(c.bsc), c.X.ravel(), device="tables", dim=4)


def test_cpu_device_is_refused():
    """There is no CPU fallback: device='cpu' is a RuntimeError."""
    from egg_amd.smoothing import structured_arrays

    c = single(n=5)
    with pytest.raises(RuntimeError, match="device='cpu'"):
        _core().CppStructuredSweepSession(c.ctx.wire, structured_arrays(c.bsc), c.X.ravel(), device="cpu")


def test_grids_valid():
    from egg_amd import grids

    g = grids.ogrid_2d(33, 17)
    assert grids.min_det(g) > 0
    t = grids.tangle(g, seed=1)
    assert grids.min_det(t) < 0
    assert np.array_equal(t.free, g.free)


def test_mg_hierarchy_shapes_match_oracle():
    """Coarse shapes (halve odd axes >= 5) of the core match the FAS oracle."""
    from oracle import fas

    from egg_amd.smoothing import structured_arrays

    for n in (9, 17, 33):
        c = single(n=n)
        s = _core().CppStructuredSweepSession(c.ctx.wire, structured_arrays(c.bsc), c.X.ravel(), device="tables")
        got = [lv[0] for lv in s.mg_levels()]
        ref = [lv.ids.shape for lv in fas.build_levels(2, c.grid.blocks[0], c.grid.free)[1:]]
        assert got == ref, (n, got, ref)


def test_capi_error_codes_without_device():
    """The C ABI reports invalid arguments before touching a device, null
    handles as EGGMI_INVALID_ARGUMENT with a message, and the learned
    proposal's constructor checks its dimension first."""
    lib = ctypes.CDLL(str(ROOT / "egg_amd" / "_lib" / "libeggmi.so"))
    lib.eggmi_last_error.restype = ctypes.c_char_p
    out = ctypes.c_void_p()
    assert lib.eggmi_mlp_create(4, ctypes.c_uint64(0), ctypes.c_int64(0), ctypes.byref(out)) == 1
    assert b"dim" in lib.eggmi_last_error()
    assert lib.eggmi_set_graph_chunk(None, 8) == 1
    assert b"session" in lib.eggmi_last_error()
    assert lib.eggmi_autotune(None, 5, ctypes.c_double(0.8), None, 0, None) == 1
    assert lib.eggmi_energy(None, 0, ctypes.c_double(0.0), None, None) == 1
    assert lib.eggmi_set_level_variants(None, None, 0) == 1
    n = ctypes.c_int()
    assert lib.eggmi_autotune_levels(None, 4, ctypes.c_double(0.8), 3, 0, None, ctypes.c_int64(0), None, 0,
                                     ctypes.byref(n)) == 1
    assert b"session" in lib.eggmi_last_error()
    assert lib.eggmi_session_set_proposal_every(None, 2) == 1
    assert lib.eggmi_session_set_control(None, None) == 1
    assert lib.eggmi_control_get_C(None, None, ctypes.c_int64(0)) == 1
    assert lib.eggmi_run_control(None, None, 0, None, None, None, ctypes.c_int64(0), None, None, None,
                                 ctypes.c_int64(0), None, None, None) == 1
    k = ctypes.c_int64()
    assert lib.eggmi_fas_cycle_kernels(None, ctypes.byref(k)) == 1
    assert b"session" in lib.eggmi_last_error()
    assert lib.eggmi_mlp_destroy(None) == 0  # deleting nothing is fine
    assert lib.eggmi_session_destroy(None) == 0


def test_level_variants_validated_without_device():
    """Per-level tile variants are range-checked (-1 = the session's; V + v
    = one-workgroup smoothing on variant v's tiles)."""
    from egg_amd.smoothing import structured_arrays

    c = single(n=17)
    s = _core().CppStructuredSweepSession(c.ctx.wire, structured_arrays(c.bsc), c.X.ravel(), device="tables")
    nv = s.set_tile_variant(-1)
    s.set_level_variants([-1, nv - 1, 0, nv, 2 * nv - 1])
    s.set_level_variants([])
    for bad in ([2 * nv], [-2], [0, 2 * nv + 3]):
        with pytest.raises(ValueError, match="level tile variant"):
            s.set_level_variants(bad)


def test_weighted_wire_validated_without_device():
    """A weighted wire is accepted with every DOF stored and the explicit
    sample reduction; a weight array of the wrong size or with a non-finite
    entry is refused."""
    from egg_amd import grids
    from egg_amd.smoothing import structured_arrays
    from tests.cases import Case

    g = grids.single_block_2d(9, seed=2, amp=0.2)
    ns = Case(g).ctx.energy_stencil["num_samples"]
    c = Case(g, weight=np.linspace(0.5, 1.5, ns))

    def make(wire):
        return _core().CppStructuredSweepSession(wire, structured_arrays(c.bsc), c.X.ravel(), device="tables")

    info = make(c.ctx.wire).info()
    assert info["n_synth"] == 0 and info["n_stored"] == g.n_free and info["structured_energy"] == 0
    grp = dict(c.ctx.wire["groups"][0])
    for bad, msg in ((grp["weight"][:-1], "weight size"), (np.where(np.arange(len(grp["weight"])) == 3, np.nan,
                                                                    grp["weight"]), "finite")):
        wire = {"groups": [dict(grp, weight=bad)], "energy_stencil": c.ctx.energy_stencil}
        with pytest.raises(ValueError, match=msg):
            make(wire)


def test_boundary_layer_target_profile_and_order():
    """Wall-normal spacing: first_height at the wall, geometric growth, the
    cap, and the smoothstep hand-over past n_layers (reference
    egg/smoothing/targets.py:196-212); W^-1 rows follow the wire's sample
    order (blocks, cells row-major, 2^d corners) with the layer counted from
    the wall side."""
    from egg_amd import grids

    kw = dict(first_height=0.1, growth=2.0, interior_spacing=1.0, n_layers=3)
    assert [grids.layer_spacing(k, **kw) for k in range(5)] == [0.1, 0.2, 0.4, 0.8, 1.0]  # capped at k = 4
    kw = dict(first_height=0.1, growth=1.5, interior_spacing=1.0, n_layers=3)
    w = (1 / 3) ** 2 * (3 - 2 / 3)  # smoothstep((4 - 3) / 3)
    assert grids.layer_spacing(4, **kw) == pytest.approx(0.1 * 1.5**4 * (1 - w) + w, rel=1e-15)
    assert grids.layer_spacing(6, **kw) == 1.0
    g = grids.single_block_2d(5)
    for side in (0, 1):
        W = grids.boundary_layer_w_inv(g, {0: (1, side)}, first_height=0.25, growth=2.0, tangential_spacing=1.0,
                                       n_layers=8)
        assert W.shape == (4 * 16, 2, 2)
        for i in range(4):
            for j in range(4):
                k = j if side == 0 else 3 - j
                rows = W[4 * (4 * i + j): 4 * (4 * i + j) + 4]
                np.testing.assert_array_equal(rows, np.broadcast_to(np.diag([1.0, 1.0 / min(0.25 * 2**k, 1.0)]),
                                                                    (4, 2, 2)))


def test_directional_eval_binding(oracle):
    """cpp_core.directional_eval (host side; reference src/bindings.cpp:1819)
    against the oracle's total and central differences of its own energy;
    a reserved kind is refused as in the reference."""
    from egg_amd import grids
    from egg_amd._cpp import cpp_core
    from tests.test_oracle import _dir_only_stencil

    g = grids.single_block_2d(9, seed=4, amp=0.25)
    s = grids.wall_directional_samples(g, weight=1.3, eps=1e-2)
    e, gr = cpp_core.directional_eval(g.X, s.nodes, s.kind, s.ref, s.weight, s.eps)
    E, _ = oracle.energy_mindet(2, "shape", 0.0, _dir_only_stencil(s), g.X)
    assert e == pytest.approx(E, rel=1e-12)
    h = 1e-6
    for node in np.unique(s.nodes[s.nodes >= 0])[::3]:
        for k in range(2):
            Xp, Xm = g.X.copy(), g.X.copy()
            Xp[node, k] += h
            Xm[node, k] -= h
            fd = (cpp_core.directional_eval(Xp, s.nodes, s.kind, s.ref, s.weight, s.eps)[0] -
                  cpp_core.directional_eval(Xm, s.nodes, s.kind, s.ref, s.weight, s.eps)[0]) / (2 * h)
            assert gr[node, k] == pytest.approx(fd, rel=1e-5, abs=1e-8)
    kind = s.kind.copy()
    kind[0] = 3
    with pytest.raises(ValueError, match="unsupported sample kind"):
        cpp_core.directional_eval(g.X, s.nodes, kind, s.ref, s.weight, s.eps)


def test_directional_wire_validation():
    """The session takes a valid directional wire (2D and 3D) and refuses a
    slot beyond its kind's nodes, a node outside the mesh, a non-finite
    weight."""
    from egg_amd import grids
    from tests.cases import Case

    g = grids.single_block_2d(9, seed=4, amp=0.2)
    c = Case(g, directional=grids.wall_directional_samples(g))
    info, _ = _tables(c)  # a valid wire passes; its windowed DOFs are the special ones
    assert info["n_stored"] > 0
    g3 = grids.single_block_3d(7, seed=2, amp=0.2)
    info3, _ = _tables(Case(g3, directional=grids.wall_directional_samples(g3)))
    assert info3["n_stored"] > 0
    for field, i, val, msg in [("dir_slot", 0, 2, "beyond its kind"), ("dir_nodes", 1, 10**6, "dir_nodes"),
                               ("dir_weight", 0, np.nan, "finite")]:
        c2 = Case(g, directional=grids.wall_directional_samples(g))
        grp = c2.ctx.wire["groups"][0]
        live = np.flatnonzero(np.repeat(grp["dir_slot"] >= 0, 4 if field == "dir_nodes" else 1))
        arr = np.array(grp[field], dtype=np.float64 if field == "dir_weight" else np.int32)
        arr[live[i]] = val
        grp[field] = arr
        with pytest.raises(ValueError, match=msg):
            _tables(c2)


def test_windowed_interior_dof_checked_like_a_plain_one():
    """A synthesized interior DOF that carries directional windows (it takes
    the special stored path) is validated exactly like a plain interior DOF:
    a logical index outside [1, n-2] or one naming another node is refused
    before any slot is computed from it."""
    from egg_amd import grids
    from tests.cases import Case

    g = grids.single_block_2d(17, seed=4, amp=0.2)
    c = Case(g, directional=grids.wall_directional_samples(g))
    grp = c.ctx.wire["groups"][0]
    K = int(grp["dir_k"])
    windowed = np.asarray(grp["dir_slot"]).reshape(-1, K).max(axis=1) >= 0
    cand = np.flatnonzero(windowed & (np.asarray(grp["interior_block"]) >= 0))
    assert cand.size, "the wall samples reach some synthesized interior DOF"
    info, _ = _tables(c)
    assert info["n_stored"] > 0
    i = int(cand[0])
    own = tuple(int(v) for v in np.asarray(grp["interior_logical"]).reshape(-1, 2)[i])
    other = (1, 1) if own != (1, 1) else (2, 2)
    for lg, msg in (((40, 40), r"outside \[1, n-2\]"), (other, "does not live at its block slot")):
        c2 = Case(g, directional=grids.wall_directional_samples(g))
        grp2 = c2.ctx.wire["groups"][0]
        il = np.array(grp2["interior_logical"], np.int32).reshape(-1, 2)
        il[i] = lg
        grp2["interior_logical"] = il.reshape(-1)
        with pytest.raises(ValueError, match=msg):
            _tables(c2)


def test_single_shared_weight_equals_full_array():
    """A length-1 ``weight`` (group and energy stencil) is one weight shared
    by every sample, as in the reference (src/bindings.cpp:247,364-366): it
    gives the same tables and weighted flag as the full array of that value."""
    from egg_amd import grids
    from egg_amd.smoothing import structured_arrays
    from tests.cases import Case

    g = grids.single_block_2d(9, seed=2, amp=0.2)
    ns = Case(g).ctx.energy_stencil["num_samples"]
    full = Case(g, weight=np.full(ns, 1.7))
    grp = dict(full.ctx.wire["groups"][0])
    es = dict(full.ctx.energy_stencil)
    one = {"groups": [dict(grp, weight=np.array([1.7]))], "energy_stencil": dict(es, weight=np.array([1.7]))}

    def make(wire):
        s = _core().CppStructuredSweepSession(wire, structured_arrays(full.bsc), full.X.ravel(), device="tables")
        return s.info(), s.tables()

    (ia, ta), (ib, tb) = make(full.ctx.wire), make(one)
    assert ia == ib
    for k in ta:
        np.testing.assert_array_equal(np.asarray(ta[k]), np.asarray(tb[k]))
    with pytest.raises(ValueError, match="weight size"):
        make({"groups": [dict(grp, weight=np.array([1.7, 1.7]))], "energy_stencil": es})


def test_raw_ctypes_session_tables_only():
    """The C ABI driven by plain ctypes structs (no pybind module): a
    tables-only session (device 3) builds the same index maps as the pybind
    path, and validation errors come back as INVALID_ARGUMENT with the
    thread's last-error message."""
    import ctypes as C

    from egg_amd.smoothing import structured_arrays
    from tests import capi_ctypes as CA

    lib = CA.load()
    c = ogrid(17, 9)
    s = CA.Session(lib, c.ctx.wire, structured_arrays(c.bsc), c.X, 2, device=3)
    info = CA.Info()  # the whole struct: the library writes every field
    s.check(lib.eggmi_session_info(s.handle, C.byref(info)))
    ref_info, _ = _tables(c)
    keys = ("n_slots", "n_overflow", "n_halo", "n_fan", "n_share", "structured_energy")
    assert [getattr(info, k) for k in keys] == [ref_info[k] for k in keys]
    s.destroy()
    bad = dict(c.ctx.wire["groups"][0])
    bad["dof_idx"] = np.array(bad["dof_idx"]) + 10**6
    with pytest.raises(ValueError, match="dof_idx"):
        CA.Session(lib, {"groups": [bad], "energy_stencil": c.ctx.energy_stencil}, structured_arrays(c.bsc), c.X, 2,
                   device=3)


def test_raw_ctypes_control_topo_arguments():
    """The multi-block control entry points through plain ctypes on a
    tables-only session: missing arguments are INVALID_ARGUMENT, calls that
    need a wire without one say so, and attaching nets needs device state."""
    import ctypes as C

    from egg_amd.smoothing import structured_arrays
    from tests import capi_ctypes as CA

    lib = CA.load()
    c = ogrid(17, 9)
    s = CA.Session(lib, c.ctx.wire, structured_arrays(c.bsc), c.X, 2, device=3)
    try:
        with pytest.raises(ValueError, match="control topology"):
            s.check(lib.eggmi_session_set_control_topo(s.handle, None))
        off = np.zeros(2, np.int32)
        with pytest.raises(ValueError, match="no multi-block control wire"):
            s.check(lib.eggmi_control_set_reduction(s.handle, C.c_int64(0), off.ctypes.data_as(C.c_void_p), None,
                                                    None))
        with pytest.raises(ValueError, match="no multi-block control wire"):
            s.check(lib.eggmi_control_set_penalty(s.handle, C.c_int64(0), None, None, None, None))
        with pytest.raises(ValueError, match="no multi-block control wire"):
            s.check(lib.eggmi_control_set_b(s.handle, 0, None, C.c_int64(0)))
        topo = (C.c_byte * 256)()  # a zeroed eggmi_control_topo (the struct is smaller)
        with pytest.raises(ValueError, match="tables only"):
            s.check(lib.eggmi_session_set_control_topo(s.handle, C.byref(topo)))
    finally:
        s.destroy()
