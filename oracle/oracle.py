"""ctypes front-end of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Loads ``oracle/liboracle.so`` (built from ``oracle/egg_oracle.c`` by
``oracle/Makefile``). Only tests/, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` import this module; the product path in
``egg_amd`` never does.

Everything is in GLOBAL node-id space (see the header of egg_oracle.c). The
per-DOF patch of an interior DOF is materialised here in the reference's
occurrence order (src/structured_patch.hpp:89-130: cell bits outer, corner
bits inner, axis 0 most significant), so the C sweep only has a stored path.
"""

from __future__ import annotations

import ctypes as C
import os
import subprocess
from itertools import product
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
_LIB = None

KIND = {"barrier": 0, "shape": 0, "untangle": 1, "shape_size": 2}

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_fp = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_lp = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")


def build() -> Path:
    """Compile the oracle library with make (gcc)."""
    subprocess.run(["make", "-s", "-C", str(_HERE)], check=True)
    return _HERE / "liboracle.so"


def lib():
    global _LIB
    if _LIB is None:
        so = _HERE / "liboracle.so"
        if not so.exists():
            build()
        L = C.CDLL(str(so))
        for sfx, rp in (("_f64", _dp), ("_f32", _fp)):
            rt = C.c_double if sfx == "_f64" else C.c_float
            f = getattr(L, "o_metric2d" + sfx)
            f.argtypes = [C.c_int, rp, rt, rp, rp, rp]
            f.restype = None
            f = getattr(L, "o_metric3d" + sfx)
            f.argtypes = [C.c_int, rp, rt, rp, rp, rp]
            f.restype = None
            f = getattr(L, "o_patch_eval" + sfx)
            f.argtypes = [C.c_int, C.c_int, rt, C.c_int, _ip, _ip, rp, _ip, rp, C.c_int, C.c_void_p, rp, rp, rp, rp,
                          rp]
            f.restype = None
            f = getattr(L, "o_solve" + sfx)
            f.argtypes = [C.c_int, rp, rp, rp]
            f.restype = None
            f = getattr(L, "o_entity_eval" + sfx)
            f.argtypes = [C.c_int, rp, rp, rp, rp]
            f.restype = None
            f = getattr(L, "o_entity_frame" + sfx)
            f.argtypes = [rp, rp, rp, rp]
            f.restype = None
            f = getattr(L, "o_entity_param" + sfx)
            f.argtypes = [rp, C.c_int, rp, C.c_void_p, rp]
            f.restype = None
            f = getattr(L, "o_curv_eval" + sfx)
            f.argtypes = [rp, rt, C.c_int, rp, rp, rp]
            f.restype = None
            f = getattr(L, "o_dir_eval" + sfx)
            f.argtypes = [C.c_int, C.c_int, rp, _dp, rt, rt, C.c_int, rp, rp, rp]
            f.restype = None
            f = getattr(L, "o_energy_mindet" + sfx)
            f.argtypes = [C.c_int, C.c_int, C.c_double, C.c_longlong, _ip, _ip, rp, rp, C.c_int, C.c_void_p, rp,
                          C.c_void_p, _dp, _dp]
            f.restype = None
            f = getattr(L, "o_jacobi" + sfx)
            f.argtypes = [
                C.c_int, C.c_int, C.c_double, C.c_double, C.c_int, C.c_int, C.c_int,
                _ip, _lp, _ip, _ip, rp, _ip, rp, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, rp, C.c_void_p,
                C.c_longlong, _ip, _ip, rp, rp, C.c_int, C.c_void_p, rp, _dp, _dp,
            ]
            f.restype = C.c_int
            f = getattr(L, "o_gauss_seidel" + sfx)
            f.argtypes = [C.c_int, C.c_int, C.c_double, C.c_int, C.c_int, _ip, _lp, _ip, _ip, rp, _ip, rp, C.c_int, rp]
            f.restype = C.c_int
            f = getattr(L, "o_patch_grads" + sfx)
            f.argtypes = [C.c_int, C.c_int, C.c_double, C.c_int, _ip, _lp, _ip, _ip, rp, _ip, rp, C.c_int,
                          C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, rp, C.c_void_p, rp]
            f.restype = None
        L.o_set_threads.argtypes = [C.c_int]
        L.o_set_threads.restype = None
        _LIB = L
    return _LIB


def _dt(prec: str):
    return np.float64 if prec == "f64" else np.float32


def _weights(w, dt):
    """Per-sample energy weights as a contiguous array of the oracle's real
    type, or None (unit weight everywhere)."""
    return None if w is None else np.ascontiguousarray(w, dt).reshape(-1)


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class _DirT(C.Structure):
    """The C oracle's directional table (dir_t)."""

    _fields_ = [("K", C.c_int), ("num", C.c_longlong), ("nodes", C.c_void_p), ("slot", C.c_void_p),
                ("kind", C.c_void_p), ("ref", C.c_void_p), ("wt", C.c_void_p), ("eps", C.c_double)]


def _dir_table(t, dim, per_dof):
    """A wire directional table (per-DOF "dir_k" windows, or the energy
    stencil's whole-sample "dir_num" table) as a (_DirT, arrays) pair, or
    (None, None) when absent. Node ids are global."""
    key = "dir_k" if per_dof else "dir_num"
    if t is None or int(t.get(key, 0) or 0) == 0:
        return None, None
    arrs = [np.ascontiguousarray(np.asarray(t["dir_nodes"], np.int32).reshape(-1)),
            np.ascontiguousarray(np.asarray(t["dir_slot"], np.int32).reshape(-1)) if per_dof else None,
            np.ascontiguousarray(np.asarray(t["dir_kind"], np.int32).reshape(-1)),
            np.ascontiguousarray(np.asarray(t["dir_ref"], np.float64).reshape(-1)),
            np.ascontiguousarray(np.asarray(t["dir_weight"], np.float64).reshape(-1))]
    n = int(t[key])
    st = _DirT(n if per_dof else 0, 0 if per_dof else n, *(_ptr(a) for a in arrs), float(t["dir_eps"]))
    assert arrs[2].size == (len(arrs[4])) and arrs[3].size == dim * arrs[2].size
    return st, arrs


def _byref(st):
    return None if st is None else C.cast(C.pointer(st), C.c_void_p)


def metric(dim: int, kind: str | int, t, delta: float = 0.0, prec: str = "f64"):
    """(value, grad, Hessian) of the per-sample metric at vec(T)."""
    dt = _dt(prec)
    vt = dim * dim
    k = KIND[kind] if isinstance(kind, str) else kind
    t = np.ascontiguousarray(t, dtype=dt)
    val = np.zeros(1, dt)
    g = np.zeros(vt, dt)
    H = np.zeros(vt * vt, dt)
    fn = getattr(lib(), f"o_metric{dim}d_{prec}")
    fn(k, t, dt(delta), val, g, H)
    return float(val[0]), g, H.reshape(vt, vt)


def patch_eval(dim, kind, delta, gc, gn, s, role, W, X, prec="f64", weight=None):
    """(grad, hess, energy, mindet) of one patch. gn, s: (dim, P);
    W: (P, dim*dim) per-occurrence rows or (dim*dim,) uniform; weight: (P,)
    per-occurrence energy weight or None (unit)."""
    dt = _dt(prec)
    P = len(gc)
    W = np.ascontiguousarray(W, dtype=dt)
    wt = _weights(weight, dt)
    ws = 0 if W.ndim == 1 else dim * dim
    grad = np.zeros(dim, dt)
    hess = np.zeros(dim * dim, dt)
    e = np.zeros(1, dt)
    m = np.zeros(1, dt)
    fn = getattr(lib(), f"o_patch_eval_{prec}")
    fn(
        dim, KIND[kind], dt(delta), P,
        np.ascontiguousarray(gc, np.int32), np.ascontiguousarray(gn, np.int32).reshape(-1),
        np.ascontiguousarray(s, dt).reshape(-1), np.ascontiguousarray(role, np.int32),
        W.reshape(-1), ws, _ptr(wt), np.ascontiguousarray(X, dt).reshape(-1), grad, hess, e, m,
    )
    return grad, hess.reshape(dim, dim), float(e[0]), float(m[0])


def solve(dim, H, g, prec="f64"):
    dt = _dt(prec)
    x = np.zeros(dim, dt)
    getattr(lib(), f"o_solve_{prec}")(
        dim, np.ascontiguousarray(H, dt).reshape(-1), np.ascontiguousarray(g, dt), x
    )
    return x


# ---- entity records: the reference wire's per-kind SoA fields and CSR
# payloads (egg/geometry/entity_soa.py:41-60, :598-688) as one record per
# entity (value 0 the tag, then fields and payloads, as the compute core
# lays them out; sphere and cylinder gain az = ax x ay, src/geometry.hpp:
# 1585-1606; a composite's segments come from its positional arena records,
# egg/geometry/entity_encoding.py:82-178).
WIRE_FIELDS = {0: 0, 1: 4, 2: 3, 3: 4, 4: 10, 5: 9, 6: 6, 7: 8, 8: 8, 9: 10, 10: 6, 11: 2, 12: 10, 13: 8, 14: 9,
               15: 4, 16: 6}


def _bspline_rec(deg, n, t0, t1, closed, has_w, knots, ctrl, w):
    deg, n = int(deg), int(n)
    assert len(knots) == n + deg + 1 and len(ctrl) == 2 * n and (not has_w or len(w) == n)
    out = [10.0, deg, n, t0, t1, closed, has_w, *knots, *ctrl]
    return out + (list(w) if has_w else [])


def entity_record(tag: int, rec, seg=()) -> np.ndarray:
    """One entity's record from its wire fields ``rec`` and payload arrays
    ``seg`` (one per CSR slot, this entity's slice)."""
    r = [float(v) for v in np.asarray(rec, np.float64).reshape(-1)]
    seg = [np.asarray(a, np.float64).reshape(-1) for a in seg]
    if tag in (4, 12):
        ax = np.array(r[4:7] if tag == 4 else r[3:6])
        ay = np.array(r[7:10] if tag == 4 else r[6:9])
        az = list(np.cross(ax, ay))
        out = [float(tag)] + (r[:10] + az if tag == 4 else r[:9] + az + [r[9]])
    elif tag == 10:
        out = _bspline_rec(*r[:6], seg[0], seg[1], seg[2] if len(seg) > 2 else [])
    elif tag == 11:
        arena = seg[0]
        m, ro = int(r[0]), int(r[1])
        subs = []
        for i in range(m):
            sr = arena[ro + 13 * i: ro + 13 * (i + 1)]
            st, q = int(sr[0]), sr[1:]
            if st == 10:
                n, deg = int(q[1]), int(q[0])
                kn = arena[int(q[2]): int(q[2]) + n + deg + 1]
                cp = arena[int(q[3]): int(q[3]) + 2 * n]
                ww = arena[int(q[6]): int(q[6]) + n] if q[7] else []
                subs.append(_bspline_rec(q[0], q[1], q[4], q[5], 0.0, q[7], kn, cp, ww))
            else:
                subs.append([float(st)] + list(q[:WIRE_FIELDS[st]]))
        out = [11.0, float(m)] + [0.0] * m
        for i, sb in enumerate(subs):
            out[2 + i] = float(len(out))
            out += sb
    elif tag == 14:
        ku, kv, cp, ww, verts, loops = seg
        out = [14.0, r[0], r[1], r[2], r[3], r[8], float(len(loops)), *ku, *kv, *cp]
        out += list(ww) if r[8] else []
        out += list(loops) + (list(verts) if len(loops) >= 2 else [])
    else:
        out = [float(tag)] + r[:WIRE_FIELDS[tag]]
    return np.asarray(out, np.float64)


def wire_entities(g: dict, ndof: int):
    """Per group-local DOF: entity tag (0 free), record offset and the
    records, from a wire group's ``entities``; (None, None, None) when no
    DOF is constrained."""
    ents = g.get("entities") or {}
    tag = np.zeros(ndof, np.int32)
    off = np.zeros(ndof, np.int64)
    blobs = []
    size = 0
    for part in ents.values():
        t, kf = int(part["tag"]), int(part["kFields"])
        loc = np.asarray(part["dof_local"], np.int64)
        recs = np.asarray(part.get("records", []), np.float64).reshape(len(loc), kf) if kf else None
        seg = part.get("seg") or []
        for i, l in enumerate(loc):
            tag[l] = t
            if t == 0:
                continue
            sl = [np.asarray(sd["data"], np.float64)[int(sd["off"][i]): int(sd["off"][i + 1])] for sd in seg]
            b = entity_record(t, recs[i], sl)
            off[l] = size
            blobs.append(b)
            size += b.size
    if not np.any(tag):
        return None, None, None
    return tag, off, np.concatenate(blobs)


def entity_frame(record, p, prec="f64"):
    """(foot, orthonormal tangent basis (dim, k)) of the point p on one
    entity record."""
    dt = _dt(prec)
    rec = np.ascontiguousarray(record, dt)
    tag = int(rec[0])
    dim = 3 if tag in (4, 5, 12, 13, 14, 16) else 2
    k = 2 if tag in (4, 5, 12, 14) else 1
    pos = np.zeros(dim, dt)
    basis = np.zeros(k * dim, dt)
    getattr(lib(), f"o_entity_frame_{prec}")(rec, np.ascontiguousarray(p, dt), pos, basis)
    return pos, basis.reshape(k, dim).T


def entity_param(record, mode, x, seed=None, prec="f64"):
    """The record's parametrization: mode 0 point at x, 1 raw frame at x
    (dim, k), 2 foot parameter of the point x (cold; from ``seed`` for the
    iterative kinds when given)."""
    dt = _dt(prec)
    rec = np.ascontiguousarray(record, dt)
    tag = int(rec[0])
    dim = 3 if tag in (4, 5, 12, 13, 14, 16) else 2
    k = 2 if tag in (4, 5, 12, 14) else 1
    out = np.zeros(dim if mode == 0 else (k * dim if mode == 1 else k), dt)
    sd = None if seed is None else np.ascontiguousarray(np.resize(np.asarray(seed, dt), 2))
    getattr(lib(), f"o_entity_param_{prec}")(rec, int(mode), np.ascontiguousarray(np.resize(np.asarray(x, dt), 3)),
                                             _ptr(sd), out)
    return out.reshape(k, dim).T if mode == 1 else out


def entity_eval(tag: int, record, p, prec="f64"):
    """Projection of the 2D point p onto entity (tag, record) and the unit
    tangent there (restating the reference's geometry_project /
    geometry_tangent for the closed-form kinds)."""
    dt = _dt(prec)
    rec = np.zeros(4, dt)
    rec[: len(record)] = record
    proj = np.zeros(2, dt)
    tang = np.zeros(2, dt)
    getattr(lib(), f"o_entity_eval_{prec}")(int(tag), rec, np.ascontiguousarray(p, dt), proj, tang)
    return proj, tang


def curv_eval(p4, w, slot, prec="f64"):
    """Gradient, Hessian (2x2) and energy of one curvature window (four 2D
    points) for its node `slot`."""
    dt = _dt(prec)
    g, H, e = np.zeros(2, dt), np.zeros(4, dt), np.zeros(1, dt)
    getattr(lib(), f"o_curv_eval_{prec}")(np.ascontiguousarray(p4, dt).reshape(-1), float(w), int(slot), g, H, e)
    return g, H.reshape(2, 2), float(e[0])


def dir_eval(dim, kind, p, ref, w, eps, slot, prec="f64"):
    """Gradient, Hessian (dim x dim) and energy of one directional sample
    (positions p, arity x dim) for its node `slot`."""
    dt = _dt(prec)
    g, H, e = np.zeros(dim, dt), np.zeros(dim * dim, dt), np.zeros(1, dt)
    getattr(lib(), f"o_dir_eval_{prec}")(int(dim), int(kind), np.ascontiguousarray(p, dt).reshape(-1),
                                         np.ascontiguousarray(ref, np.float64), float(w), float(eps), int(slot),
                                         g, H, e)
    return g, H.reshape(dim, dim), float(e[0])


def _stencil_arrays(es: dict, dim: int, dt):
    ns = int(es["num_samples"])
    gc = np.ascontiguousarray(es["gc"], np.int32)
    gn = np.ascontiguousarray(np.stack([np.asarray(es[f"gn{k}"]) for k in range(dim)]), np.int32)
    s = np.ascontiguousarray(np.stack([np.asarray(es[f"s{k}"]) for k in range(dim)]), dt)
    W = np.ascontiguousarray(es["W_inv"], dt).reshape(-1)
    ws = 0 if W.size == dim * dim else dim * dim
    return ns, gc, gn.reshape(-1), s.reshape(-1), W, ws, _weights(es.get("weight"), dt)


def energy_mindet(dim, kind, delta, es, X, prec="f64"):
    """(sum of weighted mu, min det A) over the energy stencil, accumulated
    in double (the stencil's optional "weight" scales mu, not det)."""
    dt = _dt(prec)
    ns, gc, gn, s, W, ws, wt = _stencil_arrays(es, dim, dt)
    dts, keep = _dir_table(es, dim, False)
    E = np.zeros(1)
    m = np.zeros(1)
    getattr(lib(), f"o_energy_mindet_{prec}")(
        dim, KIND[kind], float(delta), ns, gc, gn, s, W, ws, _ptr(wt),
        np.ascontiguousarray(X, dt).reshape(-1), _byref(dts), E, m,
    )
    return float(E[0]), float(m[0])


def interior_occurrences(dim: int, logical, idx):
    """Occurrences of an interior DOF's patch in the reference order
    (src/structured_patch.hpp:89-130). ``idx`` maps a logical tuple to a
    node id. Returns lists gc, gn[dim], s[dim], role."""
    nc = 1 << dim
    gcs, gns, ss, roles = [], [[] for _ in range(dim)], [[] for _ in range(dim)], []
    for occ in range(nc * nc):
        cell, o = divmod(occ, nc)
        obit = [(o >> (dim - 1 - k)) & 1 for k in range(dim)]
        corner = [logical[k] - 1 + ((cell >> (dim - 1 - k)) & 1) + obit[k] for k in range(dim)]
        gcs.append(idx(tuple(corner)))
        role = 0 if list(corner) == list(logical) else -1
        for k in range(dim):
            nbr = list(corner)
            nbr[k] += 1 if obit[k] == 0 else -1
            gns[k].append(idx(tuple(nbr)))
            ss[k].append(1.0 if obit[k] == 0 else -1.0)
            if role < 0 and nbr == list(logical):
                role = k + 1
        roles.append(role)
    return gcs, gns, ss, roles


class Patches:
    """Per-DOF patches in global id space (CSR over DOFs)."""

    def __init__(self, dim, dof, poff, gc, gn, s, role, W, wt=None, etag=None, eoff=None, eblob=None, curv=None,
                 dirw=None):
        self.dim = dim
        self.dof = np.ascontiguousarray(dof, np.int32)
        self.poff = np.ascontiguousarray(poff, np.int64)
        self.gc = np.ascontiguousarray(gc, np.int32)
        self.gn = np.ascontiguousarray(gn, np.int32)  # (dim, total)
        self.s = np.ascontiguousarray(s, np.float64)  # (dim, total)
        self.role = np.ascontiguousarray(role, np.int32)
        self.W = np.ascontiguousarray(W, np.float64)  # (dd,) uniform or (total, dd)
        self.wt = None if wt is None else np.ascontiguousarray(wt, np.float64)  # (total,) or None = unit
        # per-DOF entity a constrained boundary DOF slides on (0 free) and
        # the offset of its record in eblob (wire_entities)
        self.etag = None if etag is None else np.ascontiguousarray(etag, np.int32)
        self.eoff = None if eoff is None else np.ascontiguousarray(eoff, np.int64)
        self.eblob = None if eblob is None else np.ascontiguousarray(eblob, np.float64)
        # per-DOF curvature windows (K, nodes (ndof, K, 4), slot (ndof, K), weight (ndof, K)) or None
        self.curv = curv
        # per-DOF directional windows (a wire dict with dir_k, dir_eps, dir_nodes,
        # dir_slot, dir_kind, dir_ref, dir_weight rows in this DOF order) or None
        self.dirw = dirw


def materialize_patches(wire: dict, block_global_dof: list[np.ndarray], dim: int, dofs=None) -> Patches:
    """Expand the wire's single group into explicit per-DOF patches: stored
    occurrences for DOFs with P_of > 0, synthesized ones (reference order,
    ``interior_occurrences``) for interior DOFs (interior_block >= 0).
    Uniform W_inv only for the synthesized case, as in the reference
    (identity). The interior expansion is the per-DOF rule of
    ``interior_occurrences`` evaluated for all DOFs of a block at once
    (``tests/test_oracle.py`` checks the two agree). ``dofs`` (positions in
    the group's DOF list, ascending) restricts the expansion to a sample: a
    Jacobi sweep of the sample moves exactly those DOFs as the full sweep
    would."""
    g = wire["groups"][0]
    P_all = np.asarray(g["P_of"], np.int64)
    soff_all = np.concatenate([[0], np.cumsum(P_all)])
    sel = np.arange(len(P_all)) if dofs is None else np.asarray(dofs, np.int64)
    dof = np.asarray(g["dof_idx"], np.int32)[sel]
    n = len(dof)
    P_of = P_all[sel]
    iblk = np.asarray(g["interior_block"], np.int32)[sel]
    ilog = np.asarray(g["interior_logical"], np.int64).reshape(-1, dim)[sel]
    W = np.asarray(g["W_inv"], np.float64)
    uniform = W.size == dim * dim
    wire_wt = g.get("weight")
    nc = 1 << dim
    interior = iblk >= 0
    if interior.any() and not uniform:
        raise ValueError("interior synthesis requires a uniform W_inv")
    if interior.any() and wire_wt is not None:
        raise ValueError("interior synthesis requires unit sample weights")
    cnt = np.where(interior, nc * nc, P_of)
    poff = np.concatenate([[0], np.cumsum(cnt)]).astype(np.int64)
    total = int(poff[-1])
    gc = np.zeros(total, np.int64)
    gn = np.zeros((dim, total), np.int64)
    sg = np.zeros((dim, total), np.float64)
    role = np.zeros(total, np.int64)
    # stored DOFs: their wire occurrences in order
    st = np.nonzero(~interior)[0]
    soff = soff_all[sel]
    if len(st):
        src = np.concatenate([np.arange(soff[d], soff[d] + P_of[d]) for d in st]).astype(np.int64)
        dst = np.concatenate([np.arange(poff[d], poff[d + 1]) for d in st]).astype(np.int64)
        gc[dst] = np.asarray(g["gc"])[src]
        role[dst] = np.asarray(g["role"])[src]
        for k in range(dim):
            gn[k, dst] = np.asarray(g[f"gn{k}"])[src]
            sg[k, dst] = np.asarray(g[f"s{k}"])[src]
    # interior DOFs, block by block, occurrence by occurrence
    for b in np.unique(iblk[interior]):
        ds = np.nonzero(iblk == b)[0]
        ids = np.asarray(block_global_dof[b])
        lg = ilog[ds]
        for occ in range(nc * nc):
            cell, o = divmod(occ, nc)
            obit = [(o >> (dim - 1 - k)) & 1 for k in range(dim)]
            corner = lg - 1 + np.array([((cell >> (dim - 1 - k)) & 1) + obit[k] for k in range(dim)])
            pos = poff[ds] + occ
            gc[pos] = ids[tuple(corner.T)]
            r = np.where(np.all(corner == lg, axis=1), 0, -1)
            for k in range(dim):
                nbr = corner.copy()
                nbr[:, k] += 1 if obit[k] == 0 else -1
                gn[k, pos] = ids[tuple(nbr.T)]
                sg[k, pos] = 1.0 if obit[k] == 0 else -1.0
                r = np.where((r < 0) & np.all(nbr == lg, axis=1), k + 1, r)
            role[pos] = r
    if uniform:
        Wp = W.reshape(-1)
    else:
        Wp = np.concatenate([W.reshape(-1, dim * dim)[soff[d]:soff[d] + P_of[d]] for d in range(n)])
    wt = None
    if wire_wt is not None:
        ww = np.asarray(wire_wt, np.float64).reshape(-1)
        wt = np.concatenate([ww[soff[d]:soff[d] + P_of[d]] for d in range(n)]) if n else ww[:0]
    etag, eoff, eblob = wire_entities(g, len(P_all))
    if etag is not None:
        etag, eoff = etag[sel], eoff[sel]
        if np.any(etag[interior] != 0):
            raise ValueError("a synthesized interior DOF cannot be constrained")
    curv = None
    K = int(g.get("curv_k", 0) or 0)
    if K > 0:
        curv = (K, np.ascontiguousarray(np.asarray(g["curv_nodes"], np.int32).reshape(-1, K, 4)[sel]),
                np.ascontiguousarray(np.asarray(g["curv_slot"], np.int32).reshape(-1, K)[sel]),
                np.ascontiguousarray(np.asarray(g["curv_weight"], np.float64).reshape(-1, K)[sel]))
    dirw = None
    K = int(g.get("dir_k", 0) or 0)
    if K:
        dirw = {"dir_k": K, "dir_eps": float(g["dir_eps"]),
                "dir_nodes": np.asarray(g["dir_nodes"], np.int32).reshape(-1, K, 4)[sel],
                "dir_slot": np.asarray(g["dir_slot"], np.int32).reshape(-1, K)[sel],
                "dir_kind": np.asarray(g["dir_kind"], np.int32).reshape(-1, K)[sel],
                "dir_ref": np.asarray(g["dir_ref"], np.float64).reshape(-1, K, dim)[sel],
                "dir_weight": np.asarray(g["dir_weight"], np.float64).reshape(-1, K)[sel]}
    return Patches(dim, dof, poff, gc, gn, sg, role, Wp, wt, etag, eoff, eblob, curv, dirw)


def materialize_patches_loop(wire: dict, block_global_dof: list[np.ndarray], dim: int) -> Patches:
    """DOF-by-DOF form of materialize_patches (small grids; the check on it)."""
    g = wire["groups"][0]
    dof = np.asarray(g["dof_idx"], np.int32)
    P_of = np.asarray(g["P_of"], np.int64)
    iblk = np.asarray(g["interior_block"], np.int32)
    ilog = np.asarray(g["interior_logical"], np.int32).reshape(len(dof), dim)
    soff = np.concatenate([[0], np.cumsum(P_of)])
    gc_all, gn_all, s_all, role_all, poff = [], [[] for _ in range(dim)], [[] for _ in range(dim)], [], [0]
    W = np.asarray(g["W_inv"], np.float64)
    uniform = W.size == dim * dim
    wire_wt = g.get("weight")
    Wrows, wts = [], []
    for d in range(len(dof)):
        if iblk[d] >= 0:
            ids = block_global_dof[iblk[d]]
            gcs, gns, ss, roles = interior_occurrences(dim, tuple(ilog[d]), lambda l: int(ids[l]))
            if not uniform or wire_wt is not None:
                raise ValueError("interior synthesis requires a uniform W_inv and unit weights")
        else:
            a, b = soff[d], soff[d + 1]
            gcs = list(np.asarray(g["gc"])[a:b])
            gns = [list(np.asarray(g[f"gn{k}"])[a:b]) for k in range(dim)]
            ss = [list(np.asarray(g[f"s{k}"])[a:b]) for k in range(dim)]
            roles = list(np.asarray(g["role"])[a:b])
            if not uniform:
                Wrows.append(W.reshape(-1, dim * dim)[a:b])
            if wire_wt is not None:
                wts.append(np.asarray(wire_wt, np.float64)[a:b])
        gc_all += gcs
        for k in range(dim):
            gn_all[k] += gns[k]
            s_all[k] += ss[k]
        role_all += roles
        poff.append(poff[-1] + len(gcs))
    Wp = W.reshape(-1) if uniform else np.concatenate(Wrows)
    wt = None if wire_wt is None else (np.concatenate(wts) if wts else np.zeros(0))
    return Patches(dim, dof, poff, gc_all, gn_all, s_all, role_all, Wp, wt)


def jacobi(dim, kind, delta, omega, n_sweeps, X, patches: Patches, es: dict, *, report_every=0,
           prec="f64", tau=None, threads=None, seeds=None):
    """Run n_sweeps of block-Jacobi in global id space (one run: its first
    sweep is cold for the constrained DOFs). ``seeds`` (ndof, 2; NaN = none)
    carries the constrained DOFs' foot parameters between runs and is
    updated in place. Returns (X_out, energies, mindets)."""
    dt = _dt(prec)
    if threads is not None:
        lib().o_set_threads(int(threads))
    Xw = np.ascontiguousarray(X, dt).reshape(-1).copy()
    n_nodes = Xw.size // dim
    work = np.zeros_like(Xw)
    k = n_sweeps if report_every <= 0 else report_every
    nrep = (n_sweeps + k - 1) // k
    en = np.zeros(max(nrep, 1))
    md = np.zeros(max(nrep, 1))
    ns, egc, egn, es_, eW, ews, ewt = _stencil_arrays(es, dim, dt)
    W = np.ascontiguousarray(patches.W, dt).reshape(-1)
    ws = 0 if W.size == dim * dim else dim * dim
    pwt = _weights(patches.wt, dt)
    etag = patches.etag
    eoff = patches.eoff
    eblob = None if patches.eblob is None else np.ascontiguousarray(patches.eblob, dt)
    if seeds is None:
        seeds = np.full((len(patches.dof), 2), np.nan)
    sd = np.ascontiguousarray(seeds, dt).reshape(-1)
    ck, cn, cs, cw = _curv_args(patches, dt)
    dw, keep_w = _dir_table(patches.dirw, dim, True)
    dts, keep_s = _dir_table(es, dim, False)
    tau_arr = None if tau is None else np.ascontiguousarray(tau, dt).reshape(-1)
    got = getattr(lib(), f"o_jacobi_{prec}")(
        dim, KIND[kind], float(delta), float(omega), int(n_sweeps), int(report_every),
        len(patches.dof), patches.dof, patches.poff, patches.gc, patches.gn.reshape(-1),
        np.ascontiguousarray(patches.s, dt).reshape(-1), patches.role, W, ws, _ptr(pwt), _ptr(etag), _ptr(eoff), _ptr(eblob),
        _ptr(sd),
        ck, _ptr(cn), _ptr(cs), _ptr(cw), _byref(dw), _byref(dts), n_nodes, Xw, _ptr(tau_arr), ns, egc, egn, es_, eW, ews, _ptr(ewt),
        work, en, md,
    )
    if isinstance(seeds, np.ndarray) and seeds.dtype == sd.dtype and seeds.flags.c_contiguous:
        pass  # updated in place
    elif isinstance(seeds, np.ndarray):
        seeds[...] = sd.reshape(seeds.shape)
    return Xw.reshape(-1, dim), en[:got], md[:got]


def gauss_seidel(dim, kind, X, patches: Patches, *, delta=0.0, max_sweeps=10000, prec="f64"):
    """The reference's sequential node baseline (in-place Gauss-Seidel
    Newton, tests/smoothing/test_control_ref.py:95-135) to convergence or
    max_sweeps; returns (X_out, sweeps)."""
    dt = _dt(prec)
    Xw = np.ascontiguousarray(X, dt).reshape(-1).copy()
    W = np.ascontiguousarray(patches.W, dt).reshape(-1)
    ws = 0 if W.size == dim * dim else dim * dim
    n = getattr(lib(), f"o_gauss_seidel_{prec}")(
        dim, KIND[kind], float(delta), int(max_sweeps), len(patches.dof), patches.dof, patches.poff, patches.gc,
        patches.gn.reshape(-1), np.ascontiguousarray(patches.s, dt).reshape(-1), patches.role, W, ws, Xw)
    return Xw.reshape(-1, dim), n


def _curv_args(patches: Patches, dt):
    if patches.curv is None:
        return 0, None, None, None
    K, nodes, slot, wt = patches.curv
    return (int(K), np.ascontiguousarray(nodes, np.int32).reshape(-1), np.ascontiguousarray(slot, np.int32).reshape(-1),
            np.ascontiguousarray(wt, dt).reshape(-1))


def patch_grads(dim, kind, X, patches: Patches, *, delta=0.0, tau=None, prec="f64"):
    """Per-node patch gradient (minus tau) at every DOF; zero elsewhere."""
    dt = _dt(prec)
    Xw = np.ascontiguousarray(X, dt).reshape(-1)
    out = np.zeros_like(Xw)
    W = np.ascontiguousarray(patches.W, dt).reshape(-1)
    ws = 0 if W.size == dim * dim else dim * dim
    pwt = _weights(patches.wt, dt)
    ck, cn, cs, cw = _curv_args(patches, dt)
    dw, keep = _dir_table(patches.dirw, dim, True)
    tau_arr = None if tau is None else np.ascontiguousarray(tau, dt).reshape(-1)
    getattr(lib(), f"o_patch_grads_{prec}")(
        dim, KIND[kind], float(delta), len(patches.dof), patches.dof, patches.poff, patches.gc,
        patches.gn.reshape(-1), np.ascontiguousarray(patches.s, dt).reshape(-1), patches.role, W, ws,
        _ptr(pwt), ck, _ptr(cn), _ptr(cs), _ptr(cw), _byref(dw), Xw, _ptr(tau_arr), out,
    )
    return out.reshape(-1, dim)
