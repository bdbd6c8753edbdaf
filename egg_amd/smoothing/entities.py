"""Boundary entities on the wire: the per-kind packed records and CSR
payloads of a group's ``entities`` field, as the reference's encoder lays
them out (egg/geometry/entity_soa.py:41-60 field table, :598-688 grouping;
a composite's self-contained arena of positional segment records,
egg/geometry/entity_encoding.py:82-178; unpacked by src/bindings.cpp:
412-446).

An entity is given as ``(kind, spec)``:

* fixed-size kinds take their record as a sequence, in the wire's field
  order: ``lineseg`` / ``linerail`` (sx, sy, ex, ey), ``circle`` (cx, cy,
  r), ``ellipse`` (cx, cy, rx, ry), ``circlearc`` (cx, cy, r, t0, t1,
  closed), ``ellipsearc`` (cx, cy, a, b, phi, t0, t1, closed),
  ``quadbezier`` (P0, P1, P2, t0, t1), ``cubicbezier`` (P0..P3, t0, t1),
  ``sphere`` (c(3), r, ax(3), ay(3)), ``plane`` (o(3), ax(3), ay(3)),
  ``cylinder`` (o(3), ax(3), ay(3), r), ``line3`` (p0(3), p1(3), t0, t1),
  ``linerail3`` (p0(3), p1(3));
* ``bspline``: a dict with ``degree``, ``knots``, ``ctrl`` (n, 2) and
  optional ``weights`` (NURBS), ``t0`` / ``t1`` (default: the live knot
  domain) and ``closed``;
* ``composite``: a list of ``(kind, spec)`` 2D curve segments (no
  composite among them);
* ``bsplinesurf``: a dict with ``degree`` (pu, pv), ``knots_u``,
  ``knots_v``, ``ctrl`` (nu, nv, 3), optional ``weights`` (nu, nv) and
  ``trim`` (a list of UV loops, each (k, 2): outer loop, then holes).
"""

from __future__ import annotations

import numpy as np

__all__ = ["ENTITY_TYPES", "encode_entity", "entity_groups"]

# kind -> (wire tag, packed fields, CSR payload slots)
ENTITY_TYPES = {
    "lineseg": (1, 4, 0), "circle": (2, 3, 0), "ellipse": (3, 4, 0), "sphere": (4, 10, 0), "plane": (5, 9, 0),
    "circlearc": (6, 6, 0), "ellipsearc": (7, 8, 0), "quadbezier": (8, 8, 0), "cubicbezier": (9, 10, 0),
    "bspline": (10, 6, 3), "composite": (11, 2, 1), "cylinder": (12, 10, 0), "line3": (13, 8, 0),
    "bsplinesurf": (14, 9, 6), "linerail": (15, 4, 0), "linerail3": (16, 6, 0),
}
_PARAM_PAD = 12  # positional parameters per composite segment record


def _bspline_parts(spec: dict):
    deg = int(spec["degree"])
    knots = np.asarray(spec["knots"], np.float64).reshape(-1)
    ctrl = np.asarray(spec["ctrl"], np.float64).reshape(-1, 2)
    n = ctrl.shape[0]
    if knots.size != n + deg + 1:
        raise ValueError(f"a degree-{deg} B-spline with {n} control points needs {n + deg + 1} knots")
    w = spec.get("weights")
    w = None if w is None else np.asarray(w, np.float64).reshape(-1)
    if w is not None and w.size != n:
        raise ValueError("one NURBS weight per control point")
    t0 = float(spec.get("t0", knots[deg]))
    t1 = float(spec.get("t1", knots[n]))
    return deg, n, knots, ctrl, w, t0, t1, float(bool(spec.get("closed", False)))


def _composite(segments):
    """[n_segs, rec_off] and the self-contained arena: B-spline payloads
    first, then one [tag, params...] record per segment."""
    arena: list[float] = []
    recs = []
    for kind, spec in segments:
        tag, kf, _ = ENTITY_TYPES[kind]
        params = np.zeros(_PARAM_PAD)
        if kind == "bspline":
            deg, n, knots, ctrl, w, t0, t1, _closed = _bspline_parts(spec)
            ko = len(arena)
            arena.extend(knots.tolist())
            co = len(arena)
            arena.extend(ctrl.reshape(-1).tolist())
            wo = 0.0
            if w is not None:
                wo = len(arena)
                arena.extend(w.tolist())
            params[:8] = (deg, n, ko, co, t0, t1, wo, float(w is not None))
        elif kind in ("lineseg", "circle", "ellipse", "circlearc", "ellipsearc", "quadbezier", "cubicbezier"):
            params[:kf] = np.asarray(spec, np.float64).reshape(kf)
        else:
            raise ValueError(f"a composite segment must be a 2D curve, got {kind!r}")
        recs.append((tag, params))
    rec_off = len(arena)
    for tag, params in recs:
        arena.append(float(tag))
        arena.extend(params.tolist())
    return np.array([len(recs), rec_off], np.float64), [np.asarray(arena, np.float64)]


def encode_entity(kind: str, spec):
    """(tag, kFields, record (kFields,), payloads or None) of one entity."""
    if kind not in ENTITY_TYPES:
        raise ValueError(f"entity type {kind!r} is not supported ({', '.join(ENTITY_TYPES)})")
    tag, kf, nseg = ENTITY_TYPES[kind]
    if kind == "bspline":
        deg, n, knots, ctrl, w, t0, t1, closed = _bspline_parts(spec)
        rec = np.array([deg, n, t0, t1, closed, float(w is not None)], np.float64)
        return tag, kf, rec, [knots, ctrl.reshape(-1), np.zeros(0) if w is None else w]
    if kind == "composite":
        rec, seg = _composite(spec)
        return tag, kf, rec, seg
    if kind == "bsplinesurf":
        pu, pv = (int(v) for v in spec["degree"])
        ku = np.asarray(spec["knots_u"], np.float64).reshape(-1)
        kv = np.asarray(spec["knots_v"], np.float64).reshape(-1)
        ctrl = np.asarray(spec["ctrl"], np.float64)
        nu, nv = ctrl.shape[:2]
        w = spec.get("weights")
        loops = spec.get("trim") or []
        verts = (np.concatenate([np.asarray(lp, np.float64).reshape(-1, 2) for lp in loops]) if loops
                 else np.zeros((0, 2)))
        offs = np.concatenate([[0], np.cumsum([len(lp) for lp in loops])]) if loops else np.zeros(0)
        rec = np.array([pu, pv, nu, nv, 0, 0, 0, 0, float(w is not None)], np.float64)
        return tag, kf, rec, [ku, kv, ctrl.reshape(-1), np.zeros(0) if w is None else np.asarray(w, float).reshape(-1),
                              verts.reshape(-1), offs.astype(np.float64)]
    rec = np.asarray(spec, np.float64).reshape(-1)
    if rec.size != kf:
        raise ValueError(f"a {kind} record holds {kf} values, got {rec.size}")
    return tag, kf, rec, None


def entity_groups(dofs: np.ndarray, entities: dict | None) -> dict:
    """The group's ``entities`` field: one sub-dict per kind present,
    ordered by its first DOF, with the group-local indices of its DOFs
    (ascending), their records back to back and, for the kinds with
    payloads, one ``{"data", "off"}`` CSR table per payload slot; DOFs not
    listed slide on nothing (``free``, tag 0)."""
    if not entities:
        return {}
    kind_of = np.full(dofs.size, "free", dtype=object)
    spec_of: list = [None] * dofs.size
    for node, (kind, spec) in entities.items():
        i = int(np.searchsorted(dofs, node))
        if i < dofs.size and dofs[i] == node:
            kind_of[i] = kind
            spec_of[i] = spec
    out = {}
    kinds = sorted(set(kind_of.tolist()), key=lambda k: int(np.argmax(kind_of == k)))
    for kind in kinds:
        local = np.flatnonzero(kind_of == kind)
        if kind == "free":
            out[kind] = {"tag": 0, "kFields": 0, "dof_local": local.astype(np.int32), "records": np.zeros(0)}
            continue
        enc = [encode_entity(kind, spec_of[i]) for i in local]
        tag, kf = enc[0][0], enc[0][1]
        part = {"tag": tag, "kFields": kf, "dof_local": local.astype(np.int32),
                "records": np.concatenate([e[2] for e in enc]).astype(np.float64)}
        if enc[0][3] is not None:
            part["seg"] = []
            for j in range(len(enc[0][3])):
                arrays = [e[3][j] for e in enc]
                off = np.zeros(len(arrays) + 1, np.int32)
                off[1:] = np.cumsum([a.size for a in arrays])
                part["seg"].append({"data": np.concatenate(arrays).astype(np.float64), "off": off})
        out[kind] = part
    return out
