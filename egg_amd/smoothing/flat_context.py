"""Flat sweep wire for structured quad/hex blocks.

Builds the ``{"groups", "energy_stencil"}`` dictionary the compute core
consumes, with the field names and conventions of the reference's
``build_flat_context`` (reference egg/smoothing/flat_context.py:136-346):

* every cell contributes ``2**d`` corner samples; sample id
  ``2**d * cell + b`` over the blocks concatenated in order, cells row-major,
  corner index ``b`` with axis 0 as the most significant bit;
* a sample is a corner plus its ``d`` axis neighbours (neighbour ``k`` flips
  bit ``k`` of ``b``) with sign ``+1`` when that bit is 0 and ``-1`` otherwise;
* a DOF strictly inside a block (logical index in ``[1, n-2]`` on every
  axis) under an isotropic uniform target ships no samples: its patch is
  synthesized by the core from ``(interior_block, interior_logical)``;
* every other moving DOF carries its incident samples in ascending
  (cell, corner) order with its role (0 corner, ``1+k`` axis-``k``
  neighbour, -1 present but neither).

Only the identity-family targets and plain shape samples are supported here
(the segment, circle, ellipse and rail entities, curvature windows and
directional terms are supported; no other geometry entities).
"""

from __future__ import annotations

import numpy as np

from .directional import directional_dof_table, directional_energy_table


def _bit(d: int, k: int) -> int:
    return 1 << (d - 1 - k)


def cell_stencil(blocks: list[np.ndarray], d: int) -> dict:
    """Per-sample stencil (gc, gn[k], s[k]) and node->sample membership rows
    (node, sample, role) for the cells of all blocks."""
    nc = 1 << d
    corner_cols = []
    for ids in blocks:
        ids = np.asarray(ids)
        shp = ids.shape
        if any(s < 2 for s in shp):
            continue
        cols = []
        for b in range(nc):
            sl = tuple(slice((b >> (d - 1 - k)) & 1, ((b >> (d - 1 - k)) & 1) + shp[k] - 1) for k in range(d))
            cols.append(ids[sl].reshape(-1))
        corner_cols.append(np.stack(cols, axis=1))
    cells = np.concatenate(corner_cols) if corner_cols else np.zeros((0, nc), np.int64)
    ncell = cells.shape[0]
    b_idx = np.arange(nc)
    gc = cells.reshape(-1)
    gn = [cells[:, b_idx ^ _bit(d, k)].reshape(-1) for k in range(d)]
    sign = np.array([[-1.0 if (b & _bit(d, k)) else 1.0 for k in range(d)] for b in range(nc)])
    s = [np.tile(sign[:, k], ncell) for k in range(d)]
    # role[a, b]: local corner a in sample b
    role = np.full((nc, nc), -1, np.int32)
    role[b_idx, b_idx] = 0
    for k in range(d):
        role[b_idx ^ _bit(d, k), b_idx] = 1 + k
    sid = nc * np.arange(ncell)[:, None] + b_idx[None, :]
    m_node = np.repeat(cells[:, :, None], nc, axis=2).reshape(-1)
    m_sid = np.repeat(sid[:, None, :], nc, axis=1).reshape(-1)
    m_role = np.tile(role.reshape(-1), ncell)
    return {"gc": gc, "gn": gn, "s": s, "m_node": m_node, "m_sid": m_sid, "m_role": m_role,
            "nc": nc, "ns": gc.shape[0], "ncell": ncell}


# Entity types a constrained boundary DOF can slide on here, in the wire's
# per-type SoA form (reference egg/geometry/entity_soa.py:109-155: tag,
# record fields), 2D: segment (sx, sy, ex, ey), circle (cx, cy, r), ellipse
# (cx, cy, rx, ry), rail (sx, sy, ex, ey; releases the node beyond its ends).
ENTITY_TYPES = {"lineseg": (1, 4), "circle": (2, 3), "ellipse": (3, 4), "linerail": (15, 4)}


def _entities_wire(dofs: np.ndarray, entities: dict | None) -> dict:
    """The group's ``entities`` dict (reference egg/geometry/entity_soa.py:
    598-680 group_entities_by_type): one sub-dict per present type with the
    group-local DOF indices and packed records; free DOFs under "free"."""
    if not entities:
        return {}
    parts: dict[str, dict] = {}
    for local, node in enumerate(dofs.tolist()):
        kind, rec = entities.get(int(node), ("free", ()))
        if kind == "free":
            tag, kf = 0, 0
        elif kind in ENTITY_TYPES:
            tag, kf = ENTITY_TYPES[kind]
        else:
            raise ValueError(f"entity type {kind!r} is not supported ({', '.join(ENTITY_TYPES)})")
        if len(rec) != kf:
            raise ValueError(f"a {kind} record holds {kf} values, got {len(rec)}")
        p = parts.setdefault(kind, {"tag": tag, "kFields": kf, "dof_local": [], "records": []})
        p["dof_local"].append(local)
        p["records"].extend(float(v) for v in rec)
    return {k: {"tag": v["tag"], "kFields": v["kFields"], "dof_local": np.asarray(v["dof_local"], np.int32),
                "records": np.asarray(v["records"], np.float64)} for k, v in parts.items()}


def curvature_table(free_mask: np.ndarray, nodes: np.ndarray, weight: np.ndarray) -> dict:
    """Per-DOF curvature window table (reference
    egg/smoothing/interface_c2.py:230-279): every free node of a window
    (four global node ids, a weight) lists it with its place in the window,
    padded to a common width ``curv_k`` (place -1 = padding), in the group's
    DOF order."""
    free = np.asarray(free_mask, bool)
    dofs = np.flatnonzero(free)
    nodes = np.asarray(nodes, np.int64).reshape(-1, 4)
    weight = np.asarray(weight, np.float64).reshape(-1)
    pos = np.full(free.shape[0], -1, np.int64)
    pos[dofs] = np.arange(dofs.size)
    lists = [[] for _ in range(dofs.size)]
    for wi in range(nodes.shape[0]):
        for slot in range(4):
            n = int(nodes[wi, slot])
            if free[n]:
                lists[int(pos[n])].append((nodes[wi], slot, float(weight[wi])))
    K = max((len(x) for x in lists), default=0)
    if K == 0:
        return {}
    cn = np.full((dofs.size, K, 4), -1, np.int32)
    cs = np.full((dofs.size, K), -1, np.int32)
    cw = np.zeros((dofs.size, K))
    for i, entry in enumerate(lists):
        for j, (row, slot, w) in enumerate(entry):
            cn[i, j], cs[i, j], cw[i, j] = row, slot, w
    return {"curv_k": int(K), "curv_nodes": cn.reshape(-1), "curv_slot": cs.reshape(-1),
            "curv_weight": cw.reshape(-1)}


def build_flat_context(blocks: list[np.ndarray], free_mask: np.ndarray, d: int, *,
                       w_inv: np.ndarray | None = None, weight: np.ndarray | None = None,
                       entities: dict | None = None, curvature: tuple | None = None,
                       directional=None) -> dict:
    """Assemble the flat wire for the moving DOFs ``free_mask`` of ``blocks``.

    ``w_inv`` is one uniform ``(d, d)`` inverse target (identity when
    omitted) or ``(ns, d, d)`` per energy-stencil sample. ``weight`` is an
    optional ``(ns,)`` per-sample energy weight (it scales a sample's mu,
    gradient and Hessian, not its det); as in the reference
    (egg/smoothing/flat_context.py:186-218) a weighted wire ships every DOF's
    real stencil, so interior synthesis is off. ``entities`` maps a moving
    node id to the entity it slides on, ``("lineseg", (sx, sy, ex, ey))`` or
    ``("circle", (cx, cy, r))`` (2D): such a DOF ships its stored stencil and
    the group's ``entities`` field carries the per-type records.
    ``curvature`` = (nodes (W, 4), weight (W,)) adds block-interface C2
    curvature windows (2D; reference egg/smoothing/flat_context.py:293-304).
    ``directional`` (a ``DirectionalSamples``) adds directional soft-energy
    terms: per-DOF windows on the group, the whole-sample table on the energy
    stencil (2D in the compute core; reference flat_context.py:305-320).
    """
    st = cell_stencil(blocks, d)
    ns = st["ns"]
    N = int(free_mask.shape[0])
    dd = d * d
    w_inv = np.eye(d) if w_inv is None else np.asarray(w_inv, np.float64)
    uniform = w_inv.shape == (d, d)
    w0 = w_inv if uniform else (w_inv[0] if len(w_inv) else np.eye(d))
    synth = bool(np.allclose(w0, w0[0, 0] * np.eye(d)) and (uniform or np.allclose(w_inv, w0)))
    if weight is not None:
        weight = np.asarray(weight, np.float64).reshape(-1)
        if weight.shape != (ns,):
            raise ValueError(f"weight must hold one value per energy-stencil sample ({ns}), got {weight.shape}")
        synth = False

    node_block = np.full(N, -1, np.int32)
    node_logical = np.zeros((N, d), np.int32)
    if synth:
        for b, ids in enumerate(blocks):
            ids = np.asarray(ids)
            if any(s < 3 for s in ids.shape):
                continue
            inner = ids[tuple(slice(1, s - 1) for s in ids.shape)].reshape(-1)
            grids = np.meshgrid(*[np.arange(1, s - 1, dtype=np.int32) for s in ids.shape], indexing="ij")
            node_block[inner] = b
            node_logical[inner] = np.stack([g.reshape(-1) for g in grids], axis=1)
    if entities:
        if d != 2:
            raise ValueError("constrained boundary DOFs are 2D only here")
        node_block[np.asarray([k for k, v in entities.items() if v[0] != "free"], np.int64)] = -1
    interior = node_block >= 0

    m_node = st["m_node"]
    keep = free_mask[m_node] & ~interior[m_node]
    order = np.argsort(m_node[keep], kind="stable")
    m_node = m_node[keep][order]
    m_sid = st["m_sid"][keep][order]
    m_role = st["m_role"][keep][order]

    def fields(sid: np.ndarray) -> dict:
        out = {"gc": st["gc"][sid].astype(np.int32)}
        for k in range(d):
            out[f"gn{k}"] = st["gn"][k][sid].astype(np.int32)
            out[f"s{k}"] = st["s"][k][sid].astype(np.float64)
        out["W_inv"] = (np.asarray(w_inv, np.float64).reshape(1, dd) if uniform
                        else np.ascontiguousarray(w_inv[sid].reshape(-1, dd)))
        if weight is not None:
            out["weight"] = np.ascontiguousarray(weight[sid])
        return out

    groups = []
    dofs = np.flatnonzero(free_mask)
    if dofs.size:
        uniq, cnt = np.unique(m_node, return_counts=True)
        P_of = np.zeros(dofs.size, np.int32)
        P_of[np.searchsorted(dofs, uniq)] = cnt
        g = {
            "D": int(dofs.size),
            "role": m_role.astype(np.int32),
            "dof_idx": dofs.astype(np.int32),
            "P_of": P_of,
            "interior_block": np.ascontiguousarray(node_block[dofs]),
            "interior_logical": np.ascontiguousarray(node_logical[dofs]),
            "entities": _entities_wire(dofs, entities),
        }
        g.update(fields(m_sid))
        if curvature is not None:
            if d != 2:
                raise ValueError("curvature windows are 2D only")
            g.update(curvature_table(free_mask, *curvature))
        g.update(directional_dof_table(free_mask, directional))
        groups.append(g)
    es = {"num_samples": int(ns)}
    es.update(fields(np.arange(ns)))
    es.update(directional_energy_table(directional))
    return {"groups": groups, "energy_stencil": es}
