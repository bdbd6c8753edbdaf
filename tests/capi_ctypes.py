"""A plain ctypes binding of the eggmi C ABI (include/eggmi.h), the one a
non-Python host or a maintainer without pybind11 would write: it packs the
flat wire dict and the structured tables into the C structs and drives a
session by its handle. Used by the tests to exercise the C ABI without the
pybind module."""

from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent.parent
LIB_PATH = ROOT / "egg_amd" / "_lib" / "libeggmi.so"

OK, INVALID_ARGUMENT, NO_DEVICE, INTERNAL = 0, 1, 2, 3


class Group(C.Structure):
    _fields_ = [("ndof", C.c_int64), ("dof_idx", C.c_void_p), ("P_of", C.c_void_p),
                ("interior_block", C.c_void_p), ("interior_logical", C.c_void_p),
                ("total_samples", C.c_int64), ("role", C.c_void_p), ("gc", C.c_void_p),
                ("gn", C.c_void_p * 3), ("s", C.c_void_p * 3),
                ("W_inv", C.c_void_p), ("w_rows", C.c_int64), ("weight", C.c_void_p),
                ("entity_tag", C.c_void_p), ("entity_off", C.c_void_p),
                ("entity_data", C.c_void_p), ("entity_len", C.c_int64),
                ("curv_k", C.c_int32), ("curv_nodes", C.c_void_p), ("curv_slot", C.c_void_p),
                ("curv_weight", C.c_void_p),
                ("dir_k", C.c_int32), ("dir_eps", C.c_double), ("dir_nodes", C.c_void_p),
                ("dir_slot", C.c_void_p), ("dir_kind", C.c_void_p), ("dir_ref", C.c_void_p),
                ("dir_axis", C.c_void_p), ("dir_weight", C.c_void_p)]


class Stencil(C.Structure):
    _fields_ = [("num_samples", C.c_int64), ("gc", C.c_void_p), ("gn", C.c_void_p * 3), ("s", C.c_void_p * 3),
                ("W_inv", C.c_void_p), ("w_rows", C.c_int64), ("weight", C.c_void_p),
                ("dir_num", C.c_int64), ("dir_eps", C.c_double), ("dir_nodes", C.c_void_p),
                ("dir_kind", C.c_void_p), ("dir_ref", C.c_void_p), ("dir_axis", C.c_void_p),
                ("dir_weight", C.c_void_p)]


class Structured(C.Structure):
    _fields_ = [("num_blocks", C.c_int32), ("interior_shapes", C.c_void_p),
                ("block_global_dof", C.POINTER(C.c_void_p)), ("n_halo", C.c_int64),
                ("halo_src_block", C.c_void_p), ("halo_src_padded", C.c_void_p),
                ("halo_dst_block", C.c_void_p), ("halo_dst_padded", C.c_void_p),
                ("n_sing", C.c_int64), ("sing_block", C.c_void_p), ("sing_logical", C.c_void_p)]


class Info(C.Structure):
    _fields_ = [("n_slots", C.c_int64), ("n_overflow", C.c_int64), ("n_halo", C.c_int64), ("n_fan", C.c_int64),
                ("n_share", C.c_int64), ("n_tiles", C.c_int64), ("n_stored", C.c_int64), ("n_synth", C.c_int64),
                ("structured_energy", C.c_int32)]


def load():
    lib = C.CDLL(str(LIB_PATH))
    lib.eggmi_last_error.restype = C.c_char_p
    lib.eggmi_ping.restype = C.c_int
    return lib


class Packed:
    """The C structs of one wire; keeps every converted array alive."""

    def __init__(self, wire: dict, structured: dict, d: int):
        self._keep = []
        self.d = d
        self.groups = (Group * max(len(wire["groups"]), 1))()
        for gi, g in enumerate(wire["groups"]):
            cg = self.groups[gi]
            cg.ndof = int(g["D"])
            cg.dof_idx = self._i(g["dof_idx"])
            cg.P_of = self._i(g["P_of"])
            cg.interior_block = self._i(g["interior_block"])
            cg.interior_logical = self._i(g["interior_logical"])
            cg.total_samples = len(g["gc"])
            cg.role = self._i(g["role"])
            cg.gc = self._i(g["gc"])
            for k in range(d):
                cg.gn[k] = self._i(g[f"gn{k}"])
                cg.s[k] = self._f(g[f"s{k}"])
            W = np.asarray(g["W_inv"], np.float64).reshape(-1)
            cg.W_inv = self._f(W)
            cg.w_rows = W.size // (d * d)
            if g.get("weight") is not None:
                cg.weight = self._f(g["weight"])
        self.n_groups = len(wire["groups"])
        es = wire["energy_stencil"]
        st = Stencil()
        st.num_samples = int(es["num_samples"])
        st.gc = self._i(es["gc"])
        for k in range(d):
            st.gn[k] = self._i(es[f"gn{k}"])
            st.s[k] = self._f(es[f"s{k}"])
        W = np.asarray(es["W_inv"], np.float64).reshape(-1)
        st.W_inv = self._f(W)
        st.w_rows = W.size // (d * d)
        if es.get("weight") is not None:
            st.weight = self._f(es["weight"])
        self.stencil = st
        sd = Structured()
        shapes = np.concatenate([np.asarray(s, np.int32).reshape(-1) for s in structured["interior_shapes"]])
        sd.num_blocks = len(structured["interior_shapes"])
        sd.interior_shapes = self._i(shapes)
        ptrs = (C.c_void_p * sd.num_blocks)(*[self._i(b) for b in structured["block_global_dof"]])
        self._keep.append(ptrs)
        sd.block_global_dof = C.cast(ptrs, C.POINTER(C.c_void_p))
        sd.n_halo = len(structured["halo_src_block"])
        sd.halo_src_block = self._i(structured["halo_src_block"])
        sd.halo_src_padded = self._i(structured["halo_src_padded"])
        sd.halo_dst_block = self._i(structured["halo_dst_block"])
        sd.halo_dst_padded = self._i(structured["halo_dst_padded"])
        sd.n_sing = len(structured["sing_block"])
        sd.sing_block = self._i(structured["sing_block"])
        sd.sing_logical = self._i(structured["sing_logical"])
        self.structured = sd

    def _i(self, a):
        a = np.ascontiguousarray(np.asarray(a).reshape(-1), np.int32)
        self._keep.append(a)
        return a.ctypes.data if a.size else None

    def _f(self, a):
        a = np.ascontiguousarray(np.asarray(a).reshape(-1), np.float64)
        self._keep.append(a)
        return a.ctypes.data if a.size else None


class Session:
    """eggmi_session_create -> eggmi_run / eggmi_get_X -> eggmi_session_destroy."""

    def __init__(self, lib, wire: dict, structured: dict, X: np.ndarray, d: int, device: int = 0):
        self.lib = lib
        self.packed = Packed(wire, structured, d)
        self.n_nodes = int(np.shape(X)[0]) if np.ndim(X) == 2 else int(np.size(X) // d)
        self.d = d
        Xf = np.ascontiguousarray(X, np.float64).reshape(-1)
        self.handle = C.c_void_p()
        self.check(lib.eggmi_session_create(d, self.packed.groups, self.packed.n_groups,
                                            C.byref(self.packed.stencil), C.byref(self.packed.structured),
                                            Xf.ctypes.data_as(C.c_void_p), C.c_int64(self.n_nodes), device,
                                            C.byref(self.handle)))

    def check(self, status):
        if status == OK:
            return
        msg = self.lib.eggmi_last_error().decode()
        if status == INVALID_ARGUMENT:
            raise ValueError(msg)
        raise RuntimeError(msg)

    def run(self, n_sweeps: int, phase: int = 0, delta: float = 0.0, omega: float = 0.8, report_every: int = 0):
        cap = (n_sweeps // report_every + 1) if report_every > 0 else 1
        E = np.zeros(cap)
        M = np.zeros(cap)
        n = C.c_int64()
        self.check(self.lib.eggmi_run(self.handle, n_sweeps, phase, C.c_double(delta), C.c_double(omega),
                                      report_every, E.ctypes.data_as(C.c_void_p), M.ctypes.data_as(C.c_void_p),
                                      C.c_int64(cap), C.byref(n)))
        return E[:n.value], M[:n.value]

    def get_X(self) -> np.ndarray:
        X = np.zeros(self.n_nodes * self.d)
        self.check(self.lib.eggmi_get_X(self.handle, X.ctypes.data_as(C.c_void_p), C.c_int64(X.size)))
        return X.reshape(self.n_nodes, self.d)

    def destroy(self):
        if self.handle:
            self.check(self.lib.eggmi_session_destroy(self.handle))
            self.handle = C.c_void_p()
