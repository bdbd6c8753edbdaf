"""The allocation-free PCG entry point (eggmi_pcg_csr_noalloc) through the
C ABI: operands in device memory, work area in the library's file-scope
device array. Its results must be those of the allocating entry point
(eggmi_pcg_csr, same kernels) bit for bit, on two calls in a row with the
second one smaller, and a size beyond the work area is refused."""

from __future__ import annotations

import ctypes as C

import numpy as np
import pytest


def _lib():
    from tests import capi_ctypes as CA

    lib = CA.load()
    P, I, D = C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_double)
    lib.eggmi_pcg_csr.argtypes = [C.c_int64, P, P, P, P, P, C.c_double, C.c_int, C.c_int, I, D, I, D]
    lib.eggmi_pcg_csr_noalloc.argtypes = [C.c_int64, P, P, P, P, P, C.c_double, C.c_int, C.c_int, P, I, D, I]
    return lib


def _check(lib, status):
    assert status == 0, lib.eggmi_last_error().decode()


class _Hip:
    """Device buffers from the HIP runtime the library itself links
    (/opt/rocm/lib/libamdhip64.so.7), so one runtime serves the process."""

    def __init__(self):
        self.rt = C.CDLL("/opt/rocm/lib/libamdhip64.so.7")
        self.rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.rt.hipFree.argtypes = [C.c_void_p]
        self.rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.rt.hipDeviceSynchronize.argtypes = []
        self.bufs = []

    def upload(self, a: np.ndarray) -> int:
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        assert self.rt.hipMalloc(C.byref(p), max(a.nbytes, 8)) == 0
        self.bufs.append(p)
        assert self.rt.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0  # host to device
        return p.value

    def download(self, ptr: int, like: np.ndarray) -> np.ndarray:
        out = np.empty_like(like)
        assert self.rt.hipMemcpy(out.ctypes.data, ptr, out.nbytes, 2) == 0  # device to host
        return out

    def free(self):
        assert self.rt.hipDeviceSynchronize() == 0
        for p in self.bufs:
            self.rt.hipFree(p)
        self.bufs = []


@pytest.mark.gpu
@pytest.mark.parametrize("rec", [0, 1])
def test_pcg_noalloc_matches_allocating_entry_point(rec):
    from egg_amd.pcg import lattice_laplacian

    lib = _lib()
    hip = _Hip()
    try:
        for n, d, seed in ((33, 3, 1), (12, 3, 2)):  # the second call is smaller
            print(f"no-alloc PCG: recurrence {rec}, {n}^{d} lattice Laplacian", flush=True)
            rp, col, val = lattice_laplacian(n, d)
            N = n**d
            b = np.random.default_rng(seed).normal(size=N)
            x_ref = np.zeros(N)
            it, rr, st, ms = C.c_int(), C.c_double(), C.c_int(), C.c_double()
            _check(lib, lib.eggmi_pcg_csr(N, rp.ctypes.data, col.ctypes.data, val.ctypes.data, b.ctypes.data,
                                          x_ref.ctypes.data, 1e-10, 1000, rec, C.byref(it), C.byref(rr),
                                          C.byref(st), C.byref(ms)))
            d_rp, d_col, d_val, d_b = (hip.upload(a) for a in (rp, col, val, b))
            nan = np.full(N, np.nan)
            d_x = hip.upload(nan)
            it2, rr2, st2 = C.c_int(), C.c_double(), C.c_int()
            _check(lib, lib.eggmi_pcg_csr_noalloc(N, d_rp, d_col, d_val, d_b, d_x, 1e-10, 1000, rec, None,
                                                  C.byref(it2), C.byref(rr2), C.byref(st2)))
            x = hip.download(d_x, nan)
            assert st.value == 1 and (it2.value, st2.value) == (it.value, st.value)
            assert rr2.value == rr.value
            np.testing.assert_array_equal(x, x_ref)
    finally:
        hip.free()


@pytest.mark.gpu
def test_pcg_noalloc_refuses_beyond_its_work_area():
    lib = _lib()
    it, rr, st = C.c_int(), C.c_double(), C.c_int()
    status = lib.eggmi_pcg_csr_noalloc((1 << 18) + 1, None, None, None, None, None, 1e-8, 10, 0, None,
                                       C.byref(it), C.byref(rr), C.byref(st))
    assert status == 1 and "work area" in lib.eggmi_last_error().decode()
