"""fvk_wall_force_reports_noalloc (no allocation: device buffers in and out,
a fixed device scratch array) against fvk_wall_force_reports on the same
solver state, bit for bit, on two calls in a row, the second with fewer
wall faces; and its size limit."""
import ctypes as C

import numpy as np
import pytest

import cases
from finitevolume_amd import abi
from finitevolume_amd.solver import FVKSolver, lib

pytestmark = pytest.mark.gpu

NOALLOC_MAX = 65536


class _Hip:
    """hipMalloc / hipMemcpy / hipFree of the HIP runtime the library links,
    for the caller-side device buffers of the test."""

    def __init__(self):
        self.L = C.CDLL("libamdhip64.so.7")
        self.L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.L.hipFree.argtypes = [C.c_void_p]
        self.L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]

    def alloc(self, nbytes):
        p = C.c_void_p()
        assert self.L.hipMalloc(C.byref(p), max(int(nbytes), 8)) == 0
        return p

    def put(self, dst, arr):
        arr = np.ascontiguousarray(arr)
        assert self.L.hipMemcpy(dst, arr.ctypes.data_as(C.c_void_p), arr.nbytes, 1) == 0  # host to device

    def get(self, src, nbytes):
        out = np.empty(int(nbytes), dtype=np.uint8)
        assert self.L.hipMemcpy(out.ctypes.data_as(C.c_void_p), src, out.nbytes, 2) == 0  # device to host
        return out

    def free(self, *ps):
        for p in ps:
            self.L.hipFree(p)


def _noalloc(s, wf, ref):
    L = lib()
    vp = C.c_void_p
    L.fvk_wall_force_reports_noalloc.argtypes = [vp, vp, C.c_int64, C.POINTER(abi.WallReferenceC), vp, vp,
                                                 C.c_int32, vp]
    hip = _Hip()
    wf = np.ascontiguousarray(wf, dtype=np.int32)
    cap = len(s._bcs) + 1
    nb_s = max(len(wf), 1) * abi.WALL_SAMPLE_DTYPE.itemsize
    nb_r = cap * abi.WALL_REPORT_DTYPE.itemsize
    d_wf, d_smp, d_rep, d_n = hip.alloc(wf.nbytes), hip.alloc(nb_s), hip.alloc(nb_r), hip.alloc(4)
    try:
        hip.put(d_wf, wf)
        hip.put(d_smp, np.full(nb_s, 0xA5, dtype=np.uint8))
        hip.put(d_rep, np.full(nb_r, 0x5A, dtype=np.uint8))
        hip.put(d_n, np.array([-7], dtype=np.int32))
        rc = L.fvk_wall_force_reports_noalloc(s.h, d_wf, len(wf), C.byref(ref), d_smp, d_rep, cap, d_n)
        assert rc == 0, lib().fvk_last_error()
        s.synchronize()
        n = int(hip.get(d_n, 4).view(np.int32)[0])
        smp = hip.get(d_smp, nb_s).view(abi.WALL_SAMPLE_DTYPE)[: len(wf)]
        rep = hip.get(d_rep, nb_r).view(abi.WALL_REPORT_DTYPE)[:n]
    finally:
        hip.free(d_wf, d_smp, d_rep, d_n)
    return smp, rep


def _same(a, b):
    assert len(a) == len(b)
    for k in a.dtype.names:
        if k == "reserved0":
            continue
        assert np.array_equal(a[k], b[k], equal_nan=True), k


@pytest.mark.parametrize("equation", [abi.EQ_EULER, abi.EQ_NAVIER_STOKES, abi.EQ_RANS_SA])
def test_noalloc_matches_allocating_entry_point(equation):
    if equation == abi.EQ_EULER:
        m, bcs, p, ic, _ = cases.case_naca(64, 32, order=2, limiter=abi.LIMITER_VENKATAKRISHNAN, rk_stages=3,
                                           local_dt=1)
    elif equation == abi.EQ_NAVIER_STOKES:
        m, bcs, p, ic, _ = cases.case_ns_wavy(n=24)
    else:
        m, bcs, p, ic, _ = cases.case_rans_wavy(abi.EQ_RANS_SA, n=24)
    s = FVKSolver(m, bcs, p, ic)
    s.step(7)  # leaves later RK stages in flight: the call is ordered after them on the solver's stream
    wf = abi.collect_wall_faces(m, bcs)
    ref = abi.wall_reference(1.0, 0.3, 0.1, 0.7, 1.0, 0.25, 0.0)
    for call, faces in enumerate((wf, wf[: max(1, len(wf) // 3)])):
        smp, rep = _noalloc(s, faces, ref)
        rep_a, smp_a = s.wall_force_reports(faces, ref, with_samples=True)
        _same(smp, smp_a)
        _same(rep, rep_a)
        assert rep[-1]["patch_id"] == -1, call
        s.step(2)


def test_noalloc_rejects_oversize_input():
    m, bcs, p, ic, _ = cases.case_naca(32, 16)
    s = FVKSolver(m, bcs, p, ic)
    L = lib()
    vp = C.c_void_p
    L.fvk_wall_force_reports_noalloc.argtypes = [vp, vp, C.c_int64, C.POINTER(abi.WallReferenceC), vp, vp,
                                                 C.c_int32, vp]
    ref = abi.wall_reference()
    # rejected on the host before any device pointer is touched
    rc = L.fvk_wall_force_reports_noalloc(s.h, C.c_void_p(16), NOALLOC_MAX + 1, C.byref(ref), None,
                                          C.c_void_p(16), 8, C.c_void_p(16))
    assert rc == 1  # FVK_ERR_INVALID
    assert b"scratch capacity" in L.fvk_last_error()
