"""Test infrastructure only (the oracle): Python driver for the reference
solvers built from the reference's own sources (oracle/_ref, see
oracle/Makefile and ref_harness.cpp).  Only tests/, the smoke check and
bench.py's cpu_baseline leg use it, as a checker or a timed baseline."""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional

import numpy as np

from finitevolume_amd import abi

HERE = os.path.dirname(os.path.abspath(__file__))
_libs = {}


def available(flavour: str = "strict") -> bool:
    return os.path.exists(os.path.join(HERE, "_ref", f"libfvref_{flavour}.so"))


def lib(flavour: str = "strict"):
    if flavour in _libs:
        return _libs[flavour]
    path = os.path.join(HERE, "_ref", f"libfvref_{flavour}.so")
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path} not built (make -C oracle ref, needs the reference sources)")
    L = C.CDLL(path)
    dp, ip, vp = C.POINTER(C.c_double), C.POINTER(C.c_int32), C.c_void_p
    L.fvr_create.argtypes = [C.POINTER(abi.MeshDescC), C.POINTER(abi.BCC), C.c_int32, C.POINTER(abi.ParamsC),
                             C.POINTER(abi.InitialConditionC), C.POINTER(vp)]
    L.fvr_destroy.argtypes = [vp]
    L.fvr_set_field.argtypes = [vp, dp, dp]
    L.fvr_get_field.argtypes = [vp, dp, dp]
    L.fvr_step.argtypes = [vp, C.c_int32, dp]
    L.fvr_set_cfl.argtypes = [vp, C.c_double]
    L.fvr_flux.argtypes = [C.c_int32, C.c_int64, dp, dp, dp, C.c_double, C.c_double, C.c_int32, dp]
    L.fvr_star_pressure.argtypes = [C.c_int64, dp, C.c_double, C.c_double, C.c_int32, dp, dp]
    L.fvr_geometry.argtypes = [C.POINTER(abi.MeshDescC), dp, dp, dp]
    L.fvr_gradient.argtypes = [C.POINTER(abi.MeshDescC), C.c_int32, dp, dp, dp]
    L.fvr_face_gradient.argtypes = [C.POINTER(abi.MeshDescC), dp, dp, dp, dp, dp, dp]
    L.fvr_wall_distance.argtypes = [C.POINTER(abi.MeshDescC), ip, C.c_int64, dp]
    L.fvr_sa_source.argtypes = [C.c_int64, dp, dp, dp, dp, dp, C.POINTER(abi.SAConstantsC), dp, dp]
    L.fvr_sst_source.argtypes = [C.c_int64, dp, dp, dp, dp, dp, dp, dp, dp, dp, C.c_int32,
                                 C.POINTER(abi.SSTConstantsC), C.c_int32, dp]
    L.fvr_next_cfl.argtypes = [C.c_double, C.c_double, dp, C.c_int32, C.c_double, C.c_double, C.c_double,
                               C.c_int32, C.c_double]
    L.fvr_next_cfl.restype = C.c_double
    L.fvr_checkpoint_write.argtypes = [vp, C.c_char_p, C.c_int64, C.c_uint64]
    L.fvr_checkpoint_read.argtypes = [C.c_char_p, C.c_int32, C.c_int64, C.POINTER(C.c_int64), C.POINTER(C.c_uint64),
                                      dp, dp]
    L.fvr_mesh_read.argtypes = [C.c_char_p, C.POINTER(vp), C.POINTER(C.c_int64)]
    L.fvr_mesh_arrays.argtypes = [vp, dp, ip, ip, ip, ip, ip, ip, dp, dp, C.c_char_p, C.c_int64]
    L.fvr_mesh_free.argtypes = [vp]
    L.fvr_mesh_free.restype = None
    L.fvr_wall_reports.argtypes = [vp, ip, C.c_int64, C.POINTER(abi.WallReferenceC), C.c_int32, vp, vp, C.c_int32,
                                   C.POINTER(C.c_int32)]
    _libs[flavour] = L
    return L


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class RefSolver:
    """One reference solver object (EulerFVMSolver, NavierStokesFVMSolver,
    RANSTurbulenceSASolver, RANSTurbulenceSSTSolver or UnstructuredFVMSolver)."""

    def __init__(self, mesh, bcs, params: abi.ParamsC, ic: Optional[abi.InitialConditionC] = None,
                 flavour: str = "strict"):
        self.L = lib(flavour)
        self.md = abi.MeshDesc(mesh)
        self.n = mesh.n_cells
        self.eq = params.equation
        self.nv = abi.nvars(self.eq)
        arr, nb = abi.bc_array(bcs)
        self._bcs = arr
        self.h = C.c_void_p()
        rc = self.L.fvr_create(self.md.ref, arr, nb, C.byref(params), C.byref(ic) if ic is not None else None,
                               C.byref(self.h))
        if rc != 0:
            raise RuntimeError(f"fvr_create failed: {rc}")

    def __del__(self):
        if getattr(self, "h", None) and self.h.value:
            self.L.fvr_destroy(self.h)
            self.h = C.c_void_p()

    def set_field(self, U, turb=None):
        U = _f64(U)
        t = None if turb is None else _f64(turb)
        self.L.fvr_set_field(self.h, abi.as_dptr(U), abi.as_dptr(t))

    def field(self):
        if self.eq in (abi.EQ_DIFFUSION, abi.EQ_ADVECTION_DIFFUSION):
            U = np.empty(self.n)
            self.L.fvr_get_field(self.h, abi.as_dptr(U), abi.as_dptr(None))
            return U
        U = np.empty((self.n, 4))
        nt = self.nv - 4
        t = np.empty((self.n, nt)) if nt > 0 else None
        self.L.fvr_get_field(self.h, abi.as_dptr(U), abi.as_dptr(t))
        return U if t is None else (U, t)

    def step(self, nsteps: int = 1) -> np.ndarray:
        hist = np.empty((nsteps, self.nv))
        self.L.fvr_step(self.h, nsteps, abi.as_dptr(hist))
        return hist

    def set_cfl(self, cfl: float):
        self.L.fvr_set_cfl(self.h, cfl)

    def checkpoint_write(self, path, step, build=0):
        """Checkpoint::write with the solver's own equation tag."""
        rc = self.L.fvr_checkpoint_write(self.h, os.fsencode(path), step, build)
        if rc != 0:
            raise RuntimeError(f"Checkpoint::write failed for {path}")

    def wall_reports(self, wall_faces, ref, pressure_only: bool = False):
        """Wall samples and force reports (structured arrays of
        abi.WALL_SAMPLE_DTYPE and abi.WALL_REPORT_DTYPE)."""
        wf = np.ascontiguousarray(wall_faces, dtype=np.int32)
        smp = np.zeros(wf.shape[0], dtype=abi.WALL_SAMPLE_DTYPE)
        cap = len(self._bcs) + 1
        rep = np.zeros(cap, dtype=abi.WALL_REPORT_DTYPE)
        n = C.c_int32()
        rc = self.L.fvr_wall_reports(self.h, abi.as_iptr(wf), wf.shape[0], C.byref(ref), int(pressure_only),
                         smp.ctypes.data_as(C.c_void_p), rep.ctypes.data_as(C.c_void_p), cap, C.byref(n))
        if rc != 0:
            raise RuntimeError(f"fvr_wall_reports failed: {rc}")
        return smp, rep[: n.value]


def flux(scheme: int, UL, UR, nxy, gamma=1.4, tol=1e-6, max_iter=20, flavour="strict"):
    L = lib(flavour)
    UL, UR, nxy = _f64(UL), _f64(UR), _f64(nxy)
    n = UL.shape[0]
    out = np.empty((n, 4))
    L.fvr_flux(scheme, n, abi.as_dptr(UL), abi.as_dptr(UR), abi.as_dptr(nxy), gamma, tol, max_iter, abi.as_dptr(out))
    return out


def star_pressure(prim, gamma=1.4, tol=1e-6, max_iter=20):
    L = lib()
    prim = _f64(prim)
    n = prim.shape[0]
    ps, us = np.empty(n), np.empty(n)
    L.fvr_star_pressure(n, abi.as_dptr(prim), gamma, tol, max_iter, abi.as_dptr(ps), abi.as_dptr(us))
    return ps, us


def geometry(mesh):
    L = lib()
    md = abi.MeshDesc(mesh)
    mid, area, nxy = np.empty((mesh.n_faces, 2)), np.empty(mesh.n_faces), np.empty((mesh.n_faces, 2))
    L.fvr_geometry(md.ref, abi.as_dptr(mid), abi.as_dptr(area), abi.as_dptr(nxy))
    return mid, area, nxy


def gradient(mesh, scheme, phi, bphi):
    L = lib()
    md = abi.MeshDesc(mesh)
    g = np.empty((mesh.n_cells, 2))
    L.fvr_gradient(md.ref, scheme, abi.as_dptr(_f64(phi)), abi.as_dptr(_f64(bphi)), abi.as_dptr(g))
    return g


def face_gradient(mesh, phi, bphi, grad):
    L = lib()
    md = abi.MeshDesc(mesh)
    nf = mesh.n_faces
    dn, gv, nd = np.empty(nf), np.empty((nf, 2)), np.empty(nf)
    L.fvr_face_gradient(md.ref, abi.as_dptr(_f64(phi)), abi.as_dptr(_f64(bphi)), abi.as_dptr(_f64(grad)),
                        abi.as_dptr(dn), abi.as_dptr(gv), abi.as_dptr(nd))
    return dn, gv, nd


def wall_distance(mesh, wall_faces):
    L = lib()
    md = abi.MeshDesc(mesh)
    wf = np.ascontiguousarray(wall_faces, dtype=np.int32)
    out = np.empty(mesh.n_cells)
    L.fvr_wall_distance(md.ref, abi.as_iptr(wf), wf.shape[0], abi.as_dptr(out))
    return out


def sa_source(nut, nu, omega, dist, grad_nut, consts=None):
    L = lib()
    c = consts if consts is not None else abi.default_params().sa
    nut, nu, omega, dist, grad_nut = map(_f64, (nut, nu, omega, dist, grad_nut))
    n = nut.shape[0]
    out, eddy = np.empty((n, 3)), np.empty(n)
    L.fvr_sa_source(n, abi.as_dptr(nut), abi.as_dptr(nu), abi.as_dptr(omega), abi.as_dptr(dist),
                    abi.as_dptr(grad_nut), C.byref(c), abi.as_dptr(out), abi.as_dptr(eddy))
    return out, eddy


def sst_source(rho, k, omega, nu, dist, gu, gv, gk, gw, variant=0, consts=None, kato=0):
    L = lib()
    c = consts if consts is not None else abi.default_params().sst
    arrs = list(map(_f64, (rho, k, omega, nu, dist, gu, gv, gk, gw)))
    n = arrs[0].shape[0]
    out = np.empty((n, 10))
    L.fvr_sst_source(n, *[abi.as_dptr(a) for a in arrs], variant, C.byref(c), kato, abi.as_dptr(out))
    return out


def next_cfl(current, cfl_max, hist, cfl_min=0.05, growth=1.5, shrink=0.5, window=15, divergence_threshold=2.0):
    L = lib()
    h = _f64(hist)
    return L.fvr_next_cfl(current, cfl_max, abi.as_dptr(h), h.shape[0], cfl_min, growth, shrink, window,
                          divergence_threshold)


def checkpoint_read(path, tag, n_cells):
    """Checkpoint::read of a file with equation tag `tag`: (step, build, U, turb)."""
    L = lib()
    nt = {4: 1, 5: 2}.get(tag, 0)
    U = np.empty(n_cells if tag in (0, 2) else (n_cells, 4))
    T = np.empty((n_cells, nt)) if nt else None
    st, bd = C.c_int64(), C.c_uint64()
    rc = L.fvr_checkpoint_read(os.fsencode(path), tag, n_cells, C.byref(st), C.byref(bd), abi.as_dptr(U),
                               abi.as_dptr(T))
    if rc != 0:
        raise RuntimeError(f"Checkpoint::read failed for {path}")
    return st.value, bd.value, U, T


def read_mesh(path):
    """MeshReader::read of the reference: a dict of its mesh arrays, or
    None when the reference rejects the file."""
    L = lib()
    h = C.c_void_p()
    cnt = (C.c_int64 * 5)()
    if L.fvr_mesh_read(os.fsencode(path), C.byref(h), cnt) != 0:
        return None
    nn, nc, nf, npch, nnz = [int(x) for x in cnt]
    out = dict(node_xy=np.empty((nn, 2)), face_nodes=np.empty((nf, 2), np.int32),
               face_cells=np.empty((nf, 2), np.int32), face_patch=np.empty(nf, np.int32),
               cell_face_ptr=np.empty(nc + 1, np.int32), cell_faces=np.empty(nnz, np.int32),
               cell_nodes=np.empty(nnz, np.int32), cell_volume=np.empty(nc), cell_xy=np.empty((nc, 2)))
    names = C.create_string_buffer(1 << 16)
    try:
        rc = L.fvr_mesh_arrays(h, abi.as_dptr(out["node_xy"]), abi.as_iptr(out["face_nodes"]),
                               abi.as_iptr(out["face_cells"]), abi.as_iptr(out["face_patch"]),
                               abi.as_iptr(out["cell_face_ptr"]), abi.as_iptr(out["cell_faces"]),
                               abi.as_iptr(out["cell_nodes"]), abi.as_dptr(out["cell_volume"]),
                               abi.as_dptr(out["cell_xy"]), names, len(names))
    finally:
        L.fvr_mesh_free(h)
    assert rc == 0
    out["patches"] = names.value.decode().split("\n")[:npch]
    return out
