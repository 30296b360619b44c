"""ctypes mirror of the C ABI structs in include/fvk.h.

Shared by the product wrapper (finitevolume_amd.solver) and by test
infrastructure that drives other libraries taking the same structs.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional, Sequence

import numpy as np

# enums (include/fvk.h)
EQ_EULER, EQ_NAVIER_STOKES, EQ_RANS_SA, EQ_RANS_SST, EQ_DIFFUSION, EQ_ADVECTION_DIFFUSION = 0, 1, 2, 3, 4, 5
FLUX_RUSANOV, FLUX_HLLC, FLUX_EXACT = 0, 1, 2
GRAD_GREEN_GAUSS, GRAD_LEAST_SQUARES = 0, 1
BC_WALL, BC_FARFIELD, BC_OUTFLOW = 0, 1, 2
BC_DIRICHLET, BC_NEUMANN = 0, 1
SST_VORTICITY, SST_STRAIN_RATE = 0, 1
IC_FREESTREAM, IC_TWO_REGION, IC_NONE = 0, 1, 2
LIMITER_NONE, LIMITER_VENKATAKRISHNAN, LIMITER_FIRST_ORDER = 0, 1, 2

FLUX_NAMES = {"rusanov": FLUX_RUSANOV, "hllc": FLUX_HLLC, "exact": FLUX_EXACT}

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)


class MeshDescC(C.Structure):
    _fields_ = [
        ("n_nodes", C.c_int64), ("n_cells", C.c_int64), ("n_faces", C.c_int64),
        ("n_patches", C.c_int32), ("reserved0", C.c_int32),
        ("node_xy", _dp), ("face_nodes", _ip), ("face_cells", _ip), ("face_patch", _ip),
        ("cell_face_ptr", _ip), ("cell_faces", _ip), ("cell_volume", _dp), ("cell_xy", _dp),
        ("face_shift", _dp),
    ]


class BCC(C.Structure):
    _fields_ = [
        ("type", C.c_int32), ("is_isothermal_wall", C.c_int32),
        ("wall_u", C.c_double), ("wall_v", C.c_double), ("wall_temperature", C.c_double),
        ("farfield_state", C.c_double * 4),
        ("farfield_nut", C.c_double), ("farfield_k", C.c_double), ("farfield_omega", C.c_double),
        ("value", C.c_double),
    ]


class SAConstantsC(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("cb1", "cb2", "sigma", "kappa", "cw2", "cw3", "cv1", "cv2", "cv3")]


class SSTConstantsC(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("beta_star", "kappa", "a1", "sigma_k1", "sigma_k2", "sigma_omega1",
                                          "sigma_omega2", "beta1", "beta2")]


class ParamsC(C.Structure):
    _fields_ = [
        ("equation", C.c_int32), ("flux_scheme", C.c_int32), ("gradient_scheme", C.c_int32),
        ("exact_riemann_max_iter", C.c_int32),
        ("gamma", C.c_double), ("gas_constant", C.c_double), ("mu", C.c_double), ("prandtl", C.c_double),
        ("prandtl_t", C.c_double), ("cfl", C.c_double), ("exact_riemann_tol", C.c_double),
        ("sa", SAConstantsC), ("sst", SSTConstantsC),
        ("sst_variant", C.c_int32), ("kato_launder", C.c_int32),
        ("alpha", C.c_double), ("fixed_dt", C.c_double),
        ("order", C.c_int32), ("limiter", C.c_int32), ("venkat_k", C.c_double),
        ("rk_stages", C.c_int32), ("local_dt", C.c_int32),
        ("kernel_variant", C.c_int32), ("block_size", C.c_int32), ("tile_cells", C.c_int32),
        ("reserved1", C.c_int32), ("advection_u", C.c_double), ("advection_v", C.c_double),
    ]


class LaunchShapeC(C.Structure):
    _fields_ = [
        ("kernel_variant", C.c_int32), ("block_size", C.c_int32), ("tile_cells", C.c_int32),
        ("n_tiles", C.c_int32), ("tile_max_staged", C.c_int32), ("tile_max_ring2", C.c_int32),
        ("tile_max_faces", C.c_int32), ("persistent_grid", C.c_int32), ("tile_lds_bytes", C.c_int64),
    ]


class InitialConditionC(C.Structure):
    _fields_ = [("mode", C.c_int32), ("reserved0", C.c_int32)] + [
        (n, C.c_double) for n in ("rho", "u", "v", "p", "rho_l", "u_l", "v_l", "p_l", "rho_r", "u_r", "v_r", "p_r",
                                  "x0", "initial_nut", "initial_k", "initial_omega", "initial_value",
                                  "initial_radius")]


class WallReferenceC(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("rho_ref", "velocity_x_ref", "velocity_y_ref", "p_ref", "length_ref",
                                          "moment_reference_x", "moment_reference_y")]


WALL_SAMPLE_FIELDS = ("x_mid", "y_mid", "p", "rho", "effective_viscosity", "tau_wall", "tx", "ty", "y_wall_normal",
                      "cf", "cp", "y_plus")
WALL_REPORT_FIELDS = ("friction_drag", "pressure_drag", "total_drag", "cd_friction", "cd_pressure", "cd_total", "lift",
                      "cl", "moment", "cm")


class WallSampleC(C.Structure):
    _fields_ = [("face_index", C.c_int32), ("patch_id", C.c_int32)] + [(n, C.c_double) for n in WALL_SAMPLE_FIELDS]


class WallReportC(C.Structure):
    _fields_ = [("patch_id", C.c_int32), ("reserved0", C.c_int32)] + [(n, C.c_double) for n in WALL_REPORT_FIELDS]


# numpy views of arrays of the two structs (same layout)
WALL_SAMPLE_DTYPE = np.dtype([("face_index", np.int32), ("patch_id", np.int32)] +
                             [(n, np.float64) for n in WALL_SAMPLE_FIELDS])
WALL_REPORT_DTYPE = np.dtype([("patch_id", np.int32), ("reserved0", np.int32)] +
                             [(n, np.float64) for n in WALL_REPORT_FIELDS])


def wall_reference(rho_ref=1.0, velocity_x_ref=1.0, velocity_y_ref=0.0, p_ref=0.0, length_ref=1.0,
                   moment_reference_x=0.0, moment_reference_y=0.0) -> WallReferenceC:
    """WallReferenceQuantities with the reference's defaults (include/WallTraction.h:99-106)."""
    return WallReferenceC(rho_ref, velocity_x_ref, velocity_y_ref, p_ref, length_ref, moment_reference_x,
                          moment_reference_y)


def collect_wall_faces(mesh, bcs) -> np.ndarray:
    """Boundary faces on wall patches, in face order (collect_wall_faces,
    src/main.cpp:279-312)."""
    fp = mesh.face_patch
    b = np.nonzero(mesh.face_cells[:, 1] < 0)[0]
    return np.ascontiguousarray([f for f in b if bcs[fp[f]].type == BC_WALL], dtype=np.int32)


CKPT_DIFFUSION, CKPT_EULER, CKPT_ADVECTION_DIFFUSION, CKPT_NAVIER_STOKES, CKPT_RANS_SA, CKPT_RANS_SST = range(6)
RUN_COMPLETED, RUN_CONVERGED, RUN_DIVERGED = 0, 1, 2


class RunConfigC(C.Structure):
    _fields_ = [
        ("nsteps", C.c_int64), ("cfl_ramp", C.c_int32), ("ramp_window", C.c_int32), ("cfl", C.c_double),
        ("ramp_cfl_min", C.c_double), ("ramp_growth", C.c_double), ("ramp_shrink", C.c_double),
        ("ramp_divergence_threshold", C.c_double), ("tolerance", C.c_double * 6),
        ("residual_file", C.c_char_p), ("residual_interval", C.c_int32), ("output_precision", C.c_int32),
        ("checkpoint_file", C.c_char_p), ("build_number", C.c_uint64),
    ]


class AdjointReportC(C.Structure):
    """fvk_adjoint_report (include/fvk.h)."""
    _fields_ = [("iterations", C.c_int32), ("residual_first", C.c_double), ("residual_final", C.c_double)]


class RunResultC(C.Structure):
    _fields_ = [
        ("resume_start", C.c_int64), ("last_step", C.c_int64), ("status", C.c_int32),
        ("resumed_equation", C.c_int32), ("residual", C.c_double * 6), ("cfl", C.c_double),
    ]


def run_config(nsteps, cfl=0.5, cfl_ramp=False, tolerance=None, residual_file=None, residual_interval=1,
               output_precision=6, checkpoint_file=None, build_number=0, **ramp) -> RunConfigC:
    """A run configuration with the reference's CaseInput defaults."""
    c = RunConfigC()
    c.nsteps, c.cfl, c.cfl_ramp = nsteps, cfl, int(cfl_ramp)
    c.ramp_window = ramp.get("window", 15)
    c.ramp_cfl_min = ramp.get("cfl_min", 0.05)
    c.ramp_growth = ramp.get("growth", 1.5)
    c.ramp_shrink = ramp.get("shrink", 0.5)
    c.ramp_divergence_threshold = ramp.get("divergence_threshold", 2.0)
    tol = list(tolerance or []) + [-1.0] * 6
    for i in range(6):
        c.tolerance[i] = tol[i]
    c.residual_file = residual_file.encode() if residual_file else None
    c.residual_interval, c.output_precision = residual_interval, output_precision
    c.checkpoint_file = checkpoint_file.encode() if checkpoint_file else None
    c.build_number = build_number
    return c


def default_params(**kw) -> ParamsC:
    """Reference defaults (mirrors fvk_params_default in the library)."""
    p = ParamsC()
    p.equation = EQ_EULER
    p.flux_scheme = FLUX_RUSANOV
    p.gradient_scheme = GRAD_LEAST_SQUARES
    p.exact_riemann_max_iter = 20
    p.gamma = 1.4
    p.gas_constant = 1.0
    p.mu = 0.0
    p.prandtl = 0.72
    p.prandtl_t = 0.9
    p.cfl = 0.5
    p.exact_riemann_tol = 1e-6
    p.sa = SAConstantsC(0.1355, 0.622, 2.0 / 3.0, 0.41, 0.3, 2.0, 7.1, 0.7, 0.9)
    p.sst = SSTConstantsC(0.09, 0.41, 0.31, 0.85, 1.0, 0.5, 0.856, 0.075, 0.0828)
    p.sst_variant = SST_VORTICITY
    p.kato_launder = 0
    p.alpha = 0.0
    p.fixed_dt = 0.0
    p.order = 1
    p.limiter = LIMITER_NONE
    p.venkat_k = 5.0
    p.rk_stages = 1
    p.local_dt = 0
    for k, v in kw.items():
        if k == "flux_scheme" and isinstance(v, str):
            v = FLUX_NAMES[v]
        setattr(p, k, v)
    return p


def initial_condition(mode: int = IC_NONE, **kw) -> InitialConditionC:
    ic = InitialConditionC()
    ic.mode = mode
    ic.rho, ic.u, ic.v, ic.p = 1.0, 0.0, 0.0, 1.0
    ic.rho_l, ic.u_l, ic.v_l, ic.p_l = 1.0, 0.0, 0.0, 1.0
    ic.rho_r, ic.u_r, ic.v_r, ic.p_r = 0.125, 0.0, 0.0, 0.1
    for k, v in kw.items():
        setattr(ic, k, v)
    return ic


def make_bc(type: int, farfield_state: Optional[Sequence[float]] = None, **kw) -> BCC:
    b = BCC()
    b.type = type
    if farfield_state is not None:
        for i in range(4):
            b.farfield_state[i] = float(farfield_state[i])
    for k, v in kw.items():
        setattr(b, k, v)
    return b


def bc_array(bcs: Sequence[BCC]):
    arr = (BCC * max(1, len(bcs)))()
    for i, b in enumerate(bcs):
        arr[i] = b
    return arr, len(bcs)


def _ptr(a: np.ndarray, ct):
    return a.ctypes.data_as(C.POINTER(ct))


class MeshDesc:
    """Keeps contiguous copies of a mesh's arrays alive and exposes the
    fvk_mesh_desc struct pointing at them."""

    def __init__(self, mesh):
        self.node_xy = np.ascontiguousarray(mesh.node_xy, dtype=np.float64)
        self.face_nodes = np.ascontiguousarray(mesh.face_nodes, dtype=np.int32)
        self.face_cells = np.ascontiguousarray(mesh.face_cells, dtype=np.int32)
        self.face_patch = np.ascontiguousarray(mesh.face_patch, dtype=np.int32)
        self.cell_face_ptr = np.ascontiguousarray(mesh.cell_face_ptr, dtype=np.int32)
        self.cell_faces = np.ascontiguousarray(mesh.cell_faces, dtype=np.int32)
        self.cell_volume = np.ascontiguousarray(mesh.cell_volume, dtype=np.float64)
        self.cell_xy = np.ascontiguousarray(mesh.cell_xy, dtype=np.float64)
        self.face_shift = None if mesh.face_shift is None else np.ascontiguousarray(mesh.face_shift, dtype=np.float64)
        d = MeshDescC()
        d.n_nodes = self.node_xy.shape[0]
        d.n_cells = self.cell_volume.shape[0]
        d.n_faces = self.face_nodes.shape[0]
        d.n_patches = max(len(mesh.patches), int(self.face_patch.max()) + 1 if self.face_patch.size else 0)
        d.node_xy = _ptr(self.node_xy, C.c_double)
        d.face_nodes = _ptr(self.face_nodes, C.c_int32)
        d.face_cells = _ptr(self.face_cells, C.c_int32)
        d.face_patch = _ptr(self.face_patch, C.c_int32)
        d.cell_face_ptr = _ptr(self.cell_face_ptr, C.c_int32)
        d.cell_faces = _ptr(self.cell_faces, C.c_int32)
        d.cell_volume = _ptr(self.cell_volume, C.c_double)
        d.cell_xy = _ptr(self.cell_xy, C.c_double)
        d.face_shift = _ptr(self.face_shift, C.c_double) if self.face_shift is not None else _dp()
        self.desc = d

    @property
    def ref(self):
        return C.byref(self.desc)


def nvars(equation: int) -> int:
    return {EQ_EULER: 4, EQ_NAVIER_STOKES: 4, EQ_RANS_SA: 5, EQ_RANS_SST: 6, EQ_DIFFUSION: 1,
            EQ_ADVECTION_DIFFUSION: 1}[equation]


def as_dptr(a: Optional[np.ndarray]):
    if a is None:
        return _dp()
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def as_iptr(a: Optional[np.ndarray]):
    if a is None:
        return _ip()
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(_ip)
