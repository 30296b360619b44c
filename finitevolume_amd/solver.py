"""Python host side over the C ABI (include/fvk.h) of the HIP library.

The classes mirror the reference solver classes' constructors and methods
(include/EulerFVMSolver.h:107-160, include/NavierStokesFVMSolver.h:118-146,
include/RANSTurbulenceSASolver.h:107-153, RANSTurbulenceSSTSolver,
include/UnstructuredFVMSolver.h:34-73, GradientCalculator) so that code and
tests written against the reference read the same.  Every call goes to
libfvk.so on the GPU; there is no CPU fallback: a missing library or device
raises.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional, Sequence

import numpy as np

from . import abi
from .mesh import UnstructuredMesh

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libfvk.so")
_lib = None


class FVKError(RuntimeError):
    pass


def lib():
    """Loads libfvk.so (built by finitevolume_amd.build)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FVKError(f"{LIB_PATH} is not built: run `python -m finitevolume_amd.build` (hipcc, gfx950)")
    L = C.CDLL(LIB_PATH)
    dp, ip, vp, i64 = C.POINTER(C.c_double), C.POINTER(C.c_int32), C.c_void_p, C.c_int64
    L.fvk_create.argtypes = [C.POINTER(abi.MeshDescC), C.POINTER(abi.BCC), C.c_int32, C.POINTER(abi.ParamsC),
                             C.POINTER(abi.InitialConditionC), C.POINTER(vp)]
    L.fvk_destroy.argtypes = [vp]
    L.fvk_destroy.restype = None
    L.fvk_step.argtypes = [vp, C.c_int32, dp, dp]
    L.fvk_residual.argtypes = [vp, dp]
    L.fvk_synchronize.argtypes = [vp]
    L.fvk_set_field.argtypes = [vp, dp, dp, i64]
    L.fvk_get_field.argtypes = [vp, dp, dp, i64]
    L.fvk_set_cfl.argtypes = [vp, C.c_double]
    L.fvk_gas_gamma.argtypes = [vp]
    L.fvk_gas_gamma.restype = C.c_double
    L.fvk_get_launch_shape.argtypes = [vp, C.POINTER(abi.LaunchShapeC)]
    L.fvk_gradient.argtypes = [vp, C.c_int32, dp, dp, dp]
    L.fvk_wall_distance.argtypes = [vp, dp]
    L.fvk_wall_distance_faces.argtypes = [C.POINTER(abi.MeshDescC), ip, i64, dp]
    L.fvk_residual_eval.argtypes = [vp, dp]
    L.fvk_residual_vjp.argtypes = [vp, dp, dp, dp]
    L.fvk_wall_forces.argtypes = [vp, C.c_double, C.c_double, C.c_double, C.c_double, dp, dp]
    L.fvk_wall_forces_vjp.argtypes = [vp, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                                      dp, dp]
    L.fvk_adjoint_solve_forces.argtypes = [vp, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                                           C.c_double, C.c_double, C.c_int32, C.c_double, C.c_int32, dp, dp,
                                           C.POINTER(abi.AdjointReportC)]
    L.fvk_wall_traction.argtypes = [vp, ip, i64, C.POINTER(abi.WallReferenceC), vp]
    L.fvk_wall_force_reports.argtypes = [vp, ip, i64, C.POINTER(abi.WallReferenceC), vp, vp, C.c_int32,
                                         C.POINTER(C.c_int32)]
    L.fvk_next_cfl.argtypes = [C.c_double, C.c_double, dp, C.c_int32, C.c_double, C.c_double, C.c_double,
                               C.c_int32, C.c_double]
    L.fvk_next_cfl.restype = C.c_double
    L.fvk_checkpoint_write.argtypes = [vp, C.c_char_p, i64, C.c_uint64]
    L.fvk_checkpoint_read.argtypes = [vp, C.c_char_p, C.POINTER(i64), C.POINTER(C.c_uint64), ip]
    L.fvk_run.argtypes = [vp, C.POINTER(abi.RunConfigC), C.POINTER(abi.RunResultC)]
    L.fvk_set_kernel_timing.argtypes = [vp, C.c_int32]
    L.fvk_kernel_times.argtypes = [vp, C.c_char_p, i64, dp, C.POINTER(i64), C.c_int32, C.POINTER(C.c_int32)]
    L.fvk_device_bytes.argtypes = [vp]
    L.fvk_device_bytes.restype = i64
    L.fvk_copy_probe.argtypes = [i64, C.c_int32, dp]
    L.fvk_last_error.restype = C.c_char_p
    L.fvk_num_vars.argtypes = [C.c_int32]
    _lib = L
    return L


def _check(rc: int):
    if rc != 0:
        msg = lib().fvk_last_error().decode(errors="replace")
        raise FVKError(f"fvk status {rc}: {msg}")


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class FVKSolver:
    """One fvk_solver handle (any equation)."""

    def __init__(self, mesh: UnstructuredMesh, bcs: Sequence[abi.BCC], params: abi.ParamsC,
                 ic: Optional[abi.InitialConditionC] = None):
        self.L = lib()
        self.mesh = mesh
        self.md = abi.MeshDesc(mesh)
        self.n = mesh.n_cells
        self.eq = params.equation
        self.nv = abi.nvars(self.eq)
        self.params = params
        arr, nb = abi.bc_array(bcs)
        self._bcs = arr
        self.h = C.c_void_p()
        _check(self.L.fvk_create(self.md.ref, arr, nb, C.byref(params), C.byref(ic) if ic is not None else None,
                                 C.byref(self.h)))

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            self.L.fvk_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- reference API ----
    def step(self, nsteps: int = 1, with_dt: bool = False):
        """Advances nsteps steps; returns the per-step residual norms
        [nsteps, nvars] (and the dt per step with with_dt)."""
        hist = np.empty((nsteps, self.nv))
        dts = np.empty(nsteps)
        _check(self.L.fvk_step(self.h, nsteps, abi.as_dptr(hist), abi.as_dptr(dts)))
        return (hist, dts) if with_dt else hist

    def run(self, total_steps: int):
        self.step(total_steps)

    def synchronize(self):
        """Waits for the stages that step() left running (step() returns
        once its residual rows are final)."""
        _check(self.L.fvk_synchronize(self.h))

    def residual(self) -> np.ndarray:
        r = np.empty(self.nv)
        _check(self.L.fvk_residual(self.h, abi.as_dptr(r)))
        return r

    def field(self):
        if self.eq in (abi.EQ_DIFFUSION, abi.EQ_ADVECTION_DIFFUSION):
            U = np.empty(self.n)
            _check(self.L.fvk_get_field(self.h, abi.as_dptr(U), abi.as_dptr(None), self.n))
            return U
        U = np.empty((self.n, 4))
        nt = self.nv - 4
        t = np.empty((self.n, nt)) if nt > 0 else None
        _check(self.L.fvk_get_field(self.h, abi.as_dptr(U), abi.as_dptr(t), self.n))
        return U if t is None else (U, t)

    def set_field(self, U, turb=None):
        U = _f64(U)
        t = None if turb is None else _f64(turb)
        n = U.shape[0]
        _check(self.L.fvk_set_field(self.h, abi.as_dptr(U), abi.as_dptr(t), n))

    def set_cfl(self, cfl: float):
        _check(self.L.fvk_set_cfl(self.h, cfl))

    # ---- driver contract ----
    def checkpoint_write(self, path: str, step_index: int, build_number: int = 0):
        """Checkpoint::write of the current field (src/Checkpoint.cpp)."""
        _check(self.L.fvk_checkpoint_write(self.h, os.fsencode(path), step_index, build_number))

    def checkpoint_read(self, path: str):
        """Loads a checkpoint with the reference driver's resume rules;
        returns (step_index, build_number, source equation tag)."""
        st, bd, eq = C.c_int64(), C.c_uint64(), C.c_int32()
        _check(self.L.fvk_checkpoint_read(self.h, os.fsencode(path), C.byref(st), C.byref(bd), C.byref(eq)))
        return st.value, bd.value, eq.value

    def run_case(self, cfg: abi.RunConfigC) -> abi.RunResultC:
        """The reference driver's step loop (src/main.cpp:3491-3557) on this solver."""
        res = abi.RunResultC()
        _check(self.L.fvk_run(self.h, C.byref(cfg), C.byref(res)))
        return res

    def gas_gamma(self) -> float:
        return self.L.fvk_gas_gamma(self.h)

    def launch_shape(self) -> dict:
        """Kernel variant, workgroup size and (variant 3) tile layout in use."""
        sh = abi.LaunchShapeC()
        _check(self.L.fvk_get_launch_shape(self.h, C.byref(sh)))
        return {f: getattr(sh, f) for f, _ in abi.LaunchShapeC._fields_}

    # ---- extensions ----
    def residual_eval(self) -> np.ndarray:
        R = np.empty((self.n, self.nv))
        _check(self.L.fvk_residual_eval(self.h, abi.as_dptr(R)))
        return R

    def residual_vjp(self, lam, want_x: bool = True):
        lam = _f64(lam)
        Ub = np.empty((self.n, 4))
        Xb = np.empty((self.mesh.n_nodes, 2)) if want_x else None
        _check(self.L.fvk_residual_vjp(self.h, abi.as_dptr(lam), abi.as_dptr(Ub), abi.as_dptr(Xb)))
        return (Ub, Xb) if want_x else Ub

    def gradient(self, scheme: int, phi, boundary_phi) -> np.ndarray:
        g = np.empty((self.n, 2))
        _check(self.L.fvk_gradient(self.h, scheme, abi.as_dptr(_f64(phi)), abi.as_dptr(_f64(boundary_phi)),
                                   abi.as_dptr(g)))
        return g

    def wall_distance(self) -> np.ndarray:
        d = np.empty(self.n)
        _check(self.L.fvk_wall_distance(self.h, abi.as_dptr(d)))
        return d

    def wall_forces(self, rho_ref, vx_ref, vy_ref, length_ref=1.0):
        cl, cd = C.c_double(), C.c_double()
        _check(self.L.fvk_wall_forces(self.h, rho_ref, vx_ref, vy_ref, length_ref, C.byref(cl), C.byref(cd)))
        return cl.value, cd.value

    def wall_traction(self, wall_faces, ref: abi.WallReferenceC) -> np.ndarray:
        """compute_wall_traction_samples (src/NavierStokesFVMSolver.cpp:333-352)
        with the per-sample Cf, Cp and y+; a structured array of
        abi.WALL_SAMPLE_DTYPE."""
        wf = np.ascontiguousarray(wall_faces, dtype=np.int32)
        out = np.zeros(wf.shape[0], dtype=abi.WALL_SAMPLE_DTYPE)
        _check(self.L.fvk_wall_traction(self.h, abi.as_iptr(wf), wf.shape[0], C.byref(ref),
                                        out.ctypes.data_as(C.c_void_p)))
        return out

    def wall_force_reports(self, wall_faces, ref: abi.WallReferenceC, with_samples: bool = False):
        """compute_wall_forces (src/WallTraction.cpp:163-216): per-patch reports
        by ascending patch id, then the total (patch_id -1)."""
        wf = np.ascontiguousarray(wall_faces, dtype=np.int32)
        cap = len(self._bcs) + 1
        rep = np.zeros(cap, dtype=abi.WALL_REPORT_DTYPE)
        smp = np.zeros(wf.shape[0], dtype=abi.WALL_SAMPLE_DTYPE) if with_samples else None
        n = C.c_int32()
        _check(self.L.fvk_wall_force_reports(self.h, abi.as_iptr(wf), wf.shape[0], C.byref(ref),
                                             smp.ctypes.data_as(C.c_void_p) if smp is not None else None,
                                             rep.ctypes.data_as(C.c_void_p), cap, C.byref(n)))
        return (rep[: n.value], smp) if with_samples else rep[: n.value]

    def wall_forces_vjp(self, rho_ref, vx_ref, vy_ref, length_ref=1.0, w_cl=1.0, w_cd=0.0, want_x: bool = True):
        """Gradient of w_cl Cl + w_cd Cd with respect to the state and (optionally) the node coordinates."""
        Ub = np.empty((self.n, 4))
        Xb = np.empty((self.mesh.n_nodes, 2)) if want_x else None
        _check(self.L.fvk_wall_forces_vjp(self.h, rho_ref, vx_ref, vy_ref, length_ref, w_cl, w_cd, abi.as_dptr(Ub),
                                          abi.as_dptr(Xb)))
        return (Ub, Xb) if want_x else Ub

    def adjoint_solve_forces(self, rho_ref, vx_ref, vy_ref, length_ref=1.0, w_cl=1.0, w_cd=0.0, cfl=None,
                             max_iter=20000, tol=1e-10, check_every=20, want_lambda: bool = False):
        """dJ/dX [n_nodes, 2] of J = w_cl Cl + w_cd Cd through the flow, from one
        adjoint solve at the current (converged) state; returns (dJ_dX, report)
        or (dJ_dX, lambda, report).  cfl defaults to the solver's."""
        dX = np.empty((self.mesh.n_nodes, 2))
        lam = np.empty((self.n, 4)) if want_lambda else None
        rep = abi.AdjointReportC()
        c = self.params.cfl if cfl is None else cfl
        _check(self.L.fvk_adjoint_solve_forces(self.h, rho_ref, vx_ref, vy_ref, length_ref, w_cl, w_cd, c, max_iter,
                                               tol, check_every, abi.as_dptr(dX), abi.as_dptr(lam), C.byref(rep)))
        r = {"iterations": rep.iterations, "residual_first": rep.residual_first,
             "residual_final": rep.residual_final}
        return (dX, lam, r) if want_lambda else (dX, r)

    def set_kernel_timing(self, mode):
        """False/0: off; True/1: an event pair around every launch (per-kernel
        times); 2: one pair around each step() call's launches ("stages")."""
        _check(self.L.fvk_set_kernel_timing(self.h, int(mode)))

    def kernel_times(self) -> dict:
        cap = 32
        names = C.create_string_buffer(4096)
        tot = np.zeros(cap)
        cnt = np.zeros(cap, dtype=np.int64)
        n = C.c_int32()
        _check(self.L.fvk_kernel_times(self.h, names, 4096, abi.as_dptr(tot), cnt.ctypes.data_as(C.POINTER(C.c_int64)),
                                       cap, C.byref(n)))
        keys = names.value.decode().split("\n")[: n.value]
        return {k: (float(tot[i]), int(cnt[i])) for i, k in enumerate(keys)}

    def device_bytes(self) -> int:
        return int(self.L.fvk_device_bytes(self.h))


# ---------------------------------------------------------------------------
# reference-shaped classes
# ---------------------------------------------------------------------------

def _ic_from(ic) -> Optional[abi.InitialConditionC]:
    if ic is None or isinstance(ic, abi.InitialConditionC):
        return ic
    raise TypeError("initial_condition must be an abi.InitialConditionC")


class EulerFVMSolver(FVKSolver):
    """EulerFVMSolver(mesh, bcs, gamma, cfl, ic, flux_scheme, exact_riemann_tol,
    exact_riemann_max_iter), include/EulerFVMSolver.h:107-112; extension
    keywords (order, limiter, venkat_k, rk_stages, local_dt) default to the
    reference method."""

    def __init__(self, mesh, boundary_conditions, gamma=1.4, cfl=0.5, initial_condition=None,
                 flux_scheme=abi.FLUX_RUSANOV, exact_riemann_tol=1e-6, exact_riemann_max_iter=20, **ext):
        p = abi.default_params(equation=abi.EQ_EULER, gamma=gamma, cfl=cfl, flux_scheme=flux_scheme,
                               exact_riemann_tol=exact_riemann_tol, exact_riemann_max_iter=exact_riemann_max_iter,
                               **ext)
        super().__init__(mesh, boundary_conditions, p, _ic_from(initial_condition))


class NavierStokesFVMSolver(FVKSolver):
    """include/NavierStokesFVMSolver.h:118-121."""

    def __init__(self, mesh, boundary_conditions, gamma, gas_constant, mu, prandtl, cfl, initial_condition,
                 flux_scheme=abi.FLUX_RUSANOV, gradient_scheme=abi.GRAD_LEAST_SQUARES, exact_riemann_tol=1e-6,
                 exact_riemann_max_iter=20, **ext):
        p = abi.default_params(equation=abi.EQ_NAVIER_STOKES, gamma=gamma, gas_constant=gas_constant, mu=mu,
                               prandtl=prandtl, cfl=cfl, flux_scheme=flux_scheme, gradient_scheme=gradient_scheme,
                               exact_riemann_tol=exact_riemann_tol, exact_riemann_max_iter=exact_riemann_max_iter,
                               **ext)
        super().__init__(mesh, boundary_conditions, p, _ic_from(initial_condition))


class RANSTurbulenceSASolver(FVKSolver):
    """include/RANSTurbulenceSASolver.h:107-111 (initial_nut goes in the
    initial condition's initial_nut)."""

    def __init__(self, mesh, boundary_conditions, gamma, gas_constant, mu, prandtl, prandtl_t, cfl,
                 initial_condition, flux_scheme=abi.FLUX_RUSANOV, gradient_scheme=abi.GRAD_LEAST_SQUARES,
                 exact_riemann_tol=1e-6, exact_riemann_max_iter=20, **ext):
        p = abi.default_params(equation=abi.EQ_RANS_SA, gamma=gamma, gas_constant=gas_constant, mu=mu,
                               prandtl=prandtl, prandtl_t=prandtl_t, cfl=cfl, flux_scheme=flux_scheme,
                               gradient_scheme=gradient_scheme, exact_riemann_tol=exact_riemann_tol,
                               exact_riemann_max_iter=exact_riemann_max_iter, **ext)
        super().__init__(mesh, boundary_conditions, p, _ic_from(initial_condition))

    def nut_field(self):
        return self.field()[1][:, 0]

    def nut_residual(self):
        return self.residual()[4]


class RANSTurbulenceSSTSolver(FVKSolver):
    """src/RANSTurbulenceSSTSolver.cpp:40-46."""

    def __init__(self, mesh, boundary_conditions, gamma, gas_constant, mu, prandtl, prandtl_t, cfl,
                 initial_condition, flux_scheme=abi.FLUX_RUSANOV, gradient_scheme=abi.GRAD_LEAST_SQUARES,
                 exact_riemann_tol=1e-6, exact_riemann_max_iter=20, variant=abi.SST_VORTICITY, kato_launder=False,
                 **ext):
        p = abi.default_params(equation=abi.EQ_RANS_SST, gamma=gamma, gas_constant=gas_constant, mu=mu,
                               prandtl=prandtl, prandtl_t=prandtl_t, cfl=cfl, flux_scheme=flux_scheme,
                               gradient_scheme=gradient_scheme, exact_riemann_tol=exact_riemann_tol,
                               exact_riemann_max_iter=exact_riemann_max_iter, sst_variant=variant,
                               kato_launder=int(kato_launder), **ext)
        super().__init__(mesh, boundary_conditions, p, _ic_from(initial_condition))

    def k_field(self):
        return self.field()[1][:, 0]

    def omega_field(self):
        return self.field()[1][:, 1]


class UnstructuredFVMSolver(FVKSolver):
    """Scalar diffusion, include/UnstructuredFVMSolver.h:34-35; the patches'
    BoundaryType/value come as fvk_bc entries (type 0 Dirichlet, 1 Neumann)."""

    def __init__(self, mesh, patch_bcs, alpha, dt, initial_value, initial_radius):
        p = abi.default_params(equation=abi.EQ_DIFFUSION, alpha=alpha, fixed_dt=dt)
        ic = abi.initial_condition(abi.IC_FREESTREAM, initial_value=initial_value, initial_radius=initial_radius)
        super().__init__(mesh, patch_bcs, p, ic)

    def residual_norm(self) -> float:
        return float(self.residual()[0])


class AdvectionDiffusionFVMSolver(FVKSolver):
    """AdvectionDiffusionFVMSolver(mesh, alpha, u_adv, v_adv, dt,
    gradient_scheme, initial_value, initial_radius),
    src/AdvectionDiffusionFVMSolver.cpp:16-31; the patches' BoundaryType /
    value come as fvk_bc entries (type 0 Dirichlet, 1 Neumann)."""

    def __init__(self, mesh, patch_bcs, alpha, u_adv, v_adv, dt, gradient_scheme=abi.GRAD_LEAST_SQUARES,
                 initial_value=0.0, initial_radius=0.0):
        p = abi.default_params(equation=abi.EQ_ADVECTION_DIFFUSION, alpha=alpha, fixed_dt=dt, advection_u=u_adv,
                               advection_v=v_adv, gradient_scheme=gradient_scheme)
        ic = abi.initial_condition(abi.IC_FREESTREAM, initial_value=initial_value, initial_radius=initial_radius)
        super().__init__(mesh, patch_bcs, p, ic)

    def residual_norm(self) -> float:
        return float(self.residual()[0])


class GradientCalculator:
    """GradientCalculator(mesh, scheme).compute(phi, boundary_phi),
    src/GradientReconstruction.cpp:45-150, on the GPU."""

    def __init__(self, mesh, scheme: int):
        self.scheme = scheme
        bcs = [abi.make_bc(abi.BC_OUTFLOW) for _ in range(max(1, len(mesh.patches),
                                                               int(mesh.face_patch.max()) + 1))]
        self._s = FVKSolver(mesh, bcs, abi.default_params())

    def compute(self, phi, boundary_phi) -> np.ndarray:
        return self._s.gradient(self.scheme, phi, boundary_phi)


def wall_distance(mesh: UnstructuredMesh, wall_faces) -> np.ndarray:
    """compute_wall_distance(mesh, wall_faces), src/WallDistance.cpp:26-44, on the GPU."""
    L = lib()
    md = abi.MeshDesc(mesh)
    wf = np.ascontiguousarray(wall_faces, dtype=np.int32)
    out = np.empty(mesh.n_cells)
    _check(L.fvk_wall_distance_faces(md.ref, abi.as_iptr(wf), wf.shape[0], abi.as_dptr(out)))
    return out


def next_cfl(current, cfl_max, history, cfl_min=0.05, growth=1.5, shrink=0.5, window=15,
             divergence_threshold=2.0) -> float:
    """next_cfl (src/CflRamp.cpp:7-27), the library's host implementation."""
    h = np.ascontiguousarray(history, dtype=np.float64)
    return lib().fvk_next_cfl(current, cfl_max, abi.as_dptr(h), h.shape[0], cfl_min, growth, shrink, window,
                              divergence_threshold)


def copy_probe(n_doubles: int, reps: int = 20) -> float:
    ms = C.c_double()
    _check(lib().fvk_copy_probe(n_doubles, reps, C.byref(ms)))
    return ms.value
