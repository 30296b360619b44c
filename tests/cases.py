"""Standard test cases shared by fixture generation (tests/golden), the CPU
tests of the oracle and the GPU parity tests: each returns the mesh, the
boundary conditions, the parameters and the initial condition, with the
values the reference's own verification gates use where there is one
(src/main.cpp gates cited per case)."""
from __future__ import annotations

import math

import numpy as np

from finitevolume_amd import abi
from finitevolume_amd import mesh as M

GAMMA = 1.4


def freestream(mach, aoa_deg, gamma=GAMMA, rho=1.0, p=None):
    """rho_inf = 1, p_inf = 1/gamma (c_inf = 1) unless p is given."""
    p = 1.0 / gamma if p is None else p
    c = math.sqrt(gamma * p / rho)
    a = math.radians(aoa_deg)
    u, v = mach * c * math.cos(a), mach * c * math.sin(a)
    return rho, u, v, p


def case_1r(n=64, flux=abi.FLUX_RUSANOV, **ext):
    """Config 1 (reference mode 1R): 64x64 periodic unit square, isentropic
    vortex (strength 5, radius 0.1) on a Mach 0.5 stream at 45 degrees,
    first order, global CFL 0.5."""
    m = M.periodic_box(n, n)
    U0 = M.isentropic_vortex(m.cell_xy)
    p = abi.default_params(equation=abi.EQ_EULER, flux_scheme=flux, cfl=0.5, **ext)
    return m, [], p, None, U0


def hub_mesh(n=12, lo=4, hi=8):
    """n x n quads of side 1/4 on [0, n/4]^2 with the centre block [lo, hi)^2
    merged into one cell whose 4 (hi - lo) edges each face a small quad (16
    faces at the defaults: the widest cells the kernels take).  Patches:
    0 bottom, 1 right, 2 top, 3 left."""
    h = 0.25

    def nid(i, j):
        return j * (n + 1) + i
    I, J = np.meshgrid(np.arange(n + 1), np.arange(n + 1))
    node_xy = np.stack([I.ravel() * h, J.ravel() * h], axis=1).astype(np.float64)
    cells = [[nid(i, j), nid(i + 1, j), nid(i + 1, j + 1), nid(i, j + 1)]
             for j in range(n) for i in range(n) if not (lo <= i < hi and lo <= j < hi)]
    hub = ([nid(i, lo) for i in range(lo, hi)] + [nid(hi, j) for j in range(lo, hi)] +
           [nid(i, hi) for i in range(hi, lo, -1)] + [nid(lo, j) for j in range(hi, lo, -1)])
    cells.insert(len(cells) // 2, hub)
    ptr = np.cumsum([0] + [len(c) for c in cells])
    i = np.arange(n)
    edges = [np.stack([nid(i, 0), nid(i + 1, 0)], 1), np.stack([nid(n, i), nid(n, i + 1)], 1),
             np.stack([nid(i + 1, n), nid(i, n)], 1), np.stack([nid(0, i + 1), nid(0, i)], 1)]
    patches = np.concatenate([np.full(e.shape[0], k) for k, e in enumerate(edges)])
    return M.from_polygons(node_xy, ptr, np.concatenate(cells), np.concatenate(edges), patches,
                           ["bottom", "right", "top", "left"])


def case_euler_hub(flux=abi.FLUX_HLLC, **ext):
    """Second-order Euler (LSQ + Venkatakrishnan, SSP-RK3, local dt) on the
    hub mesh: slip wall below, farfield elsewhere, an isentropic vortex on a
    Mach 0.5 stream passing over the sixteen-face cell."""
    m = hub_mesh()
    U0 = M.isentropic_vortex(m.cell_xy, radius=0.4, center=(1.0, 1.2))
    ff = M.from_primitive(1.0, 0.5 * math.sqrt(1.4) * math.cos(math.radians(45.0)),
                          0.5 * math.sqrt(1.4) * math.sin(math.radians(45.0)), 1.0, GAMMA)
    bcs = [abi.make_bc(abi.BC_WALL)] + [abi.make_bc(abi.BC_FARFIELD, ff) for _ in range(3)]
    kw = dict(order=2, limiter=abi.LIMITER_VENKATAKRISHNAN, venkat_k=5.0, rk_stages=3, local_dt=1)
    kw.update(ext)
    p = abi.default_params(equation=abi.EQ_EULER, flux_scheme=flux, cfl=0.6, **kw)
    return m, bcs, p, None, U0


def case_sod(flux=abi.FLUX_HLLC, nx=64, ny=8):
    """Sod tube between slip walls (euler_init tworegion), x0 = 0.5."""
    m = M.cartesian(nx, ny)
    bcs = [abi.make_bc(abi.BC_WALL) for _ in range(4)]
    p = abi.default_params(equation=abi.EQ_EULER, flux_scheme=flux, cfl=0.5)
    ic = abi.initial_condition(abi.IC_TWO_REGION, x0=0.5)
    return m, bcs, p, ic, None


def case_naca(ni=128, nj=64, flux=abi.FLUX_HLLC, mach=0.8, aoa=1.25, cfl=0.8, code="0012", first=None, **ext):
    """Config 2 geometry at reduced size: NACA O-grid (slip wall + farfield),
    freestream Mach 0.8, AoA 1.25 deg.  Reference mode 2R is the default
    (first order, forward Euler, global dt)."""
    m = M.naca_ogrid(ni, nj, code=code, first=first)
    rho, u, v, pinf = freestream(mach, aoa)
    ff = M.from_primitive(rho, u, v, pinf, GAMMA)
    bcs = [abi.make_bc(abi.BC_WALL), abi.make_bc(abi.BC_FARFIELD, ff)]
    p = abi.default_params(equation=abi.EQ_EULER, flux_scheme=flux, cfl=cfl, **ext)
    ic = abi.initial_condition(abi.IC_FREESTREAM, rho=rho, u=u, v=v, p=pinf)
    return m, bcs, p, ic, None


def case_ns_wavy(gradient=abi.GRAD_LEAST_SQUARES, flux=abi.FLUX_HLLC, n=16, **ext):
    """Navier-Stokes on the wavy mesh of the gradient gate (src/main.cpp:472-478):
    isothermal no-slip bottom, farfield left/top, outflow right."""
    m = M.skewed(n, 0.15)
    ff = M.from_primitive(1.0, 0.3, 0.1, 1.0, GAMMA)
    bcs = [abi.make_bc(abi.BC_WALL, is_isothermal_wall=1, wall_temperature=1.2), abi.make_bc(abi.BC_FARFIELD, ff),
           abi.make_bc(abi.BC_FARFIELD, ff), abi.make_bc(abi.BC_OUTFLOW)]
    for b in bcs:
        b.farfield_nut, b.farfield_k, b.farfield_omega = 0.05, 2e-3, 3.0
    p = abi.default_params(equation=abi.EQ_NAVIER_STOKES, flux_scheme=flux, gradient_scheme=gradient, mu=0.01,
                           cfl=0.4, **ext)
    ic = abi.initial_condition(abi.IC_FREESTREAM, rho=1.0, u=0.3, v=0.1, p=1.0, initial_nut=0.06, initial_k=1e-3,
                               initial_omega=5.0)
    return m, bcs, p, ic, None


def case_rans_wavy(equation, gradient=abi.GRAD_LEAST_SQUARES, variant=0, kato=0, n=16, **ext):
    m, bcs, p, ic, _ = case_ns_wavy(gradient=gradient, n=n, **ext)
    p.equation = equation
    p.sst_variant = variant
    p.kato_launder = kato
    return m, bcs, p, ic, None


def case_couette(equation=abi.EQ_NAVIER_STOKES, n_x=1, n_y=16):
    """--verify-couette / --verify-rans-stability setup (src/main.cpp:1701-1730,
    2802-2830): 1x16 sheared mesh (shear 0.3), stationary bottom wall, top wall
    at U = 0.1, outflow sides, mu 0.02, Rusanov, LSQ, cfl 0.3."""
    m = M.sheared(n_x, n_y, 0.3)
    bcs = [abi.make_bc(abi.BC_WALL), abi.make_bc(abi.BC_OUTFLOW), abi.make_bc(abi.BC_WALL, wall_u=0.1),
           abi.make_bc(abi.BC_OUTFLOW)]
    p = abi.default_params(equation=equation, flux_scheme=abi.FLUX_RUSANOV, gradient_scheme=abi.GRAD_LEAST_SQUARES,
                           mu=0.02, prandtl=0.72, prandtl_t=0.9, cfl=0.3)
    nu = 0.02
    ic = abi.initial_condition(abi.IC_FREESTREAM, rho=1.0, u=0.0, v=0.0, p=1.0, initial_nut=3.0 * nu,
                               initial_k=1e-4, initial_omega=10.0)
    return m, bcs, p, ic, None


def case_diffusion(n=16):
    """UnstructuredFVMSolver on the wavy mesh: Dirichlet bottom/top, Neumann sides."""
    m = M.skewed(n, 0.15)
    bcs = [abi.make_bc(abi.BC_DIRICHLET, value=1.0), abi.make_bc(abi.BC_NEUMANN, value=0.3),
           abi.make_bc(abi.BC_DIRICHLET, value=0.0), abi.make_bc(abi.BC_NEUMANN, value=0.0)]
    p = abi.default_params(equation=abi.EQ_DIFFUSION, alpha=0.1, fixed_dt=1e-3)
    ic = abi.initial_condition(abi.IC_FREESTREAM, initial_value=2.0, initial_radius=0.5)
    return m, bcs, p, ic, None


def case_advdiff(gradient=abi.GRAD_LEAST_SQUARES, n=16, u=0.8, v=-0.3):
    """AdvectionDiffusionFVMSolver on the wavy mesh: Dirichlet bottom/top,
    Neumann sides, a blob of radius 0.5 around the origin, a uniform
    velocity that leaves through two sides and enters through the others."""
    m = M.skewed(n, 0.15)
    bcs = [abi.make_bc(abi.BC_DIRICHLET, value=1.0), abi.make_bc(abi.BC_NEUMANN, value=0.3),
           abi.make_bc(abi.BC_DIRICHLET, value=0.0), abi.make_bc(abi.BC_NEUMANN, value=0.0)]
    p = abi.default_params(equation=abi.EQ_ADVECTION_DIFFUSION, alpha=0.05, fixed_dt=2e-3, advection_u=u,
                           advection_v=v, gradient_scheme=gradient)
    ic = abi.initial_condition(abi.IC_FREESTREAM, initial_value=2.0, initial_radius=0.5)
    return m, bcs, p, ic, None


def adjoint_naca(flux=abi.FLUX_HLLC, order=2, **launch):
    """Config-2 numerics on a small NACA 0012 O-grid (40 x 20 triangle pairs,
    farfield at 6 chords, growth 1.2: cells of moderate aspect ratio).
    Returns (mesh, bcs, params, ic, initial state or None)."""
    m = M.naca_ogrid(40, 20, r_far=6.0, ratio=1.2)
    rho, u, v, pinf = freestream(0.8, 1.25)
    ff = M.from_primitive(rho, u, v, pinf, GAMMA)
    bcs = [abi.make_bc(abi.BC_WALL), abi.make_bc(abi.BC_FARFIELD, ff)]
    ext = dict(order=2, limiter=abi.LIMITER_VENKATAKRISHNAN, rk_stages=3, local_dt=1) if order == 2 else {}
    p = abi.default_params(equation=abi.EQ_EULER, flux_scheme=flux, cfl=0.8, **ext, **launch)
    ic = abi.initial_condition(abi.IC_FREESTREAM, rho=rho, u=u, v=v, p=pinf)
    return m, bcs, p, ic, None


def adjoint_wavy(order=2, flux=abi.FLUX_HLLC, **launch):
    """Skewed quads, slip wall below, farfield left, outflow top and right,
    with a smooth non-uniform state (every ghost type, four-face cells)."""
    m = M.skewed(12, 0.15)
    ff = M.from_primitive(1.0, 0.4, 0.1, 1.0, GAMMA)
    bcs = [abi.make_bc(abi.BC_WALL), abi.make_bc(abi.BC_FARFIELD, ff), abi.make_bc(abi.BC_OUTFLOW),
           abi.make_bc(abi.BC_OUTFLOW)]
    ext = dict(order=2, limiter=abi.LIMITER_VENKATAKRISHNAN) if order == 2 else {}
    p = abi.default_params(equation=abi.EQ_EULER, flux_scheme=flux, cfl=0.5, **ext, **launch)
    ic = abi.initial_condition(abi.IC_FREESTREAM, rho=1.0, u=0.4, v=0.1, p=1.0)
    x, y = m.cell_xy[:, 0], m.cell_xy[:, 1]
    U0 = M.from_primitive(1.0 + 0.2 * np.sin(3 * x) * np.cos(2 * y), 0.4 + 0.1 * np.cos(4 * y), 0.1 * np.sin(5 * x),
                          1.0 + 0.15 * np.cos(3 * x + y), GAMMA)
    return m, bcs, p, ic, U0


ADJOINT_CASES = {
    "naca_hllc_o2": lambda: adjoint_naca(abi.FLUX_HLLC, 2),
    "naca_rusanov_o2": lambda: adjoint_naca(abi.FLUX_RUSANOV, 2),
    "naca_hllc_o1": lambda: adjoint_naca(abi.FLUX_HLLC, 1),
    "naca_rusanov_o1": lambda: adjoint_naca(abi.FLUX_RUSANOV, 1),
    "wavy_outflow_o2": lambda: adjoint_wavy(2),
    "wavy_outflow_rusanov_o2": lambda: adjoint_wavy(2, abi.FLUX_RUSANOV),
    # a solver with cell tiles (the adjoint itself runs the same passes)
    "naca_hllc_o2_tiles": lambda: adjoint_naca(abi.FLUX_HLLC, 2, kernel_variant=3, block_size=128),
    # the free stream with a density bump at the leading edge: away from it
    # the state is exactly uniform (zero gradients, zero MUSCL increments)
    "naca_hllc_o2_bump": lambda: adjoint_naca_bump(),
    # the hub mesh: a sixteen-face cell (the widest slot instances), hanging
    # nodes on its edges
    "hub_hllc_o2": lambda: case_euler_hub(),
}
# steps of the oracle from the case's start to the state the adjoint is taken at
ADJOINT_STEPS = {"naca_hllc_o2_bump": 0}


def adjoint_naca_bump():
    m, bcs, p, ic, _ = adjoint_naca(abi.FLUX_HLLC, 2)
    rho, u, v, pinf = freestream(0.8, 1.25)
    r2 = (m.cell_xy[:, 0] - 0.0) ** 2 + m.cell_xy[:, 1] ** 2
    bump = np.where(r2 < 0.25, 0.2 * np.exp(-8.0 * r2), 0.0)
    U0 = M.from_primitive(rho * (1.0 + bump), np.full(m.n_cells, u), np.full(m.n_cells, v), np.full(m.n_cells, pinf),
                          GAMMA)
    return m, bcs, p, ic, U0


def developed_state(case, steps=40):
    """The oracle's state after `steps` steps from the case's start."""
    from oracle.pyoracle import OracleSolver
    m, bcs, p, ic, U0 = case
    o = OracleSolver(m, bcs, p, ic)
    if U0 is not None:
        o.set_field(U0)
    if steps:
        o.step(steps)
    return o.field()


def rel_state_diff(a, b, floor=1e-300):
    """Per-cell norm-wise relative difference max_k |a_k - b_k| / max_k |b_k|,
    maximised over cells (a component that is zero by symmetry is judged
    against the cell's state scale, not against its own round-off)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    if a.ndim == 1:
        a, b = a[:, None], b[:, None]
    num = np.max(np.abs(a - b), axis=1)
    den = np.maximum(np.max(np.abs(b), axis=1), floor)
    return float(np.max(num / den))


def residual_close(r_gpu, r_ref, rtol=1e-10):
    """|r - r_ref| <= rtol |r_ref| + rtol max_k r_ref(step 1)_k.  The absolute
    floor uses the largest first-step component: a component that is zero by
    symmetry (rho_v in a 1D problem) carries only round-off, which differs at
    O(1) relatively even between two builds of the reference."""
    r_gpu = np.asarray(r_gpu)
    r_ref = np.asarray(r_ref)
    floor = rtol * np.max(np.abs(r_ref[0]))
    return bool(np.all(np.abs(r_gpu - r_ref) <= rtol * np.abs(r_ref) + floor))
