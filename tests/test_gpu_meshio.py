"""A run on meshes read from files: the library reads the file (FVMESH and
Gmsh, the reference's face order), the GPU solver and the oracle start from
the same read mesh, and their trajectories agree to the parity bar.  Patch
boundary conditions are assigned by patch name, as the reference driver
assigns them (src/main.cpp:3405-3429)."""
import os

import numpy as np
import pytest

import cases
from finitevolume_amd import abi
from finitevolume_amd.meshio import read_mesh
from finitevolume_amd.solver import FVKSolver
from oracle.pyoracle import OracleSolver

pytestmark = pytest.mark.gpu

MESHES = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "meshes")


def bcs_by_name(mesh, table):
    return [table[name] for name in mesh.patches]


@pytest.mark.parametrize("name", ["voronoi.fvmesh", "naca_tri.msh", "naca_mixed_v41.msh"])
def test_euler_on_read_mesh(name):
    m = read_mesh(os.path.join(MESHES, name))
    rho, u, v, pinf = cases.freestream(0.7, 2.0)
    ff = abi.make_bc(abi.BC_FARFIELD, cases.M.from_primitive(rho, u, v, pinf, 1.4))
    bcs = bcs_by_name(m, {"wall": abi.make_bc(abi.BC_WALL), "farfield": ff})
    ic = abi.initial_condition(abi.IC_FREESTREAM, rho=rho, u=u, v=v, p=pinf)
    for cfl, ext in ((0.8, {}), (0.4, dict(order=2, limiter=abi.LIMITER_VENKATAKRISHNAN, rk_stages=3, local_dt=1))):
        p = abi.default_params(equation=abi.EQ_EULER, flux_scheme=abi.FLUX_HLLC, cfl=cfl, **ext)
        g, o = FVKSolver(m, bcs, p, ic), OracleSolver(m, bcs, p, ic)
        rg, ro = g.step(40), o.step(40)
        assert np.all(np.isfinite(ro)), (name, ext)
        assert cases.residual_close(rg, ro, 1e-10), (name, ext)
        if ext:
            assert cases.rel_state_diff(g.field(), o.field(), floor=1e-12) <= 1e-12, name
        else:
            assert np.array_equal(g.field(), o.field()), name


def test_navier_stokes_on_read_mesh():
    m = read_mesh(os.path.join(MESHES, "wavy_quads_v22.msh"))
    ff = abi.make_bc(abi.BC_FARFIELD, cases.M.from_primitive(1.0, 0.3, 0.1, 1.0, 1.4))
    bcs = bcs_by_name(m, {"bottom": abi.make_bc(abi.BC_WALL, is_isothermal_wall=1, wall_temperature=1.2),
                          "left": ff, "top": ff, "right": abi.make_bc(abi.BC_OUTFLOW)})
    p = abi.default_params(equation=abi.EQ_NAVIER_STOKES, flux_scheme=abi.FLUX_HLLC, mu=0.01, cfl=0.4)
    ic = abi.initial_condition(abi.IC_FREESTREAM, rho=1.0, u=0.3, v=0.1, p=1.0)
    g, o = FVKSolver(m, bcs, p, ic), OracleSolver(m, bcs, p, ic)
    rg, ro = g.step(50), o.step(50)
    assert cases.residual_close(rg, ro, 1e-12)
    assert np.array_equal(g.field(), o.field())
