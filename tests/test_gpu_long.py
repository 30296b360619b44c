"""Long-horizon parity on the GPU: the residual-history bar of the plain path
(per-step conserved-state difference <= 1e-12 relative, L2 residual curves
within 1e-10 |r| + 1e-10 r(step 1) over 1000 steps) and configuration 1 as
stated (first-order Rusanov, SSP-RK2, CFL 0.5, 200 steps).

Every run compares the HIP library with the oracle after every step, in the
caller's (reference) cell order.  First-order paths are compared bit for
bit; the second-order path to the 1e-12 bar.
"""
import numpy as np
import pytest

import cases
from finitevolume_amd import abi
from finitevolume_amd.solver import FVKSolver
from oracle.pyoracle import OracleSolver

pytestmark = pytest.mark.gpu


def pair(case):
    m, bcs, p, ic, U0 = case
    g = FVKSolver(m, bcs, p, ic)
    o = OracleSolver(m, bcs, p, ic)
    if U0 is not None:
        g.set_field(U0)
        o.set_field(U0)
    return g, o


def run_lockstep(case, nsteps, bitwise, label):
    """Steps both solvers one step at a time; returns the residual and dt
    histories and the largest per-step state difference seen."""
    g, o = pair(case)
    rg, ro, dg, do = [], [], [], []
    worst = 0.0
    for k in range(nsteps):
        a, da = g.step(1, with_dt=True)
        b, db = o.step(1, with_dt=True)
        rg.append(a[0]); ro.append(b[0]); dg.append(da[0]); do.append(db[0])
        ug, uo = g.field(), o.field()
        if bitwise:
            assert np.array_equal(ug, uo), f"{label}: state differs at step {k + 1}"
        else:
            d = cases.rel_state_diff(ug, uo, floor=1e-12)
            worst = max(worst, d)
            assert d <= 1e-12, f"{label}: state difference {d:.3e} at step {k + 1}"
    return np.array(rg), np.array(ro), np.array(dg), np.array(do), worst


def test_config1_as_stated():
    """64x64 periodic torus, isentropic vortex, first-order Rusanov, SSP-RK2
    (Heun), global CFL 0.5, 200 steps: states and dt bit for bit every step,
    the density residual history within the curve bar."""
    case = cases.case_1r(flux=abi.FLUX_RUSANOV, rk_stages=2)
    assert case[2].cfl == 0.5 and case[2].order == 1 and case[2].local_dt == 0
    rg, ro, dg, do, _ = run_lockstep(case, 200, True, "config 1")
    assert np.array_equal(dg, do)
    assert cases.residual_close(rg, ro, 1e-10)
    assert np.all(np.isfinite(rg)) and rg[-1, 0] > 0.0


@pytest.mark.parametrize("flux", [abi.FLUX_RUSANOV, abi.FLUX_HLLC])
def test_1r_1000_steps(flux):
    """Reference mode 1R (forward Euler, global dt) for 1000 steps."""
    rg, ro, dg, do, _ = run_lockstep(cases.case_1r(flux=flux), 1000, True, f"1R flux {flux}")
    assert np.array_equal(dg, do)
    assert cases.residual_close(rg, ro, 1e-10)


def test_2r_naca_128x64_1000_steps():
    """Reference mode 2R on the 128 x 64 NACA 0012 O-grid (16,384 triangles):
    first-order HLLC, forward Euler, global CFL 0.8, 1000 steps."""
    case = cases.case_naca(128, 64)
    assert case[0].n_cells == 16384
    rg, ro, dg, do, _ = run_lockstep(case, 1000, True, "2R 128x64")
    assert np.array_equal(dg, do)
    assert cases.residual_close(rg, ro, 1e-10)


def test_config2_reduced_1000_steps():
    """Configuration 2's numerics (second-order HLLC, LSQ + Venkatakrishnan
    K = 5, SSP-RK3, local dt at CFL 0.8) on the 256 x 128 O-grid (65,536
    triangles) for 1000 steps, with the fused tile kernel the benchmark
    runs."""
    case = cases.case_naca(256, 128, order=2, limiter=abi.LIMITER_VENKATAKRISHNAN, rk_stages=3, local_dt=1,
                           kernel_variant=3, block_size=256, tile_cells=216)
    assert case[0].n_cells == 65536
    rg, ro, dg, do, worst = run_lockstep(case, 1000, False, "config 2 reduced")
    print(f"config 2 reduced: worst per-step state difference {worst:.3e}")
    assert cases.residual_close(rg, ro, 1e-10)
    assert np.allclose(dg, do, rtol=1e-12, atol=0.0)
