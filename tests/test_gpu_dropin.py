"""The C++ drop-ins (include/fvk_dropin.hpp: HipEulerFVMSolver,
HipNavierStokesFVMSolver, with the constructors and public API of the
reference's EulerFVMSolver and NavierStokesFVMSolver) run on the GPU through
a program compiled against the reference's own headers
(tests/dropin_check.cpp, built by `make -C oracle dropin`).  The reference
types carry the mesh, boundary conditions and initial condition; the
per-step residual norms and the final field must be those of the library's
own binding on the same case, bit for bit, and (first-order paths) those of
the oracle."""
import os
import subprocess

import numpy as np
import pytest

import cases
from finitevolume_amd import abi
from finitevolume_amd.solver import FVKSolver
from oracle.pyoracle import OracleSolver

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
EXE = os.path.join(os.path.dirname(HERE), "oracle", "_ref", "dropin_check")


def case_text(m, bcs, p, ic, steps):
    f = "%.17g"
    out = [f"{m.n_nodes} {m.n_faces} {m.n_cells} {len(bcs)}"]
    out += [f"{f % x} {f % y}" for x, y in m.node_xy]
    out += [f"{a} {b} {l} {r} {q}" for (a, b), (l, r), q in zip(m.face_nodes, m.face_cells, m.face_patch)]
    fp, np_ = m.cell_face_ptr, m.cell_node_ptr
    for c in range(m.n_cells):
        fc = m.cell_faces[fp[c]:fp[c + 1]]
        nc = m.cell_nodes[np_[c]:np_[c + 1]]
        out.append(f"{f % m.cell_volume[c]} {f % m.cell_xy[c, 0]} {f % m.cell_xy[c, 1]} {len(fc)} "
                   + " ".join(map(str, fc)) + f" {len(nc)} " + " ".join(map(str, nc)))
    eq = 0 if p.equation == abi.EQ_EULER else 1
    out.append(f"{eq} {p.flux_scheme} {f % p.gamma} {f % p.cfl} {f % p.gas_constant} {f % p.mu} {f % p.prandtl} "
               f"{p.gradient_scheme} {f % p.exact_riemann_tol} {p.exact_riemann_max_iter}")
    for b in bcs:
        out.append(f"{b.type} " + " ".join(f % v for v in b.farfield_state) +
                   f" {f % b.wall_u} {f % b.wall_v} {b.is_isothermal_wall} {f % b.wall_temperature}")
    vals = [ic.rho, ic.u, ic.v, ic.p, ic.rho_l, ic.u_l, ic.v_l, ic.p_l, ic.rho_r, ic.u_r, ic.v_r, ic.p_r, ic.x0]
    out.append(f"{ic.mode} " + " ".join(f % v for v in vals))
    out.append(str(steps))
    return "\n".join(out) + "\n"


def run_dropin(case, steps):
    m, bcs, p, ic, _ = case
    r = subprocess.run([EXE], input=case_text(m, bcs, p, ic, steps), capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    R = np.array([[float(x) for x in ln.split()[1:]] for ln in r.stdout.splitlines() if ln.startswith("R ")])
    U = np.array([[float(x) for x in ln.split()[1:]] for ln in r.stdout.splitlines() if ln.startswith("U ")])
    return R, U


CASES = {
    "naca_2r_hllc": (lambda: cases.case_naca(32, 16), 40, True),
    "naca_2r_rusanov": (lambda: cases.case_naca(32, 16, flux=abi.FLUX_RUSANOV), 40, True),
    "sod_exact": (lambda: cases.case_sod(flux=abi.FLUX_EXACT), 30, False),
    "ns_wavy_lsq": (lambda: cases.case_ns_wavy(), 30, False),
    "ns_wavy_gg": (lambda: cases.case_ns_wavy(gradient=abi.GRAD_GREEN_GAUSS), 30, False),
}


@pytest.mark.skipif(not os.path.exists(EXE), reason="drop-in check program not built (make -C oracle dropin)")
@pytest.mark.parametrize("name", sorted(CASES))
def test_dropin_matches_library_binding(name):
    make, steps, bitwise_oracle = CASES[name]
    case = make()
    m, bcs, p, ic, _ = case
    R, U = run_dropin(case, steps)
    g = FVKSolver(m, bcs, p, ic)
    rg = g.step(steps)
    assert R.shape == (steps, 4) and U.shape == (m.n_cells, 4)
    assert np.array_equal(R, rg[:, :4]), name
    assert np.array_equal(U, g.field()), name
    o = OracleSolver(m, bcs, p, ic)
    ro = o.step(steps)
    if bitwise_oracle:
        assert np.array_equal(U, o.field()), name
        assert cases.residual_close(R, ro[:, :4], 1e-13), name
    else:
        assert cases.rel_state_diff(U, o.field(), floor=1e-12) <= 1e-12, name
        assert cases.residual_close(R, ro[:, :4], 1e-10), name
