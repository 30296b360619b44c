"""CPU tests of the oracle (the C restatement in oracle/fvoracle.c).

It is pinned three ways before the GPU parity tests trust it:
  * against the golden fixtures written by the reference itself
    (tests/golden, made from the reference's own sources);
  * against the reference build live (oracle/_ref), where it is built;
  * against closed forms and invariants that the mathematics fixes and
    that the reference's own verification gates check (src/main.cpp).
"""
import math
import os

import numpy as np
import pytest

import cases
from finitevolume_amd import abi
from finitevolume_amd import mesh as M
from oracle import pyoracle, pyref

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def gold(name):
    return np.load(os.path.join(GOLD, name))


def run_oracle(case, nsteps, snapshots=()):
    m, bcs, p, ic, U0 = case
    o = pyoracle.OracleSolver(m, bcs, p, ic)
    if U0 is not None:
        o.set_field(U0)
    res, snaps, done = [], {}, 0
    for s in sorted(set(list(snapshots) + [nsteps])):
        res.append(o.step(s - done))
        done = s
        snaps[s] = o.field()
    return np.concatenate(res), snaps, o


# ---------------------------------------------------------------------------
# golden fixtures (reference outputs)
# ---------------------------------------------------------------------------

def test_flux_vectors_bitwise():
    import sys
    sys.path.insert(0, GOLD)
    from make_golden import flux_inputs
    g = gold("flux_vectors.npz")
    UL, UR, nxy = flux_inputs(int(g["seed"]), int(g["n"]))
    assert np.array_equal(UL[:8], g["UL"]) and np.array_equal(UR[:8], g["UR"]), "input generator drifted"
    for scheme, key in ((abi.FLUX_RUSANOV, "rusanov"), (abi.FLUX_HLLC, "hllc"), (abi.FLUX_EXACT, "exact")):
        out = pyoracle.flux(scheme, UL, UR, nxy)
        assert np.array_equal(out, g[key]), key


TRAJ = [
    ("euler_1r_rusanov.npz", lambda: cases.case_1r(), 200, (50, 100, 150)),
    ("euler_1r_hllc.npz", lambda: cases.case_1r(flux=abi.FLUX_HLLC), 100, ()),
    ("euler_sod_hllc.npz", lambda: cases.case_sod(flux=abi.FLUX_HLLC), 100, ()),
    ("euler_sod_exact.npz", lambda: cases.case_sod(flux=abi.FLUX_EXACT), 100, ()),
    ("euler_sod_rusanov.npz", lambda: cases.case_sod(flux=abi.FLUX_RUSANOV), 100, ()),
    ("euler_naca_2r.npz", lambda: cases.case_naca(64, 32), 100, ()),
    ("ns_wavy_ls.npz", lambda: cases.case_ns_wavy(gradient=abi.GRAD_LEAST_SQUARES), 100, ()),
    ("ns_wavy_gg.npz", lambda: cases.case_ns_wavy(gradient=abi.GRAD_GREEN_GAUSS), 100, ()),
    ("sa_wavy_ls.npz", lambda: cases.case_rans_wavy(abi.EQ_RANS_SA, gradient=abi.GRAD_LEAST_SQUARES), 100, ()),
    ("sa_wavy_gg.npz", lambda: cases.case_rans_wavy(abi.EQ_RANS_SA, gradient=abi.GRAD_GREEN_GAUSS), 100, ()),
    ("sst_wavy_v0k0.npz", lambda: cases.case_rans_wavy(abi.EQ_RANS_SST, variant=0, kato=0), 100, ()),
    ("sst_wavy_v1k1.npz", lambda: cases.case_rans_wavy(abi.EQ_RANS_SST, variant=1, kato=1), 100, ()),
    ("ns_couette.npz", lambda: cases.case_couette(abi.EQ_NAVIER_STOKES), 2000, ()),
    ("sa_couette.npz", lambda: cases.case_couette(abi.EQ_RANS_SA), 2000, ()),
    ("sst_couette.npz", lambda: cases.case_couette(abi.EQ_RANS_SST), 2000, ()),
    ("diffusion_wavy.npz", lambda: cases.case_diffusion(), 300, ()),
    ("advdiff_wavy_ls.npz", lambda: cases.case_advdiff(gradient=abi.GRAD_LEAST_SQUARES), 300, ()),
    ("advdiff_wavy_gg.npz", lambda: cases.case_advdiff(gradient=abi.GRAD_GREEN_GAUSS), 300, ()),
]


@pytest.mark.parametrize("name,make,nsteps,snaps", TRAJ, ids=[t[0] for t in TRAJ])
def test_trajectory_matches_reference(name, make, nsteps, snaps):
    g = gold(name)
    res, fields, _ = run_oracle(make(), nsteps, snaps)
    final = fields[nsteps]
    U = final[0] if isinstance(final, tuple) else final
    # state: bit for bit (same IEEE operation sequence as the reference)
    assert np.array_equal(U, g["U"])
    if isinstance(final, tuple):
        assert np.array_equal(final[1], g["T"])
    for s in snaps:
        f = fields[s]
        assert np.array_equal(f[0] if isinstance(f, tuple) else f, g[f"U_{s}"])
    # residual norms: the reference sums with an OpenMP reduction (order
    # varies), so relative 1e-12 with the first-step floor
    assert cases.residual_close(res, g["res"], 1e-12)


def test_gradients_and_face_gradient_match_reference():
    g = gold("gradient_wavy.npz")
    m = M.skewed(8, 0.15)
    o = pyoracle.OracleSolver(m, [abi.make_bc(abi.BC_OUTFLOW)] * 4, abi.default_params())
    assert np.array_equal(o.gradient(abi.GRAD_LEAST_SQUARES, g["phi"], g["bphi"]), g["g_ls"])
    assert np.array_equal(o.gradient(abi.GRAD_GREEN_GAUSS, g["phi"], g["bphi"]), g["g_gg"])
    assert np.array_equal(o.gradient(abi.GRAD_LEAST_SQUARES, g["phi_r"], g["bphi_r"]), g["g_ls_r"])
    assert np.array_equal(o.gradient(abi.GRAD_GREEN_GAUSS, g["phi_r"], g["bphi_r"]), g["g_gg_r"])


def test_wall_distance_matches_reference():
    g = gold("wall_distance.npz")
    m = M.sheared(1, 16, 0.3)
    assert np.array_equal(pyoracle.wall_distance_faces(m, np.nonzero(m.face_patch == 0)[0]), g["sheared"])
    n = M.naca_ogrid(64, 32)
    assert np.array_equal(pyoracle.wall_distance_faces(n, np.nonzero(n.face_patch == 0)[0]), g["naca"])


def test_turbulence_sources_match_reference():
    g = gold("turbulence_sources.npz")
    sa, eddy = pyoracle.sa_source(g["sa_in_nut"], g["sa_in_nu"], g["sa_in_omega"], g["sa_in_dist"], g["sa_in_grad"])
    assert np.array_equal(sa, g["sa_random"]) and np.array_equal(eddy, g["sa_random_eddy"])
    args = [g[f"sst_in_{k}"] for k in ("rho", "k", "omega", "nu", "dist", "gu", "gv", "gk", "gw")]
    for v in (0, 1):
        for k in (0, 1):
            assert np.array_equal(pyoracle.sst_source(*args, variant=v, kato=k), g[f"sst_v{v}k{k}"])


def test_cfl_ramp_matches_reference():
    g = gold("cfl_ramp.npz")
    for key in ("short", "falling", "flat", "mild_rise", "diverging"):
        out = np.array([pyoracle.next_cfl(cur, 10.0, g[f"seq_{key}"]) for cur in (0.05, 1.0, 9.0)])
        assert np.array_equal(out, g[f"out_{key}"]), key


# ---------------------------------------------------------------------------
# live reference (where oracle/_ref is built)
# ---------------------------------------------------------------------------

needs_ref = pytest.mark.skipif(not pyref.available(), reason="reference build (oracle/_ref) not present")


@needs_ref
@pytest.mark.parametrize("seed", [1, 2, 3])
def test_random_states_vs_reference(seed):
    """Random smooth states on a triangulated box with all three boundary
    types, every equation: oracle and reference agree bit for bit."""
    rng = np.random.default_rng(seed)
    m = M.tri_box(6, 5)
    ff = M.from_primitive(1.0, 0.4, -0.2, 0.8, 1.4)
    bcs = [abi.make_bc(abi.BC_WALL, wall_u=0.05, is_isothermal_wall=1, wall_temperature=0.9),
           abi.make_bc(abi.BC_FARFIELD, ff), abi.make_bc(abi.BC_OUTFLOW), abi.make_bc(abi.BC_WALL)]
    for b in bcs:
        b.farfield_nut, b.farfield_k, b.farfield_omega = 0.02, 1e-3, 2.0
    x, y = m.cell_xy[:, 0], m.cell_xy[:, 1]
    rho = 1.0 + 0.2 * np.sin(3 * x + rng.uniform()) * np.cos(2 * y)
    U0 = M.from_primitive(rho, 0.3 + 0.1 * y, -0.1 + 0.05 * x, 0.8 + 0.1 * x * y, 1.4)
    for eq in (abi.EQ_EULER, abi.EQ_NAVIER_STOKES, abi.EQ_RANS_SA, abi.EQ_RANS_SST):
        for fl in (0, 1, 2):
            for gs in (0, 1):
                p = abi.default_params(equation=eq, flux_scheme=fl, gradient_scheme=gs, mu=0.01, cfl=0.4)
                ic = abi.initial_condition(abi.IC_NONE, initial_nut=0.03, initial_k=2e-3, initial_omega=4.0)
                r = pyref.RefSolver(m, bcs, p, ic)
                o = pyoracle.OracleSolver(m, bcs, p, ic)
                nt = abi.nvars(eq) - 4
                T0 = None
                if nt:
                    T0 = np.stack([0.03 + 0.01 * np.sin(5 * x)] + ([4.0 + y] if nt == 2 else []), 1)
                    if nt == 2:
                        T0[:, 0] = 2e-3 * (1 + x)
                r.set_field(U0, T0)
                o.set_field(U0, T0)
                hr, ho = r.step(20), o.step(20)
                fr, fo = r.field(), o.field()
                fr = fr if isinstance(fr, tuple) else (fr,)
                fo = fo if isinstance(fo, tuple) else (fo,)
                for a, b in zip(fr, fo):
                    assert np.array_equal(a, b), (eq, fl, gs)
                assert cases.residual_close(ho, hr, 1e-12)


@needs_ref
def test_sixteen_face_cell_vs_reference():
    """The hub mesh (a sixteen-face cell with hanging nodes on its edges,
    surrounded by quads), every equation and gradient scheme: oracle and
    reference agree bit for bit after 20 steps."""
    m = cases.hub_mesh()
    ff = M.from_primitive(1.0, 0.4, -0.2, 0.8, 1.4)
    bcs = [abi.make_bc(abi.BC_WALL, is_isothermal_wall=1, wall_temperature=0.9), abi.make_bc(abi.BC_FARFIELD, ff),
           abi.make_bc(abi.BC_OUTFLOW), abi.make_bc(abi.BC_WALL)]
    for b in bcs:
        b.farfield_nut, b.farfield_k, b.farfield_omega = 0.02, 1e-3, 2.0
    x, y = m.cell_xy[:, 0], m.cell_xy[:, 1]
    U0 = M.from_primitive(1.0 + 0.2 * np.sin(2 * x) * np.cos(y), 0.3 + 0.05 * y, -0.1 + 0.03 * x,
                          0.8 + 0.05 * x * y, 1.4)
    for eq in (abi.EQ_EULER, abi.EQ_NAVIER_STOKES, abi.EQ_RANS_SA, abi.EQ_RANS_SST):
        for gs in (0, 1):
            p = abi.default_params(equation=eq, flux_scheme=abi.FLUX_HLLC, gradient_scheme=gs, mu=0.01, cfl=0.4)
            ic = abi.initial_condition(abi.IC_NONE, initial_nut=0.03, initial_k=2e-3, initial_omega=4.0)
            r = pyref.RefSolver(m, bcs, p, ic)
            o = pyoracle.OracleSolver(m, bcs, p, ic)
            nt = abi.nvars(eq) - 4
            T0 = None
            if nt:
                T0 = np.stack([0.03 + 0.01 * np.sin(3 * x)] + ([4.0 + y] if nt == 2 else []), 1)
                if nt == 2:
                    T0[:, 0] = 2e-3 * (1 + 0.3 * x)
            r.set_field(U0, T0)
            o.set_field(U0, T0)
            hr, ho = r.step(20), o.step(20)
            fr, fo = r.field(), o.field()
            fr = fr if isinstance(fr, tuple) else (fr,)
            fo = fo if isinstance(fo, tuple) else (fo,)
            for a, b in zip(fr, fo):
                assert np.array_equal(a, b), (eq, gs)
            assert cases.residual_close(ho, hr, 1e-12)


# ---------------------------------------------------------------------------
# closed forms and invariants
# ---------------------------------------------------------------------------

def _rand_states(rng, n):
    return M.from_primitive(rng.uniform(0.2, 2, n), rng.normal(0, 1, n), rng.normal(0, 1, n), rng.uniform(0.1, 3, n),
                            1.4)


@pytest.mark.parametrize("scheme", [0, 1, 2])
def test_flux_consistency(scheme):
    """F*(U, U) = F(U) (docs/hllc-and-exact-riemann-plan.md:151-155)."""
    rng = np.random.default_rng(5)
    U = _rand_states(rng, 500)
    th = rng.uniform(0, 2 * np.pi, 500)
    nxy = np.stack([np.cos(th), np.sin(th)], 1)
    F = pyoracle.flux(scheme, U, U, nxy)
    rho, u, v = U[:, 0], U[:, 1] / U[:, 0], U[:, 2] / U[:, 0]
    p = 0.4 * (U[:, 3] - 0.5 * rho * (u * u + v * v))
    Vn = u * nxy[:, 0] + v * nxy[:, 1]
    exact = np.stack([rho * Vn, U[:, 1] * Vn + p * nxy[:, 0], U[:, 2] * Vn + p * nxy[:, 1], (U[:, 3] + p) * Vn], 1)
    scale = np.abs(exact).max(1, keepdims=True) + 1.0
    assert np.max(np.abs(F - exact) / scale) < 1e-13


@pytest.mark.parametrize("scheme", [0, 1, 2])
def test_flux_reflection_symmetry(scheme):
    """F(L, R, n) = -F(R, L, -n)."""
    rng = np.random.default_rng(6)
    L, R = _rand_states(rng, 500), _rand_states(rng, 500)
    th = rng.uniform(0, 2 * np.pi, 500)
    nxy = np.stack([np.cos(th), np.sin(th)], 1)
    a = pyoracle.flux(scheme, L, R, nxy)
    b = pyoracle.flux(scheme, R, L, -nxy)
    assert np.max(np.abs(a + b) / (np.abs(a) + 1.0)) < 1e-12


def test_exact_sod_star_state():
    """Sod: p* = 0.30313, u* = 0.92745, rho*_L = 0.42632 (Toro ch. 4,
    docs/hllc-and-exact-riemann-plan.md:199-201).  At x/t = 0 the exact
    solution sits in the left star region, so the mass flux is rho*_L u*."""
    g = gold("exact_sod_star.npz")
    assert abs(g["p_star"][0] - 0.30313) < 5e-6 and abs(g["u_star"][0] - 0.92745) < 5e-6
    L = M.from_primitive(1.0, 0.0, 0.0, 1.0, 1.4)[None]
    R = M.from_primitive(0.125, 0.0, 0.0, 0.1, 1.4)[None]
    F = pyoracle.flux(abi.FLUX_EXACT, L, R, np.array([[1.0, 0.0]]))
    assert abs(F[0, 0] - 0.42632 * 0.92745) < 2e-5
    p_star = F[0, 1] - F[0, 0] * 0.92745  # momentum flux = rho u^2 + p
    assert abs(p_star - 0.30313) < 5e-5


def test_supersonic_upwind_and_vacuum():
    rng = np.random.default_rng(8)
    L = M.from_primitive(np.ones(50), rng.uniform(3, 5, 50), np.zeros(50), np.ones(50), 1.4)
    R = _rand_states(rng, 50)
    R[:, 1] = np.abs(R[:, 1]) + 4 * R[:, 0]
    nxy = np.tile([1.0, 0.0], (50, 1))
    FL = pyoracle.flux(abi.FLUX_HLLC, L, L, nxy)
    assert np.array_equal(pyoracle.flux(abi.FLUX_HLLC, L, R, nxy), FL)
    V = M.from_primitive(1.0, -8.0, 0.0, 0.4, 1.4)[None], M.from_primitive(1.0, 8.0, 0.0, 0.4, 1.4)[None]
    assert np.all(pyoracle.flux(abi.FLUX_EXACT, V[0], V[1], np.array([[1.0, 0.0]])) == 0.0)


def test_gradient_gate_numbers():
    """--verify-gradient (src/main.cpp:1443-1527): LSQ exact for a linear field,
    Green-Gauss error 7.34492e-2, Jasak face gradient exact, naive 1.73695
    (docs/references.md:199-204)."""
    g = gold("gradient_wavy.npz")
    m = M.skewed(8, 0.15)
    o = pyoracle.OracleSolver(m, [abi.make_bc(abi.BC_OUTFLOW)] * 4, abi.default_params())
    ls = o.gradient(abi.GRAD_LEAST_SQUARES, g["phi"], g["bphi"])
    gg = o.gradient(abi.GRAD_GREEN_GAUSS, g["phi"], g["bphi"])
    assert np.abs(ls - [2.3, -1.1]).max() < 1e-9
    assert abs(np.abs(gg - [2.3, -1.1]).max() - 7.34492e-2) < 5e-8
    exact_dn = 2.3 * g["face_normal"][:, 0] - 1.1 * g["face_normal"][:, 1]
    assert np.abs(g["dphidn"] - exact_dn).max() < 1e-9


def test_wall_distance_gate():
    """--verify-wall-distance: distance to the flat bottom wall = y_centroid."""
    m = M.sheared(1, 16, 0.3)
    d = pyoracle.wall_distance_faces(m, np.nonzero(m.face_patch == 0)[0])
    assert np.abs(d - m.cell_xy[:, 1]).max() < 1e-9


def test_sa_source_gate():
    """--verify-sa-source (src/main.cpp:2490-2563): the zero state is a fixed
    point; the manufactured maxima are P 0.0439688, D 0.707294, CD 9.33e-05."""
    out, _ = pyoracle.sa_source(np.zeros(16), np.full(16, 0.02), np.zeros(16), np.linspace(0.03, 0.97, 16),
                                np.zeros((16, 2)))
    assert np.all(out == 0.0)
    g = gold("turbulence_sources.npz")
    mx = g["sa_manufactured"].max(0)
    assert abs(mx[0] - 0.0439688) < 5e-8 and abs(mx[1] - 0.707294) < 5e-7 and abs(mx[2] - 9.33e-05) < 5e-8


def test_ns_uniform_flow_stays_uniform():
    """--verify-ns-uniform (src/main.cpp:1620-1662): exactly uniform after 20 steps."""
    m = M.skewed(8, 0.15)
    st = M.from_primitive(1.0, 1.0, 0.5, 1.0, 1.4)
    bcs = [abi.make_bc(abi.BC_FARFIELD, st) for _ in range(4)]
    p = abi.default_params(equation=abi.EQ_NAVIER_STOKES, mu=0.5, cfl=0.3)
    ic = abi.initial_condition(abi.IC_FREESTREAM, rho=1.0, u=1.0, v=0.5, p=1.0)
    o = pyoracle.OracleSolver(m, bcs, p, ic)
    o.step(20)
    assert np.abs(o.field() - st).max() < 1e-9


@pytest.mark.slow
def test_couette_profile():
    """--verify-couette (src/main.cpp:1701-1762): u = U y / H within 5% L2 (the
    reference reaches 1.4e-14 after 20000 steps)."""
    m, bcs, p, ic, _ = cases.case_couette(abi.EQ_NAVIER_STOKES)
    o = pyoracle.OracleSolver(m, bcs, p, ic)
    o.step(20000)
    U = o.field()
    u = U[:, 1] / U[:, 0]
    err = np.sqrt(np.mean((u - 0.1 * m.cell_xy[:, 1]) ** 2)) / 0.1
    assert err < 1e-6


def test_mass_conservation_torus():
    m, bcs, p, ic, U0 = cases.case_1r(flux=abi.FLUX_HLLC, order=2, limiter=abi.LIMITER_VENKATAKRISHNAN, rk_stages=3)
    o = pyoracle.OracleSolver(m, bcs, p, ic)
    o.set_field(U0)
    o.step(20)
    U = o.field()
    for q in range(4):
        a, b = (U0[:, q] * m.cell_volume).sum(), (U[:, q] * m.cell_volume).sum()
        assert abs(a - b) <= 1e-13 * (abs(a) + 1.0)


# ---- extensions: reductions to the reference path ----

@pytest.mark.parametrize("scheme", [abi.FLUX_HLLC, abi.FLUX_RUSANOV])
def test_second_order_with_zero_slopes_is_first_order(scheme):
    """The first-order limiter (phi = 0) of the second-order path is the
    reference step: its face states are the cells' primitives, and the fluxes
    on primitive states equal the reference fluxes on the conserved states up
    to rounding.  Checked every step (state within 1e-13 relative, residual
    within 1e-12), on a mesh with walls and far field, where a dropped term, a
    wrong sign or a swapped side in the primitive fluxes or ghost states gives
    an O(1) difference."""
    base = cases.case_naca(32, 16)
    m, bcs, p1, ic, _ = base
    p1.flux_scheme = scheme
    p2 = abi.default_params(equation=abi.EQ_EULER, flux_scheme=scheme, cfl=0.8, order=2,
                            limiter=abi.LIMITER_FIRST_ORDER)
    a = pyoracle.OracleSolver(m, bcs, p1, ic)
    b = pyoracle.OracleSolver(m, bcs, p2, ic)
    for _ in range(20):
        ra, rb = a.step(1), b.step(1)
        ua, ub = a.field(), b.field()
        assert np.max(np.abs(ua - ub) / np.maximum(np.abs(ua), 1e-12)) <= 1e-13
        assert np.all(np.abs(ra - rb) <= 1e-12 * np.abs(ra) + 1e-300)
    assert np.all(ra > 0.0)  # the flow is moving


def test_local_dt_minimum_is_reference_dt():
    m, bcs, p, ic, _ = cases.case_naca(32, 16)
    pl = abi.default_params(equation=abi.EQ_EULER, flux_scheme=abi.FLUX_HLLC, cfl=0.8, local_dt=1)
    a = pyoracle.OracleSolver(m, bcs, p, ic)
    b = pyoracle.OracleSolver(m, bcs, pl, ic)
    _, da = a.step(1, with_dt=True)
    _, db = b.step(1, with_dt=True)
    assert da[0] == db[0]


def test_muscl_exact_for_linear_fields():
    """LSQ gradients of linear primitive fields are exact in cells whose
    neighbours are all cells, so the face values are exact there."""
    m = M.tri_box(12, 10)
    x, y = m.cell_xy[:, 0], m.cell_xy[:, 1]
    U = M.from_primitive(1.0 + 0.1 * x + 0.05 * y, 0.3 - 0.2 * y, 0.1 + 0.1 * x, 1.0 + 0.2 * x - 0.1 * y, 1.4)
    bcs = [abi.make_bc(abi.BC_OUTFLOW)] * 4
    p = abi.default_params(equation=abi.EQ_EULER, order=2, limiter=abi.LIMITER_NONE)
    o = pyoracle.OracleSolver(m, bcs, p)
    o.set_field(U)
    o.residual_eval()
    g, phi = o.last_gradients()
    interior = np.ones(m.n_cells, bool)
    interior[m.face_cells[m.face_cells[:, 1] < 0, 0]] = False
    want = np.array([0.1, 0.05, 0.0, -0.2, 0.1, 0.0, 0.2, -0.1])
    assert np.abs(g[interior] - want).max() < 1e-12
    assert np.all(phi == 1.0)


def test_venkatakrishnan_range():
    rng = np.random.default_rng(3)
    for _ in range(2000):
        d1 = rng.uniform(0, 1) * (1 if rng.uniform() < 0.5 else -1)
        d2 = abs(rng.normal()) * math.copysign(1, d1)
        ph = pyoracle.venkatakrishnan(d1, d2, 0.0)
        # (r^2 + 2r) / (r^2 + r + 2), r = d1/d2 >= 0, peaks at r = 2 + 2 sqrt(2)
        r = 2.0 + 2.0 * math.sqrt(2.0)
        assert -1e-15 <= ph <= (r * r + 2 * r) / (r * r + r + 2) + 1e-15
    assert abs(pyoracle.venkatakrishnan(0.3, 0.1, 1e6) - 1.0) < 1e-6


def test_venkatakrishnan_limiter_brute_force():
    """phi per cell and variable is min(1, min over the cell's faces of the
    limiter function), recomputed here face by face from the unlimited
    gradients.  Outflow boundaries make the ghost equal to the cell, so a
    boundary face adds its midpoint to the minimum but no new extreme."""
    m = M.tri_box(14, 12)
    rng = np.random.default_rng(5)
    x, y = m.cell_xy[:, 0], m.cell_xy[:, 1]
    n = m.n_cells
    W = np.stack([1.0 + 0.3 * np.sin(5 * x) + 0.05 * rng.standard_normal(n),
                  0.3 + 0.2 * np.cos(4 * y) + 0.05 * rng.standard_normal(n),
                  -0.1 + 0.2 * x * y + 0.05 * rng.standard_normal(n),
                  1.0 + 0.2 * np.sin(3 * x + 2 * y) + 0.05 * rng.standard_normal(n)], 1)
    U = M.from_primitive(W[:, 0], W[:, 1], W[:, 2], W[:, 3], 1.4)
    bcs = [abi.make_bc(abi.BC_OUTFLOW)] * 4
    out = {}
    for lim in (abi.LIMITER_NONE, abi.LIMITER_VENKATAKRISHNAN):
        o = pyoracle.OracleSolver(m, bcs, abi.default_params(equation=abi.EQ_EULER, order=2, limiter=lim,
                                                             venkat_k=0.3))
        o.set_field(U)
        o.residual_eval()
        out[lim] = o.last_gradients()
    g0, _ = out[abi.LIMITER_NONE]
    g, phi = out[abi.LIMITER_VENKATAKRISHNAN]
    Wc = np.stack([U[:, 0], U[:, 1] / U[:, 0], U[:, 2] / U[:, 0], 0.4 * (U[:, 3] - 0.5 * U[:, 0] * (
        (U[:, 1] / U[:, 0]) ** 2 + (U[:, 2] / U[:, 0]) ** 2))], 1)
    want = np.ones((n, 4))
    for c in range(n):
        faces = m.cell_faces[m.cell_face_ptr[c]:m.cell_face_ptr[c + 1]]
        nb = [int(a if b == c else b) for a, b in m.face_cells[faces] if a >= 0 and b >= 0]
        wmin = np.min(Wc[[c] + nb], axis=0)
        wmax = np.max(Wc[[c] + nb], axis=0)
        eps2 = (0.3 * math.sqrt(m.cell_volume[c])) ** 3
        for f in faces:
            r = 0.5 * (m.node_xy[m.face_nodes[f, 0]] + m.node_xy[m.face_nodes[f, 1]]) - m.cell_xy[c]
            for q in range(4):
                d2 = g0[c, 2 * q] * r[0] + g0[c, 2 * q + 1] * r[1]
                if d2 == 0.0:
                    continue
                d1 = (wmax[q] if d2 > 0 else wmin[q]) - Wc[c, q]
                ph = (d1 * d1 + eps2 + 2 * d1 * d2) / (d1 * d1 + 2 * d2 * d2 + d1 * d2 + eps2)
                want[c, q] = min(want[c, q], ph)
    assert (want < 0.999).mean() > 0.3  # the limiter is active in most cells
    assert np.abs(phi - want).max() <= 1e-12
    assert np.array_equal(g, np.repeat(phi, 2, axis=1) * g0)


def _vortex_error(n, order, stages):
    """L1 density error of the isentropic vortex after it is advected a
    fixed time (an exact translating solution)."""
    m = M.periodic_box(n, n)
    U0 = M.isentropic_vortex(m.cell_xy)
    p = abi.default_params(equation=abi.EQ_EULER, flux_scheme=abi.FLUX_HLLC, cfl=0.4, order=order,
                           limiter=abi.LIMITER_NONE, rk_stages=stages)
    o = pyoracle.OracleSolver(m, [], p)
    o.set_field(U0)
    t, T = 0.0, 0.05
    while t < T - 1e-14:
        _, dt = o.step(1, with_dt=True)
        t += dt[0]
        if t > T:
            break
    c = math.sqrt(1.4)
    ux, uy = 0.5 * c / math.sqrt(2), 0.5 * c / math.sqrt(2)
    exact = M.isentropic_vortex(m.cell_xy, center=(0.5 + ux * t, 0.5 + uy * t))
    return np.abs(o.field()[:, 0] - exact[:, 0]).mean()


def test_vortex_convergence_order():
    e1 = [_vortex_error(n, 1, 1) for n in (32, 64)]
    e2 = [_vortex_error(n, 2, 3) for n in (32, 64)]
    o1 = math.log2(e1[0] / e1[1])
    o2 = math.log2(e2[0] / e2[1])
    assert e2[1] < e1[1]
    assert o2 > 1.6, (o1, o2)
    assert 0.5 < o1 < 1.4, (o1, o2)


# ---------------------------------------------------------------------------
# wall diagnostics (src/WallTraction.cpp:95-216)
# ---------------------------------------------------------------------------

def _wall_fixture_cases():
    import sys
    sys.path.insert(0, GOLD)
    from make_golden import WALL_REF, wall_cases
    return WALL_REF, wall_cases()


def test_wall_reports_match_reference_fixture():
    """Samples (traction, Cf, Cp, y+) and per-patch force reports of the
    oracle against the reference's compute_wall_traction_samples and
    compute_wall_forces, from the same states, bit for bit."""
    g = gold("wall_reports.npz")
    wref, wcases = _wall_fixture_cases()
    ref = abi.wall_reference(**wref)
    for key, case, _ in wcases:
        m, bcs, p, ic, _ = case
        o = pyoracle.OracleSolver(m, bcs, p, ic)
        o.set_field(g[f"{key}_U"], g[f"{key}_T"] if f"{key}_T" in g.files else None)
        wf = abi.collect_wall_faces(m, bcs)
        for po in (0, 1):
            smp, rep = o.wall_reports(wf, ref, bool(po))
            gs, gr = g[f"{key}_p{po}_samples"], g[f"{key}_p{po}_reports"]
            for k in smp.dtype.names:
                assert np.array_equal(smp[k], gs[k], equal_nan=True), (key, po, k)
            assert len(rep) == len(gr), key
            for k in rep.dtype.names:
                assert np.array_equal(rep[k], gr[k]), (key, po, k)
        if p.equation != abi.EQ_EULER:
            assert np.any(g[f"{key}_p0_samples"]["tau_wall"] != 0.0), key


def test_wall_forces_vanish_on_closed_body_under_uniform_pressure():
    """A uniform pressure on a closed section gives zero net force and zero
    moment: the sum of A n over a closed polygon is zero, and so is the sum
    of r x (A n)."""
    m, bcs, p, ic, _ = cases.case_naca(64, 32)
    o = pyoracle.OracleSolver(m, bcs, p, ic)
    U = np.tile(M.from_primitive(1.3, 0.0, 0.0, 2.5, 1.4), (m.n_cells, 1))
    o.set_field(U)
    wf = abi.collect_wall_faces(m, bcs)
    ref = abi.wall_reference(1.0, 1.0, 0.2, 0.0, 1.0, 0.25, 0.05)
    smp, rep = o.wall_reports(wf, ref)
    tot = rep[-1]
    assert tot["patch_id"] == -1
    scale = 2.5 * len(wf)  # pressure times the number of terms (chord 1, every |r| <= 1)
    for k in ("total_drag", "lift", "moment"):
        assert abs(tot[k]) <= 1e-14 * scale, (k, tot[k])
    cl, cd = o.wall_forces(1.0, 1.0, 0.2)
    assert abs(cl) <= 1e-13 and abs(cd) <= 1e-13
    # and a non-uniform pressure does not vanish
    U[:, 3] *= 1.0 + 0.1 * m.cell_xy[:, 0]
    o.set_field(U)
    assert abs(o.wall_forces(1.0, 1.0, 0.2)[1]) > 1e-3
