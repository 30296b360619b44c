"""GPU parity: the HIP library against the oracle (the C restatement,
itself pinned bit for bit to the reference in tests/test_oracle.py) on the
same inputs.  First-order Rusanov/HLLC paths are compared bit for bit;
paths that call pow/tanh (exact Riemann, SA, SST) use the per-step state
tolerance 1e-12 relative and the residual-curve tolerance 1e-10 relative
with an absolute floor of 1e-10 times the first-step residual."""
import numpy as np
import pytest

import cases
from finitevolume_amd import abi
from finitevolume_amd.solver import FVKSolver
from oracle.pyoracle import OracleSolver

pytestmark = pytest.mark.gpu


def pair(case, **kw):
    m, bcs, p, ic, U0 = case
    g = FVKSolver(m, bcs, p, ic)
    o = OracleSolver(m, bcs, p, ic)
    if U0 is not None:
        g.set_field(U0)
        o.set_field(U0)
    return g, o


def fields(s):
    f = s.field()
    return f if isinstance(f, tuple) else (f,)


def assert_bitwise(g, o):
    for a, b in zip(fields(g), fields(o)):
        assert np.array_equal(a, b), f"max abs diff {np.abs(a - b).max()}"


def assert_close(g, o, rtol=1e-12):
    for a, b in zip(fields(g), fields(o)):
        d = cases.rel_state_diff(a, b, floor=1e-12)
        assert d <= rtol, d


@pytest.mark.parametrize("flux", [abi.FLUX_RUSANOV, abi.FLUX_HLLC])
def test_1r_bitwise(flux):
    g, o = pair(cases.case_1r(flux=flux))
    for _ in range(4):
        rg, dg = g.step(50, with_dt=True)
        ro, do = o.step(50, with_dt=True)
        assert np.array_equal(dg, do)
        assert cases.residual_close(rg, ro, 1e-12)
        assert_bitwise(g, o)


def test_1r_exact():
    g, o = pair(cases.case_1r(flux=abi.FLUX_EXACT))
    rg = g.step(100)
    ro = o.step(100)
    assert cases.residual_close(rg, ro, 1e-10)
    assert_close(g, o, 1e-12)


@pytest.mark.parametrize("flux", [abi.FLUX_RUSANOV, abi.FLUX_HLLC, abi.FLUX_EXACT])
def test_sod(flux):
    g, o = pair(cases.case_sod(flux=flux))
    rg = g.step(100)
    ro = o.step(100)
    assert cases.residual_close(rg, ro, 1e-10)
    if flux == abi.FLUX_EXACT:
        assert_close(g, o, 1e-12)
    else:
        assert_bitwise(g, o)


def test_naca_2r_bitwise():
    g, o = pair(cases.case_naca(64, 32))
    rg = g.step(50)
    ro = o.step(50)
    assert cases.residual_close(rg, ro, 1e-12)
    assert_bitwise(g, o)


@pytest.mark.parametrize("stages,local", [(1, 0), (2, 0), (3, 1)])
def test_naca_second_order(stages, local):
    case = cases.case_naca(64, 32, order=2, limiter=abi.LIMITER_VENKATAKRISHNAN, rk_stages=stages, local_dt=local)
    g, o = pair(case)
    rg = g.step(30)
    ro = o.step(30)
    assert cases.residual_close(rg, ro, 1e-10)
    assert_close(g, o, 1e-12)


def test_1r_second_order_periodic():
    g, o = pair(cases.case_1r(flux=abi.FLUX_HLLC, order=2, limiter=abi.LIMITER_VENKATAKRISHNAN, rk_stages=2))
    rg = g.step(40)
    ro = o.step(40)
    assert cases.residual_close(rg, ro, 1e-10)
    assert_close(g, o, 1e-12)


@pytest.mark.parametrize("gradient", [abi.GRAD_LEAST_SQUARES, abi.GRAD_GREEN_GAUSS])
def test_ns_wavy_bitwise(gradient):
    g, o = pair(cases.case_ns_wavy(gradient=gradient))
    rg = g.step(100)
    ro = o.step(100)
    assert cases.residual_close(rg, ro, 1e-12)
    assert_bitwise(g, o)


@pytest.mark.parametrize("gradient", [abi.GRAD_LEAST_SQUARES, abi.GRAD_GREEN_GAUSS])
def test_sa_wavy(gradient):
    g, o = pair(cases.case_rans_wavy(abi.EQ_RANS_SA, gradient=gradient))
    rg = g.step(100)
    ro = o.step(100)
    assert cases.residual_close(rg, ro, 1e-10)
    assert_close(g, o, 1e-12)


@pytest.mark.parametrize("variant,kato", [(0, 0), (1, 1)])
def test_sst_wavy(variant, kato):
    g, o = pair(cases.case_rans_wavy(abi.EQ_RANS_SST, variant=variant, kato=kato))
    rg = g.step(100)
    ro = o.step(100)
    assert cases.residual_close(rg, ro, 1e-10)
    assert_close(g, o, 1e-12)


def test_diffusion_bitwise():
    g, o = pair(cases.case_diffusion())
    rg = g.step(200)
    ro = o.step(200)
    assert cases.residual_close(rg, ro, 1e-12)
    assert_bitwise(g, o)


@pytest.mark.parametrize("gradient", [abi.GRAD_LEAST_SQUARES, abi.GRAD_GREEN_GAUSS])
def test_advection_diffusion_bitwise(gradient):
    """AdvectionDiffusionFVMSolver (src/AdvectionDiffusionFVMSolver.cpp:42-104)
    on the device: 300 steps, the same state bits as the oracle."""
    g, o = pair(cases.case_advdiff(gradient=gradient))
    rg = g.step(300)
    ro = o.step(300)
    assert cases.residual_close(rg, ro, 1e-12)
    assert_bitwise(g, o)
    assert np.ptp(g.field()) > 0.1  # the blob is still there, moved and spread


@pytest.mark.parametrize("order", [1, 2])
def test_kernel_variants_bit_identical(order):
    """The two Euler kernel variants (face sweep + gather, fused per cell) and
    every workgroup size give the same state bits; the residual norms are
    block sums whose grouping follows the workgroup size, so they agree to
    round-off (and bit for bit between the variants at one size)."""
    ext = dict(order=order, limiter=abi.LIMITER_VENKATAKRISHNAN, rk_stages=3, local_dt=1) if order == 2 else {}
    ref = None
    by_bs = {}
    # variant 3: (256, 208) runs four workgroups per CU (the 128-VGPR tile kernel)
    shapes3 = ((64, 0), (128, 0), (192, 0), (256, 0), (256, 208), (320, 224))
    for variant in (1, 2, 3):
        for bs, tc in (shapes3 if variant == 3 else ((64, 0), (128, 0), (256, 0), (512, 0))):
            m, bcs, p, ic, _ = cases.case_naca(64, 32, kernel_variant=variant, block_size=bs, tile_cells=tc, **ext)
            g = FVKSolver(m, bcs, p, ic)
            r, dt = g.step(10, with_dt=True)
            out = (g.field(), r, dt)
            if ref is None:
                ref = out
            else:
                assert np.array_equal(out[0], ref[0]), (variant, bs)
                assert np.allclose(out[1], ref[1], rtol=1e-13, atol=0.0), (variant, bs)
                # the step's dt: an atomic minimum (variants 1, 2) or one per tile reduced by the finalize
                assert np.array_equal(out[2], ref[2]), (variant, bs)
            if tc in (0, 224):
                by_bs.setdefault(bs, []).append(r)
    for bs, rs in by_bs.items():
        if len(rs) > 1:
            assert np.array_equal(rs[0], rs[1]), bs


@pytest.mark.parametrize("case", ["1r_o2_periodic", "sod_exact", "naca_o1"])
def test_tile_variant_bit_identical(case):
    """The fused tile kernels against the face sweep + gather on periodic,
    exact-Riemann and first-order cases, several tile sizes."""
    for bs3 in (64, 256, 320, 512):
        out = []
        for variant in (1, 3):
            bs = bs3 if variant == 3 else min(bs3, 256)  # the tile kernels also take 5 wavefronts
            tc = {320: 224, 512: 448}.get(bs3, 0) if variant == 3 else 0
            if case == "1r_o2_periodic":
                c = cases.case_1r(flux=abi.FLUX_HLLC, order=2, limiter=abi.LIMITER_VENKATAKRISHNAN, rk_stages=2,
                                  kernel_variant=variant, block_size=bs, tile_cells=tc)
            elif case == "sod_exact":
                c = cases.case_sod(flux=abi.FLUX_EXACT)
                c[2].kernel_variant, c[2].block_size, c[2].tile_cells = variant, bs, tc
            else:
                c = cases.case_naca(64, 32, local_dt=1, kernel_variant=variant, block_size=bs, tile_cells=tc)
            g, _ = pair(c)
            r = g.step(20)
            out.append((g.field(), r))
        for o in out[1:]:
            assert np.array_equal(out[0][0], o[0]), (case, bs3)
            assert np.allclose(out[0][1], o[1], rtol=1e-13, atol=0.0), (case, bs3)


def test_global_dt_variants_match_oracle():
    for variant in (1, 2, 3):
        g, o = pair(cases.case_naca(64, 32, kernel_variant=variant))
        rg, dg = g.step(20, with_dt=True)
        ro, do = o.step(20, with_dt=True)
        assert np.array_equal(dg, do)
        assert_bitwise(g, o)


@pytest.mark.parametrize("gradient", [abi.GRAD_GREEN_GAUSS, abi.GRAD_LEAST_SQUARES])
@pytest.mark.parametrize("equation", [abi.EQ_NAVIER_STOKES, abi.EQ_RANS_SA, abi.EQ_RANS_SST])
def test_viscous_tile_bit_identical(equation, gradient):
    """The fused viscous tile pass (gradients, faces and cell updates per
    tile) against the gradient, face and cell passes: the same state bits at
    several tile sizes; residual norms to round-off (their block grouping
    differs).  With a global dt the tile pass follows a dt pass over the
    faces (first stage), and the step's dt must be the same bits too."""
    def run(variant, bs, local_dt):
        if equation == abi.EQ_NAVIER_STOKES:
            m, bcs, p, ic, _ = cases.case_ns_wavy(gradient=gradient, n=24, rk_stages=3, local_dt=local_dt)
        else:
            m, bcs, p, ic, _ = cases.case_rans_wavy(equation, gradient=gradient, n=24, rk_stages=3,
                                                    local_dt=local_dt)
        p.kernel_variant, p.block_size = variant, bs
        g = FVKSolver(m, bcs, p, ic)
        g.set_kernel_timing(True)
        r, dt = g.step(40, with_dt=True)
        names = g.kernel_times()
        f = g.field()
        return np.concatenate([x.ravel() for x in (f if isinstance(f, tuple) else (f,))]), r, names, dt
    for local_dt in (1, 0):
        ref, rref, _, dref = run(1, 256, local_dt)
        for bs in (64, 128, 256):
            out, r, names, dt = run(3, bs, local_dt)
            assert "visc_tile" in names and "visc_grad" not in names, (bs, sorted(names))
            assert ("visc_dt" in names) == (local_dt == 0), (bs, sorted(names))
            assert np.array_equal(out, ref), (equation, gradient, bs, local_dt)
            assert np.allclose(r, rref, rtol=1e-13, atol=0.0), (equation, gradient, bs, local_dt)
            assert np.array_equal(dt, dref), (equation, gradient, bs, local_dt)


def test_sixteen_face_cells_match_oracle():
    """Cells wider than nine faces take the sixteen-slot kernels (the per-cell
    passes and the tile kernel reading its slots where it uses them): the
    hub mesh's centre cell has sixteen faces.  Second-order Euler with local
    dt against the oracle bit for bit, for the split and the fused passes
    at two tile sizes."""
    for variant, bs in ((1, 256), (3, 64), (3, 256)):
        m, bcs, p, ic, U0 = cases.case_euler_hub()
        p.kernel_variant, p.block_size = variant, bs
        g, o = pair((m, bcs, p, ic, U0))
        assert g.launch_shape()["kernel_variant"] == variant
        for _ in range(3):
            rg = g.step(10)
            ro = o.step(10)
            assert cases.residual_close(rg, ro, 1e-12), (variant, bs)
            assert_bitwise(g, o)


@pytest.mark.parametrize("gradient", [abi.GRAD_GREEN_GAUSS, abi.GRAD_LEAST_SQUARES])
@pytest.mark.parametrize("kind", ["periodic", "triangles", "voronoi", "hub"])
def test_viscous_tile_meshes(kind, gradient):
    """The fused viscous tile (gradients of the owned and ring-1 cells in the
    tile) on the other mesh kinds: a doubly periodic quad box (the LSQ offsets
    carry the periodic shift), triangles (cells of three faces in four-slot
    registers) and Voronoi polygons of up to nine faces, with walls the ring-1
    cells touch, and the hub mesh's sixteen-face cell (the widest slot
    instance).  Same state bits as the split passes after 30 SSP-RK3 steps
    with local dt.  The periodic box has no wall, so no wall distance for the
    turbulence models: Navier-Stokes only there."""
    from finitevolume_amd import mesh as M, voronoi
    ff = M.from_primitive(1.0, 0.3, 0.1, 1.0, 1.4)
    if kind == "periodic":
        m, bcs = M.periodic_box(20, 16), []
    elif kind in ("triangles", "hub"):
        m = M.tri_box(12, 10) if kind == "triangles" else cases.hub_mesh()
        bcs = [abi.make_bc(abi.BC_WALL, is_isothermal_wall=1, wall_temperature=1.2), abi.make_bc(abi.BC_OUTFLOW),
               abi.make_bc(abi.BC_FARFIELD, ff), abi.make_bc(abi.BC_FARFIELD, ff)]
    else:
        m = voronoi.naca_voronoi(1500, seed=3)
        bcs = [abi.make_bc(abi.BC_WALL), abi.make_bc(abi.BC_FARFIELD, ff)]
    for b in bcs:
        b.farfield_nut, b.farfield_k, b.farfield_omega = 0.05, 2e-3, 3.0
    equations = (abi.EQ_NAVIER_STOKES,) if kind == "periodic" else (
        abi.EQ_NAVIER_STOKES, abi.EQ_RANS_SA, abi.EQ_RANS_SST)
    for equation in equations:
        out = []
        for variant, bs in ((1, 256), (3, 64), (3, 256)):
            p = abi.default_params(equation=equation, flux_scheme=abi.FLUX_HLLC, gradient_scheme=gradient, mu=0.01,
                                   cfl=0.4, rk_stages=3, local_dt=1)
            p.kernel_variant, p.block_size = variant, bs
            ic = abi.initial_condition(abi.IC_FREESTREAM, rho=1.0, u=0.3, v=0.1, p=1.0, initial_nut=0.06,
                                       initial_k=1e-3, initial_omega=5.0)
            g = FVKSolver(m, bcs, p, ic)
            r = g.step(30)
            f = g.field()
            out.append((np.concatenate([x.ravel() for x in (f if isinstance(f, tuple) else (f,))]), r))
            del g
        assert np.isfinite(out[0][0]).all(), (kind, gradient, equation)
        for o, r in out[1:]:
            assert np.array_equal(o, out[0][0]), (kind, gradient, equation)
            assert np.allclose(r, out[0][1], rtol=1e-13, atol=0.0), (kind, gradient, equation)


@pytest.mark.parametrize("equation", [abi.EQ_NAVIER_STOKES, abi.EQ_RANS_SA, abi.EQ_RANS_SST])
def test_viscous_rk3_local_dt(equation):
    """SSP-RK3 with local time steps on the viscous paths (as the RANS
    workload runs them) against the oracle."""
    if equation == abi.EQ_NAVIER_STOKES:
        case = cases.case_ns_wavy(rk_stages=3, local_dt=1)
    else:
        case = cases.case_rans_wavy(equation, rk_stages=3, local_dt=1)
    g, o = pair(case)
    rg, dg = g.step(60, with_dt=True)
    ro, do = o.step(60, with_dt=True)
    assert cases.residual_close(rg, ro, 1e-10)
    assert_close(g, o, 1e-12)
    # the step's dt (the smallest local dt), reduced from per-block minima
    assert np.allclose(dg, do, rtol=1e-12, atol=0.0)
