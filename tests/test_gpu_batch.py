"""Batched cases on the GPU: every case of the batch must step exactly as it
does alone (same kernels, same per-cell arithmetic and summation order), so
states are compared bit for bit with single-case solvers, which are pinned
to the oracle; the per-case residual norms are block sums grouped
differently and agree to round-off.  Also: frozen cases keep their state,
and replaying a captured hipGraph gives the same bits as direct launches."""
import numpy as np
import pytest

import cases
from finitevolume_amd import abi
from finitevolume_amd import mesh as M
from finitevolume_amd import voronoi as V
from finitevolume_amd.batch import FVKBatch
from finitevolume_amd.solver import FVKSolver
from oracle.pyoracle import OracleSolver

pytestmark = pytest.mark.gpu


def small_cases(n=3, stages=3):
    meshes, bcs, ics = [], [], []
    for k in range(n):
        prm = V.batch_case_params(k)
        m = V.naca_voronoi(1500 + 700 * k, prm["seed"])
        rho, u, v, p = cases.freestream(prm["mach"], prm["aoa"])
        ff = M.from_primitive(rho, u, v, p, cases.GAMMA)
        meshes.append(m)
        bcs.append([abi.make_bc(abi.BC_WALL), abi.make_bc(abi.BC_FARFIELD, ff)])
        ics.append(abi.initial_condition(abi.IC_FREESTREAM, rho=rho, u=u, v=v, p=p))
    params = abi.default_params(equation=abi.EQ_EULER, flux_scheme=abi.FLUX_HLLC, cfl=0.8, order=2,
                                limiter=abi.LIMITER_VENKATAKRISHNAN, rk_stages=stages, local_dt=1)
    return meshes, bcs, params, ics


@pytest.fixture(scope="module")
def three():
    return small_cases()


def with_variant(params, variant, block=0, tile_cells=0):
    p = abi.ParamsC.from_buffer_copy(params)
    p.kernel_variant, p.block_size, p.tile_cells = variant, block, tile_cells
    return p


@pytest.mark.parametrize("variant,block,tc", [(1, 0, 0), (3, 0, 0), (3, 256, 200), (3, 512, 448), (4, 0, 0),
                                           (4, 256, 200), (4, 512, 448)])
@pytest.mark.parametrize("graph", [False, True])
def test_batch_matches_single_cases(three, graph, variant, block, tc):
    """The tile launch over the active cases' tiles (variant 3, the default
    for SSP-RK), its persistent form pulling tiles from a queue (variant 4;
    25 steps are 75 launches, so the queue must reset itself every launch,
    under graph replay too) and the face sweep + gather (variant 1), against
    each case alone (the single-case tile kernel for variants 3 and 4)."""
    meshes, bcs, params0, ics = three
    params = with_variant(params0, variant, block, tc)
    single = with_variant(params0, 3 if variant == 4 else variant, block, tc)
    b = FVKBatch(meshes, bcs, params, ics)
    b.set_graph(graph)
    rb = b.step(25)
    for k in range(len(meshes)):
        g = FVKSolver(meshes[k], bcs[k], single, ics[k])
        rg = g.step(25)
        assert np.array_equal(b.field(k), g.field()), k
        assert np.allclose(rb[:, k, :], rg, rtol=1e-13, atol=1e-13 * np.abs(rg).max()), k


def test_batch_case_matches_oracle(three):
    meshes, bcs, params, ics = three
    b = FVKBatch(meshes, bcs, params, ics)
    rb = b.step(20)
    o = OracleSolver(meshes[1], bcs[1], params, ics[1])
    ro = o.step(20)
    assert cases.residual_close(rb[:, 1, :], ro, 1e-10)
    assert cases.rel_state_diff(b.field(1), o.field(), floor=1e-12) <= 1e-12


@pytest.mark.parametrize("variant", [1, 3, 4])
def test_frozen_case_keeps_state(three, variant):
    meshes, bcs, params0, ics = three
    params = with_variant(params0, variant)
    b = FVKBatch(meshes, bcs, params, ics)
    b.set_graph(True)  # freezing changes the device tile list, not the captured launches
    b.step(10)
    U1 = b.field(1).copy()
    b.set_active([1, 0, 1])
    r = b.step(10)
    assert np.array_equal(b.field(1), U1)
    assert np.all(r[:, 1, :] == 0.0)
    g0 = FVKSolver(meshes[0], bcs[0], with_variant(params0, 3 if variant == 4 else variant), ics[0])
    g0.step(20)
    assert np.array_equal(b.field(0), g0.field())


def test_forward_euler_graph_pairs():
    meshes, bcs, params, ics = small_cases(2, stages=1)
    a = FVKBatch(meshes, bcs, params, ics)
    c = FVKBatch(meshes, bcs, params, ics)
    c.set_graph(True)
    a.step(7)
    c.step(7)  # three captured pairs and one direct step
    c.step(4)
    a.step(4)
    for k in range(2):
        assert np.array_equal(a.field(k), c.field(k)), k


def test_autotuner_picks_a_valid_shape(three, tmp_path, monkeypatch):
    from finitevolume_amd import autotune
    monkeypatch.setenv("FVK_TUNE_CACHE", str(tmp_path / "tune.json"))
    meshes, bcs, params, ics = three
    p = abi.ParamsC.from_buffer_copy(params)
    best = autotune.tune(meshes[0], bcs[0], p, ics[0], steps=5, warmup=1)
    assert (p.kernel_variant, p.block_size) == (best["variant"], best["block"])
    shapes = {(v, b, 0) for v in (1, 2) for b in autotune.BLOCKS} | {(3, b, t) for b, t in autotune.TILE_SHAPES}
    assert (best["variant"], best["block"], best["tile_cells"]) in shapes
    again = autotune.tune(meshes[0], bcs[0], abi.ParamsC.from_buffer_copy(params), ics[0])
    assert again == best  # served from the cache
    g = FVKSolver(meshes[0], bcs[0], p, ics[0])
    ref = FVKSolver(meshes[0], bcs[0], params, ics[0])
    g.step(10)
    ref.step(10)
    assert np.array_equal(g.field(), ref.field())
    pb = abi.ParamsC.from_buffer_copy(params)
    bb = autotune.tune_batch(meshes, bcs, pb, ics, steps=3, warmup=1)
    assert pb.block_size == bb["block"] and pb.kernel_variant == bb["variant"] in (1, 3)
    assert pb.tile_cells == bb["tile_cells"]
