"""Parity at the benchmark sizes, on the GPU.

The workloads are the ones `bench.py` times: the 1,048,576-cell config-2
NACA 0012 O-grid, its first-order reference mode 2R, and the 601,088-cell
config-3 SA RANS mesh. The GPU runs against the oracle for a few steps from
the free-stream start, to the same bars as the small cases:
- bit for bit on the first-order path (states and dt);
- 1e-12 per-step state difference and 1e-10 residual curves elsewhere.

The size-independent property checked further on is that every kernel
variant gives the same state bits after 100 steps.
"""
import numpy as np
import pytest

import bench
import cases
from finitevolume_amd import abi
from finitevolume_amd.solver import FVKSolver
from oracle.pyoracle import OracleSolver

pytestmark = pytest.mark.gpu


def fields(s):
    f = s.field()
    return f if isinstance(f, tuple) else (f,)


def with_shape(p, variant, block, tile_cells=0):
    q = abi.ParamsC.from_buffer_copy(p)
    q.kernel_variant, q.block_size, q.tile_cells = variant, block, tile_cells
    return q


@pytest.fixture(scope="module")
def config2():
    return bench.naca_case(1024, 512, order2=True)


def test_config2_full_size_against_oracle(config2):
    m, bcs, p, ic = config2
    assert m.n_cells == 1048576
    o = OracleSolver(m, bcs, p, ic)
    ro = o.step(3)
    ref = fields(o)
    # tiles of 224 cells run three workgroups per CU, tiles of 208 four (the
    # 128-VGPR tile kernel)
    for variant, block, tc in ((1, 512, 0), (3, 256, 224), (3, 256, 208), (3, 512, 448)):
        g = FVKSolver(m, bcs, with_shape(p, variant, block, tc), ic)
        rg = g.step(3)
        for a, b in zip(fields(g), ref):
            assert cases.rel_state_diff(a, b, floor=1e-12) <= 1e-12, variant
        assert cases.residual_close(rg, ro, 1e-10), variant


def test_2r_full_size_bitwise():
    m, bcs, p, ic = bench.naca_case(1024, 512, order2=False)
    o = OracleSolver(m, bcs, p, ic)
    ro, do = o.step(10, with_dt=True)
    for variant, block in ((1, 512), (3, 256)):
        g = FVKSolver(m, bcs, with_shape(p, variant, block), ic)
        rg, dg = g.step(10, with_dt=True)
        assert np.array_equal(dg, do), variant
        for a, b in zip(fields(g), fields(o)):
            assert np.array_equal(a, b), (variant, float(np.abs(a - b).max()))
        assert cases.residual_close(rg, ro, 1e-12), variant


def test_config3_full_size_against_oracle():
    m, bcs, p, ic = bench.rans_case()
    assert m.n_cells == 601088
    o = OracleSolver(m, bcs, p, ic)
    ro = o.step(3)
    g = FVKSolver(m, bcs, p, ic)
    rg = g.step(3)
    for a, b in zip(fields(g), fields(o)):
        assert cases.rel_state_diff(a, b, floor=1e-12) <= 1e-12
    assert cases.residual_close(rg, ro, 1e-10)


def test_config2_variants_identical_after_100_steps(config2):
    m, bcs, p, ic = config2
    out = []
    for variant, block, tc in ((1, 512, 0), (2, 256, 0), (3, 256, 224), (3, 256, 208), (3, 192, 160)):
        g = FVKSolver(m, bcs, with_shape(p, variant, block, tc), ic)
        r, dt = g.step(100, with_dt=True)
        out.append((fields(g)[0], r, dt))
    for f, r, dt in out[1:]:
        assert np.array_equal(f, out[0][0])
        assert np.allclose(r, out[0][1], rtol=1e-12, atol=0.0)
        assert np.array_equal(dt, out[0][2])
