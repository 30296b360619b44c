"""fvk_step returns once its residual rows are final; with three or more RK
stages the last step's later stages may still be running.  Every call that
reads or writes the device waits for them first.  These runs interleave
such calls with single steps; they must give the same bits as one
uninterrupted run."""
import numpy as np
import pytest

import cases
from finitevolume_amd import abi
from finitevolume_amd.solver import FVKSolver

pytestmark = pytest.mark.gpu


def _naca(variant):
    return cases.case_naca(64, 32, order=2, limiter=abi.LIMITER_VENKATAKRISHNAN, rk_stages=3, local_dt=1,
                           kernel_variant=variant)


@pytest.mark.parametrize("variant", [1, 3])
def test_reads_between_steps(variant):
    m, bcs, p, ic, _ = _naca(variant)
    ref = FVKSolver(m, bcs, p, ic)
    r_ref = ref.step(12)
    ref.synchronize()
    U_ref = ref.field()
    R_ref = ref.residual_eval()

    g = FVKSolver(m, bcs, p, ic)
    rows = []
    for k in range(12):
        rows.append(g.step(1)[0])
        assert np.array_equal(g.residual(), rows[-1])
        if k % 3 == 2:
            g.field()  # a read while the step's last stages may still run
    assert np.array_equal(np.array(rows), r_ref)
    assert np.array_equal(g.field(), U_ref)
    assert np.array_equal(g.residual_eval(), R_ref)


def test_write_after_step():
    """set_field right after a step must not be overwritten by the stages
    still in flight."""
    m, bcs, p, ic, _ = _naca(3)
    g = FVKSolver(m, bcs, p, ic)
    g.step(5)
    U5 = g.field()
    g.step(1)
    g.set_field(U5)
    assert np.array_equal(g.field(), U5)
    r_a = g.step(4)
    g.set_field(U5)
    r_b = g.step(4)
    assert np.array_equal(r_a, r_b)


@pytest.mark.parametrize("variant", [1, 3])
def test_viscous_reads_between_steps(variant):
    """The SA solver, split passes (1) and the fused tile pass (3): reads,
    a residual evaluation and a write between single steps give the bits of
    an uninterrupted run."""
    m, bcs, p, ic, U0 = cases.case_rans_wavy(abi.EQ_RANS_SA, rk_stages=3, local_dt=1)
    p.kernel_variant = variant
    ref = FVKSolver(m, bcs, p, ic)
    g = FVKSolver(m, bcs, p, ic)
    if U0 is not None:
        ref.set_field(U0)
        g.set_field(U0)
    r_ref = ref.step(10)
    U_ref = ref.field()
    R_ref = ref.residual_eval()
    rows = []
    for k in range(10):
        rows.append(g.step(1)[0])
        assert np.array_equal(g.residual(), rows[-1])
        if k % 3 == 1:
            g.field()  # a read while the step's last stages may still run
        if k == 5:
            U6, T6 = g.field()
            g.set_field(U6, T6)  # a write of the same state: nothing in flight may overwrite it
    assert np.array_equal(np.array(rows), r_ref)
    for a, b in zip(g.field(), U_ref):
        assert np.array_equal(a, b)
    assert np.array_equal(g.residual_eval(), R_ref)
