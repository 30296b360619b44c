"""The torch.autograd.Function wrapper of the residual and its discrete
adjoint (a checking harness; the compute is the HIP library's)."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")  # imported at collection, outside any test's time limit

from finitevolume_amd import abi  # noqa: E402
from finitevolume_amd.torch_adjoint import ResidualFunction  # noqa: E402
from cases import adjoint_naca as naca_case, developed_state  # noqa: E402
from oracle.pyoracle import residual_jvp  # noqa: E402
from test_gpu_adjoint import gpu_vjp  # noqa: E402

pytestmark = pytest.mark.gpu


def test_torch_autograd_function():
    """backward() of the wrapper returns the library's VJP; forward()
    returns the residual, and along joint state and node directions the
    backward product agrees with the oracle's forward-mode derivative
    (dot-product identity, 1e-10 relative)."""
    case = naca_case(abi.FLUX_HLLC, 2)
    m, bcs, p, ic, _ = case
    U = developed_state(case)
    lam = np.random.default_rng(0).standard_normal(U.shape)
    Ut = torch.tensor(U, dtype=torch.float64, requires_grad=True)
    Xt = torch.tensor(m.node_xy, dtype=torch.float64, requires_grad=True)
    R = ResidualFunction.apply(Ut, Xt, (m, bcs, p, ic))
    (R * torch.from_numpy(lam)).sum().backward()
    Ub, Xb = gpu_vjp(case, U, lam)
    assert np.array_equal(Ut.grad.numpy(), Ub)
    assert np.array_equal(Xt.grad.numpy(), Xb)
    with torch.no_grad():
        r = ResidualFunction.apply(torch.from_numpy(U), torch.from_numpy(m.node_xy), (m, bcs, p, ic)).numpy()
    rng = np.random.default_rng(4)
    for _ in range(3):
        V = rng.standard_normal(U.shape) * np.abs(U).mean(axis=0)
        W = rng.standard_normal(m.node_xy.shape) * 1e-2
        R0, JV = residual_jvp(m, bcs, p, U, V, W)
        assert np.abs(r - R0).max() <= 1e-10 * np.abs(R0).max()
        exact = float(np.sum(lam * JV))
        ad = float(np.sum(Ub * V) + np.sum(Xb * W))
        assert abs(exact - ad) <= 1e-10 * abs(exact), (exact, ad)
