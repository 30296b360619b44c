"""The torch.autograd.Function wrapper of the residual and its discrete
adjoint (a checking harness; the compute is the HIP library's)."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")  # imported at collection, outside any test's time limit

from finitevolume_amd import abi  # noqa: E402
from finitevolume_amd.torch_adjoint import ResidualFunction  # noqa: E402
from test_gpu_adjoint import H, developed_state, gpu_vjp, naca_case  # noqa: E402

pytestmark = pytest.mark.gpu


def test_torch_autograd_function():
    """backward() of the wrapper returns the library's VJP, and central
    differences (step 1e-7) of forward() along joint state and node
    directions agree with it."""
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
    rng = np.random.default_rng(4)
    for _ in range(3):
        V = rng.standard_normal(U.shape) * np.abs(U).mean(axis=0)
        W = rng.standard_normal(m.node_xy.shape) * 1e-2
        def J(h):
            with torch.no_grad():
                r = ResidualFunction.apply(torch.from_numpy(U + h * V), torch.from_numpy(m.node_xy + h * W),
                                           (m, bcs, p, ic))
            return float((r * torch.from_numpy(lam)).sum())
        fd = (J(H) - J(-H)) / (2 * H)
        ad = float(np.sum(Ub * V) + np.sum(Xb * W))
        scale = float(np.sum(np.abs(Ub * V)) + np.sum(np.abs(Xb * W)))
        assert abs(fd - ad) <= 1e-6 * scale, (fd, ad)
