"""torch.autograd.Function over the residual and its discrete adjoint, for
checking the adjoint against autograd-style use (gradcheck-like tests); the
residual and the VJP themselves run in the HIP library.

    R = ResidualFunction.apply(U, X, case)

U: [n_cells, 4] float64 conserved state, X: [n_nodes, 2] float64 node
coordinates, `case` = (mesh, bcs, params, ic) with the mesh topology.  The
forward builds a solver on the mesh moved to X (cell volumes and centroids
recomputed), evaluates R(U, X); the backward returns (dR/dU)^T g and
(dR/dX)^T g from fvk_residual_vjp.  Tensors cross as host float64 arrays:
torch is the harness here, not the compute path.
"""
from __future__ import annotations

import numpy as np
import torch

from .solver import FVKSolver


class ResidualFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, U: torch.Tensor, X: torch.Tensor, case):
        mesh, bcs, params, ic = case
        m = mesh.with_nodes(X.detach().cpu().numpy().astype(np.float64))
        s = FVKSolver(m, bcs, params, ic)
        s.set_field(U.detach().cpu().numpy().astype(np.float64))
        R = s.residual_eval()
        ctx.solver = s
        return torch.from_numpy(R).to(U.device)

    @staticmethod
    def backward(ctx, gR: torch.Tensor):
        s = ctx.solver
        Ub, Xb = s.residual_vjp(np.ascontiguousarray(gR.detach().cpu().numpy(), dtype=np.float64), want_x=True)
        return torch.from_numpy(Ub).to(gR.device), torch.from_numpy(Xb).to(gR.device), None
