"""torch.autograd.Function wrapper of the SPAI-BiCGSTAB solve, used by the
tests to check the hand-written adjoint (spai_system_adjoint in the C ABI).

Forward: x = A^{-1} b through the library (FP64 SPAI-preconditioned
BiCGSTAB).  Backward: lambda = A^{-T} dL/dx through the library's adjoint
solve (BiCGSTAB preconditioned by M^T), dL/db = lambda and
dL/dA_ij = -lambda_i x_j on A's CSC pattern.  M only accelerates the solves;
the converged x does not depend on it, so no gradient flows through the SPAI
build.
"""
from __future__ import annotations

import torch

from . import System


class SpaiSolve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, values: torch.Tensor, b: torch.Tensor, system: System, rel_tol: float = 1e-12,
                max_iter: int = 5000):
        if not (values.is_cuda and b.is_cuda and values.dtype == torch.float64 and b.dtype == torch.float64):
            raise TypeError("values and b must be float64 CUDA tensors")
        values = values.contiguous()
        b = b.contiguous()
        torch.cuda.synchronize()
        system.set_values(values.data_ptr(), on_device=True)
        x = torch.zeros_like(b)
        torch.cuda.synchronize()
        rep = system.solve_device(b.data_ptr(), x.data_ptr(), rel_tol=rel_tol, max_iter=max_iter)
        if not rep["all_converged"]:
            raise RuntimeError(f"forward solve did not converge: {rep}")
        ctx.system = system
        ctx.rel_tol = rel_tol
        ctx.max_iter = max_iter
        ctx.save_for_backward(x)
        return x

    @staticmethod
    def backward(ctx, gx: torch.Tensor):
        (x,) = ctx.saved_tensors
        system: System = ctx.system
        gx = gx.contiguous()
        dA = torch.empty(system.nnz, dtype=torch.float64, device=x.device)
        db = torch.empty_like(x)
        torch.cuda.synchronize()
        rep = system.adjoint_device(x.data_ptr(), gx.data_ptr(), dA.data_ptr(), db.data_ptr(), rel_tol=ctx.rel_tol,
                                    max_iter=ctx.max_iter)
        if not rep["all_converged"]:
            raise RuntimeError(f"adjoint solve did not converge: {rep}")
        return dA, db, None, None, None


def spai_solve(values: torch.Tensor, b: torch.Tensor, system: System, rel_tol: float = 1e-12,
               max_iter: int = 5000) -> torch.Tensor:
    return SpaiSolve.apply(values, b, system, rel_tol, max_iter)
