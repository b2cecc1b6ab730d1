"""Design-to-compliance chain as a ``torch.autograd.Function``.

forward : x -> filter -> Heaviside -> protected clamp -> SIMP -> solve -> c
backward: dc/dx = H^T((dc/drho * drho/drho_f)/Hs) with
          dc/drho = -p (1 - Emin) rho^(p-1) u_e.KE.u_e (compliance is
          self-adjoint, so the adjoint field is -u).  Element energies are
          recomputed from the saved displacement in the backward kernel rather
          than stored.
This is the gradient top3d uses (pytopo3d/core/optimizer.py:596-625), packaged
for autograd; all arithmetic runs in the HIP kernels.
"""

from __future__ import annotations

import torch

from .optimizer import SIMPProblem


class ComplianceFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x: torch.Tensor, problem: SIMPProblem, penal: float, beta: float,
                eta: float, solve_kw: dict):
        xc = x.detach().contiguous().to(torch.float64)
        _, rho, drho = problem.physical_density(xc, beta, eta)
        U = problem.solve(rho, penal, **solve_kw).clone()
        c = problem.sensitivities(U, rho, None, penal)
        ctx.problem = problem
        ctx.penal = penal
        ctx.save_for_backward(U, rho, drho)
        return torch.tensor(c, dtype=torch.float64, device=x.device)

    @staticmethod
    def backward(ctx, grad_c: torch.Tensor):
        U, rho, drho = ctx.saved_tensors
        problem = ctx.problem
        dc_dx = problem.new_elem()
        problem.plan.energy(U, rho, ctx.penal, drho=drho, dc_dx=dc_dx, sync=False)
        return grad_c * dc_dx, None, None, None, None, None


def compliance(x: torch.Tensor, problem: SIMPProblem, penal: float = 3.0, beta: float = 1.0,
               eta: float = 0.5, **solve_kw) -> torch.Tensor:
    """Differentiable compliance of the design vector ``x`` (element order)."""
    return ComplianceFunction.apply(x, problem, float(penal), float(beta), float(eta), solve_kw)
