"""A torch.autograd.Function over the PDHG solve, for checking its gradients
with torch's own machinery.

forward solves the standard-form LP (K from a template LP, c, q and the
bounds from the inputs) on the device and returns the unscaled primal-dual
pair; backward runs the device adjoint (implicit differentiation of the PDHG
fixed point on the final active set) with the incoming gradients as seeds.
Tensors are CPU float64: the LP data crosses to the device inside the solve.
"""
from __future__ import annotations

import torch

from .api import LP, pdhg_adjoint, solve_pdhg


class PdhgSolution(torch.autograd.Function):
    @staticmethod
    def forward(ctx, c, q, lower, upper, template: LP, context, tolerance: float):
        e = template.export()
        lp = LP.from_standard(template.rows, template.cols, template.num_equalities, e["start"],
                              e["index"], e["value"], c.detach().numpy(), q.detach().numpy(),
                              lower.detach().numpy(), upper.detach().numpy(),
                              template.objective_scale, template.objective_offset)
        r = solve_pdhg(context, lp, tolerance=tolerance)
        if r.status != "optimal":
            raise RuntimeError(f"PDHG solve ended with status {r.status!r}")
        x, y = torch.from_numpy(r.x), torch.from_numpy(r.y)
        ctx.lp, ctx.context = lp, context
        ctx.save_for_backward(x, y)
        return x, y

    @staticmethod
    def backward(ctx, grad_x, grad_y):
        x, y = ctx.saved_tensors
        a = pdhg_adjoint(ctx.context, ctx.lp, x.numpy(), y.numpy(),
                         w_x=grad_x.detach().numpy(), w_y=grad_y.detach().numpy())
        return (torch.from_numpy(a.grad_c), torch.from_numpy(a.grad_q),
                torch.from_numpy(a.grad_lower), torch.from_numpy(a.grad_upper), None, None, None)


def pdhg_solution(c, q, lower, upper, template: LP, context, tolerance: float = 1e-8):
    """(x, y) = the PDHG solution of the LP with K from `template`, differentiable
    in c, q, lower and upper."""
    return PdhgSolution.apply(c, q, lower, upper, template, context, tolerance)
