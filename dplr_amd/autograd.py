"""torch.autograd wrappers of the adjoint kernels, for checking gradients
(finite differences, dense references) -- the kernels themselves are the
C-ABI calls dplrx_solve (transposed) and dplrx_grad_outer.

PlrSolveFn:  y = A^{-1} b for a factored Plan.  backward gives
             bbar = A^{-T} ybar, and leaves the gradient with respect to A's
             factors, G = -A^{-T} ybar y^T projected onto the structure, in
             plan.last_factor_grad = (dU, dV, dS, dD) (plan layouts).
PlrApplyFn:  y = A x.  backward gives xbar = A^T ybar and the factor
             gradient of G = ybar x^T.
float32 inputs run the FP32 path (the plan's FP32 factors from
factor_single("f32"): dplrx_solve_single, dplrx_gepdmm_single and
dplrx_grad_outer_f32); float64 inputs the FP64 one.
"""
import torch


class PlrSolveFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, b, plan):
        y = torch.zeros_like(b)
        if b.dtype == torch.float32:
            plan.solve_single("N", b.contiguous(), y)
        else:
            plan.solve("N", b.contiguous(), y)
        ctx.plan = plan
        ctx.save_for_backward(y)
        return y

    @staticmethod
    def backward(ctx, ybar):
        (y,) = ctx.saved_tensors
        plan = ctx.plan
        lam = torch.zeros_like(ybar)
        if ybar.dtype == torch.float32:
            plan.solve_single("T", ybar.contiguous(), lam)
            plan.last_factor_grad = plan.grad_outer_f32(-1.0, lam, y)
        else:
            plan.solve("T", ybar.contiguous(), lam)
            plan.last_factor_grad = plan.grad_outer(-1.0, lam, y)
        return lam, None


class PlrApplyFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, plan):
        y = torch.zeros_like(x)
        if x.dtype == torch.float32:
            plan.gepdmm_single("N", 1.0, x.contiguous(), 0.0, y)
        else:
            plan.gepdmm("N", 1.0, x.contiguous(), 0.0, y)
        ctx.plan = plan
        ctx.save_for_backward(x)
        return y

    @staticmethod
    def backward(ctx, ybar):
        (x,) = ctx.saved_tensors
        plan = ctx.plan
        xbar = torch.zeros_like(ybar)
        if ybar.dtype == torch.float32:
            plan.gepdmm_single("T", 1.0, ybar.contiguous(), 0.0, xbar)
            plan.last_factor_grad = plan.grad_outer_f32(1.0, ybar.contiguous(), x)
        else:
            plan.gepdmm("T", 1.0, ybar.contiguous(), 0.0, xbar)
            plan.last_factor_grad = plan.grad_outer(1.0, ybar.contiguous(), x)
        return xbar, None


def plr_solve(b, plan):
    """y = A^{-1} b (rows of b are right-hand sides), differentiable in b."""
    return PlrSolveFn.apply(b, plan)


def plr_apply(x, plan):
    """y = A x (rows of x are vectors), differentiable in x."""
    return PlrApplyFn.apply(x, plan)
