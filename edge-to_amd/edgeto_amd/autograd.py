"""Compliance as a differentiable torch function with a hand-written adjoint.

Forward (TopOpt/_cuda.py:282-291): x~ = H x, E = clip(x~^p, void, 1), solve
A_hat(E) u = f with the GMG-PCG solver, c = f . u.
Backward: the problem is self-adjoint, so the adjoint state is u itself and
    dc/dx = H^T [ p clip(x~, void, 1)^(p-1) * (-u_e^T K0 u_e) ]
(TopOpt/_cuda.py:293-300), evaluated by two library kernels: the fused SIMP
sensitivity (element energy times the SIMP factor) and the exact filter
transpose.  No autograd tape is recorded through the solver.
"""
import torch

from . import _lib


class ComplianceFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, topopt, iteration=0):
        tp = topopt
        mm = tp.material_model
        xt = tp.filter_kernel.dot(x.detach().contiguous())
        E = mm(xt, iteration)
        tp.kernel.set_rho(E)
        U, residual = tp.solver.solve(tp.f, use_last=tp.reuse_sol)
        U = U.clone()  # the solver keeps its warm-start buffer
        out = torch.zeros(1, dtype=torch.float64, device=x.device)
        _lib.call("et_dot_f64", tp.f.numel(), _lib.ptr(tp.f), _lib.ptr(U), _lib.ptr(out), _lib.stream())
        ctx.save_for_backward(U, xt)
        ctx.tp = tp
        ctx.iteration = iteration
        ctx.residual = residual
        return out.reshape(())

    @staticmethod
    def backward(ctx, grad_c):
        U, xt = ctx.saved_tensors
        tp = ctx.tp
        mm = tp.material_model
        pen = mm.penalty_schedule(mm.penalty, ctx.iteration) if mm.penalty_schedule else mm.penalty
        dr = tp.kernel.sensitivity(U, xt, pen, mm.void)
        dx = tp.filter_kernel._rmatvec(dr)
        return dx * grad_c, None, None


def compliance(x, topopt, iteration=0):
    """c(x) for the problem held by a CuTopOpt instance; differentiable in x."""
    return ComplianceFunction.apply(x, topopt, iteration)


def energy_compliance(u, topopt):
    """2 f.u - u^T A_hat u: equals f.u at the solution and is stationary in u,
    so solve errors enter only quadratically (used for finite-difference checks)."""
    tp = topopt
    au = tp.kernel.dot(u)
    out = torch.zeros(2, dtype=torch.float64, device=u.device)
    _lib.call("et_dot_f64", u.numel(), _lib.ptr(tp.f), _lib.ptr(u), _lib.ptr(out[0:1]), _lib.stream())
    _lib.call("et_dot_f64", u.numel(), _lib.ptr(u), _lib.ptr(au), _lib.ptr(out[1:2]), _lib.stream())
    return float(2.0 * out[0].item() - out[1].item())
