"""Compliance as a differentiable torch function with a hand-written adjoint.

Forward (TopOpt/_cuda.py:282-291): x~ = H x, E = clip(x~^p, void, 1), solve
A_hat(E) u = f with the GMG-PCG solver, c = f . u.
Backward: the problem is self-adjoint, so the adjoint state is u itself and
    dc/dx = H^T [ p clip(x~, void, 1)^(p-1) * (-u_e^T K0 u_e) ]
(TopOpt/_cuda.py:293-300), evaluated by one library kernel (et_simp_adjoint_f64:
element energy times the SIMP factor, divided by the filter normaliser while
staged in LDS, then the transpose stencil).  No autograd tape is recorded
through the solver.
"""
import torch

from . import _lib


class ComplianceFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, topopt, iteration=0):
        tp = topopt
        mm = tp.material_model
        ctx.in_dtype = x.dtype
        # the path computes in FP64; other input dtypes are cast (and the gradient cast back)
        xt = tp.filter_kernel.dot(x.detach().to(torch.float64).contiguous())
        E = mm(xt, iteration)
        tp.kernel.set_rho(E)
        U, residual = tp.solver.solve(tp.f, use_last=tp.reuse_sol)
        U = U.clone()  # the solver keeps its warm-start buffer
        out = torch.zeros(1, dtype=torch.float64, device=x.device)
        _lib.call("et_dot_f64", tp.f.numel(), _lib.f64(tp.f), _lib.f64(U, tp.f.numel()), _lib.f64(out),
                  _lib.stream())
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
        if hasattr(tp.filter_kernel, "simp_adjoint"):
            dx = tp.filter_kernel.simp_adjoint(tp.kernel, U, xt, pen, mm.void)
        else:
            dx = tp.filter_kernel._rmatvec(tp.kernel.sensitivity(U, xt, pen, mm.void))
        return (dx * grad_c).to(ctx.in_dtype), None, None


def compliance(x, topopt, iteration=0):
    """c(x) for the problem held by a CuTopOpt instance; differentiable in x."""
    return ComplianceFunction.apply(x, topopt, iteration)


def energy_compliance(u, topopt):
    """2 f.u - u^T A_hat u: equals f.u at the solution and is stationary in u,
    so solve errors enter only quadratically (used for finite-difference checks)."""
    tp = topopt
    au = tp.kernel.dot(u)
    out = torch.zeros(2, dtype=torch.float64, device=u.device)
    n = u.numel()
    _lib.call("et_dot_f64", n, _lib.f64(tp.f, n), _lib.f64(u), _lib.f64(out[0:1]), _lib.stream())
    _lib.call("et_dot_f64", n, _lib.f64(u), _lib.f64(au, n), _lib.f64(out[1:2]), _lib.stream())
    return float(2.0 * out[0].item() - out[1].item())
