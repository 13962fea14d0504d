"""torch.autograd.Function wrappers of the flux residual and the force
coefficients, for checking the hand-written adjoint (lskum_flux_vjp,
lskum_forces_grad) with torch's machinery, e.g. torch.autograd.gradcheck.

Tensors are double, shape [4, n] (prim rows rho, u1, u2, p), on the CPU;
the library computes on the GPU through the C ABI.  The local time step is
held at the solver's last func_delta value.
"""
import numpy as np
import torch


class FluxResidual(torch.autograd.Function):
    """prim -> flux_res of a Solver (forward: its kernels; backward: its
    discrete adjoint)."""

    @staticmethod
    def forward(ctx, prim, solver):
        p = prim.detach().cpu().numpy().astype(np.float64)
        out = solver.residual_of(p)
        ctx.solver = solver
        ctx.save_for_backward(prim)
        return torch.from_numpy(out.copy())

    @staticmethod
    def backward(ctx, grad_out):
        (prim,) = ctx.saved_tensors
        solver = ctx.solver
        solver.set_state(prim.detach().cpu().numpy().astype(np.float64))
        g = solver.flux_vjp(grad_out.detach().cpu().numpy().astype(np.float64))
        return torch.from_numpy(g), None


class ForceCoefficients(torch.autograd.Function):
    """prim -> (Cl, Cd, Cm) of a Solver's shape."""

    @staticmethod
    def forward(ctx, prim, solver):
        solver.set_state(prim.detach().cpu().numpy().astype(np.float64))
        cl, cd, cm = solver.compute_cl_cd_cm()
        ctx.grads = [torch.from_numpy(g) for g in solver.forces_grad()]
        return torch.tensor([cl, cd, cm], dtype=torch.float64)

    @staticmethod
    def backward(ctx, grad_out):
        g = sum(grad_out[j] * ctx.grads[j] for j in range(3))
        return g, None
