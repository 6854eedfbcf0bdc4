"""Gradient harness over the C-ABI adjoint.

``rk3_function()`` returns a ``torch.autograd.Function`` whose forward runs
fixed-dt SSP-RK3 steps on the GPU through ``cfd_adjoint_forward`` and whose
backward pulls the cotangent of the final state back to the initial state
through ``cfd_adjoint_backward``.  Torch is only the harness (chain rule
through the initial condition and the objective, optimisers); the solver and
its adjoint are the HIP library.  Tensors crossing the boundary are host
float64 [cell][4] arrays.

One solver holds the checkpoints of its last forward: call backward before
the next forward on the same solver.
"""
import numpy as np

from . import cases


def rk3_function():
    import torch

    class RK3Steps(torch.autograd.Function):
        @staticmethod
        def forward(ctx, u0, solver, dt, nsteps, checkpoint_every):
            solver.set_state(u0.detach().cpu().numpy())
            solver.adjoint_forward(dt, nsteps, checkpoint_every)
            ctx.solver = solver
            return torch.from_numpy(solver.get_state()).to(u0.dtype)

        @staticmethod
        def backward(ctx, grad_out):
            g = ctx.solver.adjoint_backward(grad_out.detach().cpu().numpy().astype(np.float64))
            return torch.from_numpy(g), None, None, None, None

    return RK3Steps


def initial_state_torch(case, rho0):
    """Differentiable conservative state of an adjoint_shock case (torch)."""
    import torch
    r = rho0 * torch.from_numpy(case["rho_factor"])
    u = torch.from_numpy(case["u"])
    v = torch.from_numpy(case["v"])
    p = torch.from_numpy(case["p"])
    E = p / (cases.GAMMA - 1) + 0.5 * r * (u * u + v * v)
    return torch.stack([r, r * u, r * v, E], -1).reshape(-1, 4)


def window_pressure_objective_torch(cons, case):
    """J = integral of p^2 over the case's probe window (torch)."""
    n = case["rho0"].shape[0]
    c = cons.reshape(n, n, 4)[case["window"]]
    r = c[..., 0]
    p = (cases.GAMMA - 1) * (c[..., 3] - 0.5 * (c[..., 1] ** 2 + c[..., 2] ** 2) / r)
    return (p * p).sum() * case["h"] ** 2


def window_pressure_objective(cons, case):
    """Numpy value and gradient d J / d cons of the probe-window objective."""
    n = case["rho0"].shape[0]
    c = np.asarray(cons).reshape(n, n, 4)
    w = c[case["window"]]
    r, mx, my, E = w[..., 0], w[..., 1], w[..., 2], w[..., 3]
    g1 = cases.GAMMA - 1
    p = g1 * (E - 0.5 * (mx * mx + my * my) / r)
    h2 = case["h"] ** 2
    J = float((p * p).sum() * h2)
    grad = np.zeros_like(c)
    dp = 2 * p * h2
    gw = grad[case["window"]]
    gw[..., 0] = dp * g1 * 0.5 * (mx * mx + my * my) / (r * r)
    gw[..., 1] = -dp * g1 * mx / r
    gw[..., 2] = -dp * g1 * my / r
    gw[..., 3] = dp * g1
    grad[case["window"]] = gw
    return J, grad.reshape(-1, 4)
