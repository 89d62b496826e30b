"""Gradients of a time-mean force objective with respect to the contact
parameters (k_n, k_t, gamma_n, mu), as a torch.autograd.Function over the
library's hand-written adjoint (sph_adjoint_forward / sph_adjoint_backward).

The simulation itself runs in the library on the GPU; torch only carries the
parameter tensor and the gradient.  The objective is

    J = 1/nsteps  sum_steps  sum_i  w_i . F_i,

e.g. the mean shear stress of a sheared layer (python/sphere.py:2718-2724,
divided by the area): w_i = (-1/(L0 L1), 0, 0) on the driven particles.
"""

from __future__ import annotations

import numpy as np
import torch

from .dem import DEM
from .spherebin import Spherebin

PARAMS = ("k_n", "k_t", "gamma_n", "mu")


def shear_stress_weights(sb: Spherebin) -> np.ndarray:
    """Weights of the mean shear stress: -1/(L0 L1) on the x force of every
    fixed-velocity particle moving in +x."""
    w = np.zeros((sb.np, 3))
    top = (np.asarray(sb.fixvel) > 0.0) & (np.asarray(sb.vel)[:, 0] > 0.0)
    w[top, 0] = -1.0 / (float(sb.L[0]) * float(sb.L[1]))
    return w


def theta_of(sb: Spherebin) -> torch.Tensor:
    """The parameter vector (k_n, k_t, gamma_n, mu) of an input."""
    return torch.tensor([sb.k_n, sb.k_t, sb.gamma_n, sb.mu_s], dtype=torch.float64)


class MeanForceObjective(torch.autograd.Function):
    """J(theta) for a fixed input, step count and weights; backward runs the
    adjoint over the checkpoints the forward run left in its handle."""

    @staticmethod
    def forward(ctx, theta, source, nsteps, weights, seg_len, precision):
        th = [float(v) for v in theta.detach().cpu().double().tolist()]
        dem = DEM(source, verbosity=0)
        try:
            if precision != "fp64":
                dem.set_precision(precision)
            dem.set_params(k_n=th[0], k_t=th[1], gamma_n=th[2], mu_s=th[3], mu_d=th[3])
            J = dem.adjoint_forward(nsteps, weights, seg_len)
        except Exception:
            dem.close()
            raise
        ctx.dem = dem
        return theta.new_tensor(J)

    @staticmethod
    def backward(ctx, grad_J):
        dem = ctx.dem
        try:
            g = dem.adjoint_backward()
        finally:
            dem.close()
            ctx.dem = None
        grad = torch.as_tensor(g, dtype=grad_J.dtype, device=grad_J.device) * grad_J
        return grad, None, None, None, None, None


def mean_shear_stress(theta: torch.Tensor, sb: Spherebin, nsteps: int, seg_len: int = 100,
                      precision: str = "fp64") -> torch.Tensor:
    """Mean shear stress over `nsteps` steps of `sb` with contact parameters
    theta = (k_n, k_t, gamma_n, mu); differentiable in theta."""
    return MeanForceObjective.apply(theta, sb, nsteps, shear_stress_weights(sb), seg_len, precision)


def objective_value(sb: Spherebin, theta, nsteps: int, weights, precision: str = "fp64") -> float:
    """J alone (for finite differences): a plain run with the objective
    accumulated, no gradient."""
    th = [float(v) for v in theta]
    dem = DEM(sb, verbosity=0)
    try:
        if precision != "fp64":
            dem.set_precision(precision)
        dem.set_params(k_n=th[0], k_t=th[1], gamma_n=th[2], mu_s=th[3], mu_d=th[3])
        return dem.adjoint_forward(nsteps, weights, seg_len=max(1, nsteps))
    finally:
        dem.close()
