"""torch.autograd binding of the hand-written adjoint (test binding).

GradientEnergy.apply(state0, lambda0) runs `nsteps` RK4 steps on a level of
the HIP library from the initial state `state0` (float64 tensor of shape
(4, nz, ny, nx): psi1, psi2, Pi1, Pi2) and returns the final-time total
gradient energy J = sum |D psi|^2 dx^3 / R_f^4.  backward() returns the
gradients with respect to state0 and lambda0 computed by
axs_adjoint_grad_energy (checkpointed forward in the level's precision,
FP64 adjoint), scaled by the incoming gradient.  All arithmetic runs in
libaxs_hip.so; torch only carries the tensors.
"""

from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import capi


class GradEnergyProblem:
    """A fixed level shape, background and step schedule."""

    def __init__(self, ctx: capi.Context, n, dx: float = 0.5,
                 precision: int = capi.AXS_FP64, bg=None, tau0: float = 1.0,
                 dt: float = 0.1, nsteps: int = 20, checkpoint_every: int = 20):
        self.level = capi.Level(ctx, n, dx, precision)
        self.bg = bg if bg is not None else capi.make_background(2.0)
        self.tau0, self.dt, self.nsteps = tau0, dt, nsteps
        self.checkpoint_every = checkpoint_every

    def forward_value(self, state0: np.ndarray, lambda0: float) -> float:
        self.level.upload(state0)
        J = C.c_double()
        capi.check(self.level.lib.axs_forward_grad_energy(
            self.level.h, C.byref(self.bg), lambda0, self.tau0, self.dt,
            self.nsteps, C.byref(J)))
        return J.value

    def value_and_grad(self, state0: np.ndarray, lambda0: float):
        self.level.upload(state0)
        J = C.c_double()
        gl = C.c_double()
        g = np.zeros(4 * self.level.ncell)
        capi.check(self.level.lib.axs_adjoint_grad_energy(
            self.level.h, C.byref(self.bg), lambda0, self.tau0, self.dt,
            self.nsteps, self.checkpoint_every, C.byref(J), capi.dptr(g),
            C.byref(gl)))
        return J.value, g.reshape(4, self.level.nz, self.level.ny,
                                  self.level.nx), gl.value


class GradientEnergy(torch.autograd.Function):
    @staticmethod
    def forward(ctx, state0: torch.Tensor, lambda0: torch.Tensor,
                problem: GradEnergyProblem):
        s = state0.detach().to("cpu", torch.float64).contiguous().numpy()
        J, g, gl = problem.value_and_grad(s, float(lambda0))
        ctx.save_for_backward(torch.from_numpy(g).to(state0.device),
                              torch.tensor(gl, dtype=torch.float64,
                                           device=lambda0.device))
        return torch.tensor(J, dtype=torch.float64, device=state0.device)

    @staticmethod
    def backward(ctx, gout):
        g, gl = ctx.saved_tensors
        return gout * g, gout * gl, None
