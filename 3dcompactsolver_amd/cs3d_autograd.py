"""torch.autograd binding of the discrete adjoint (test and scripting aid).

`KineticEnergy.apply(U0, V0, W0)` runs the GPU solver forward from the
initial velocities (density and pressure held by the closure's solver
configuration) and returns the final volume-averaged kinetic energy
J = <rho |u|^2 / 2>; its backward returns dJ/dU0, dJ/dV0, dJ/dW0 from the
discrete adjoint of the steps taken (cs3d_adjoint_ke_gradient).  Tensors
cross the C ABI as host float64 arrays; torch only carries the autograd
graph here.
"""
import numpy as np
import torch

import cs3d


class KineticEnergy(torch.autograd.Function):
    @staticmethod
    def forward(ctx, U0, V0, W0, cfg, rho0, p0, nsteps, checkpoint):
        g = cs3d.Solver(cfg)
        try:
            vel = [t.detach().cpu().double().numpy() for t in (U0, V0, W0)]
            g.set_initial_conditions(rho0, vel[0], vel[1], vel[2], p0)
            J, gU, gV, gW = g.ke_gradient(nsteps, checkpoint)
        finally:
            g.close()
        ctx.save_for_backward(*[torch.from_numpy(a).to(U0.dtype) for a in (gU, gV, gW)])
        return torch.tensor(J, dtype=U0.dtype)

    @staticmethod
    def backward(ctx, gJ):
        gU, gV, gW = ctx.saved_tensors
        return gJ * gU, gJ * gV, gJ * gW, None, None, None, None, None


def kinetic_energy(U0, V0, W0, cfg, rho0, p0, nsteps, checkpoint=8):
    """J(U0, V0, W0) after nsteps RK4 steps, differentiable through the
    discrete adjoint.  U0, V0, W0: torch tensors shaped (Nz, Ny, Nx);
    rho0, p0: numpy arrays of the same shape."""
    return KineticEnergy.apply(U0, V0, W0, cfg, np.asarray(rho0, dtype=np.float64),
                               np.asarray(p0, dtype=np.float64), nsteps, checkpoint)
