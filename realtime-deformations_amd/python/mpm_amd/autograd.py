"""torch.autograd over the C ABI's reverse mode.

    Rollout.apply(sim, dt, nsteps, every, x0, v0, E0) -> (x_T, v_T)

runs `nsteps` substeps of the context from (x0, v0, E0) (float32 CPU
tensors; x0, v0: n x 3, E0: n) with checkpoints every `every` substeps (0:
the library's choice, every substep when the records fit in device memory), and
backpropagates dL/dx_T, dL/dv_T into dL/dx0, dL/dv0, dL/dE0 through
mpm_adjoint_backward.  The context's initial volumes are computed from x0 on
every forward and held fixed in the derivative.  Used by the gradient checks
and the parameter fit; torch is imported only here.

A context holds one reverse-mode record at a time.  Backward of a Rollout
whose context has run another adjoint forward since (a second Rollout.apply
on the same context, or a direct adjoint_forward) raises instead of
returning the gradients of the newer trajectory.
"""
from __future__ import annotations

import numpy as np
import torch


class Rollout(torch.autograd.Function):
    @staticmethod
    def forward(ctx, sim, dt, nsteps, every, x0, v0, E0):
        x = x0.detach().cpu().numpy().astype(np.float32)
        v = v0.detach().cpu().numpy().astype(np.float32)
        e = E0.detach().cpu().numpy().astype(np.float32)
        m = getattr(sim, "rollout_mass")
        sim.set_particles(x, v, m, e)
        sim.init_volumes()
        sim.adjoint_forward(dt, nsteps, every)
        out = sim.particles()
        ctx.sim = sim
        ctx.generation = sim.adjoint_generation
        return torch.from_numpy(out["x"].copy()), torch.from_numpy(out["v"].copy())

    @staticmethod
    def backward(ctx, gx, gv):
        sim = ctx.sim
        if sim.adjoint_generation != ctx.generation:
            raise RuntimeError("Rollout.backward: the context ran another adjoint forward after this rollout; "
                               "its reverse-mode record now belongs to the newer trajectory (one record per "
                               "context: use one Simulation per live rollout)")
        gx = None if gx is None else gx.detach().cpu().numpy().astype(np.float32)
        gv = None if gv is None else gv.detach().cpu().numpy().astype(np.float32)
        gx0, gv0, gE0 = sim.adjoint_backward(gx, gv)
        return None, None, None, None, torch.from_numpy(gx0), torch.from_numpy(gv0), torch.from_numpy(gE0)
