"""torch.autograd.Function wrapper of the checkpointed forward and the adjoint
(fs_forward_checkpointed / fs_step_adjoint), used as a gradient-check harness.

`FlipRollout.apply(vel0, gravity_y, pos0, params, steps, every)` runs `steps`
substeps of the fused step from (pos0, vel0) with gravity (0, gravity_y, 0)
and returns the final particle positions and velocities; `backward` runs the
reverse pass on the GPU and returns dL/d(vel0) and dL/d(gravity_y).  Tensors
live on the host (the step keeps its own device buffers); the harness exists
to compare the adjoint with finite differences and with the FP64 restatement,
not to train anything.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

import fluidstep as fsm


def _params_with_gravity(params: fsm.Params, gy: float) -> fsm.Params:
    p = fsm.Params()
    C.pointer(p)[0] = params
    p.gravity[0], p.gravity[1], p.gravity[2] = 0.0, float(gy), 0.0
    return p


def rollout(params, pos0, vel0, gy, steps, every=10):
    """Forward only: final (pos, vel) as float32 arrays, plus the tank (open)."""
    t = fsm.FluidStep(_params_with_gravity(params, gy), pos0, vel0)
    t.forward_checkpointed(steps, every)
    pos, vel = t.particles()
    return pos, vel, t


class FlipRollout(torch.autograd.Function):
    @staticmethod
    def forward(ctx, vel0, gravity_y, pos0, params, steps, every):
        v = vel0.detach().cpu().numpy().astype(np.float32)
        gy = float(gravity_y.detach().cpu())
        pos, vel, tank = rollout(params, pos0, v, gy, steps, every)
        ctx.tank = tank
        return torch.from_numpy(pos.copy()), torch.from_numpy(vel.copy())

    @staticmethod
    def backward(ctx, gpos, gvel):
        tank = ctx.tank
        gp = None if gpos is None else gpos.detach().cpu().numpy()
        gv = None if gvel is None else gvel.detach().cpu().numpy()
        _, gvel0, ggrav = tank.adjoint(gp, gv)
        tank.close()
        ctx.tank = None
        return (torch.from_numpy(gvel0), torch.tensor(ggrav[1], dtype=torch.float64), None, None, None, None)


def centroid_loss(pos: torch.Tensor, target=(0.7, 0.3, 0.5)) -> torch.Tensor:
    """Squared distance of the particle centroid to `target` (in float64)."""
    c = pos.to(torch.float64).mean(0)
    return ((c - torch.tensor(target, dtype=torch.float64)) ** 2).sum()
