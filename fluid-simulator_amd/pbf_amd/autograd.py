"""torch.autograd wrapper of the adjoint (gradient checks and optimisation).

`rollout(v0, viscosity, vorticity, scene, frames, checkpoint_every)` runs
the library's solver on `scene` from its positions with initial velocities
`v0` and returns the final positions; its backward is the library's adjoint
(pbf_adjoint_*), with neighbour lists and discrete decisions held fixed.
Tensors stay on the host: torch is the differentiation front end only, the
step and its adjoint run in the HIP library.
"""
from __future__ import annotations

import copy

import numpy as np
import torch

from . import PBFAdjoint, PBFSolver


class Rollout(torch.autograd.Function):
    @staticmethod
    def forward(ctx, v0, viscosity, vorticity, scene, frames, checkpoint_every):
        params = copy.copy(scene.params)
        params.xsphViscosity = float(viscosity)
        params.vorticityStrength = float(vorticity)
        vel = np.zeros((scene.n, 4), np.float32)
        vel[:, :3] = v0.detach().cpu().numpy()
        solver = PBFSolver()
        solver.initialize(scene.n, scene.minBounds, scene.maxBounds, params)
        if scene.spheres:
            solver.setSpheres(scene.spheres)
        if scene.boxes:
            solver.setBoxes(scene.boxes)
        if scene.planes:
            solver.setPlanes(scene.planes)
        solver.setParticles(scene.positions, vel)
        adj = PBFAdjoint(solver)
        adj.forward(int(frames), int(checkpoint_every))
        ctx.adj = adj
        ctx.n = scene.n
        return torch.from_numpy(solver.copyPositionsToHost()[:, :3].astype(np.float64))

    @staticmethod
    def backward(ctx, gx):
        xbar = np.zeros((ctx.n, 4), np.float32)
        xbar[:, :3] = gx.detach().cpu().numpy()
        _, v0b, cb, eb = ctx.adj.backward(xbar, None)
        return (torch.from_numpy(v0b[:, :3].astype(np.float64)), torch.tensor(cb, dtype=torch.float64),
                torch.tensor(eb, dtype=torch.float64), None, None, None)


def rollout(v0, viscosity, vorticity, scene, frames: int, checkpoint_every: int = 10):
    """Final positions (n, 3) of `frames` frames from (scene.positions, v0)."""
    return Rollout.apply(v0, viscosity, vorticity, scene, frames, checkpoint_every)
