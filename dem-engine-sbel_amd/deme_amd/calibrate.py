"""Calibration harness: per-material E, nu, CoR, mu and initial velocities fitted to observed final positions, with
gradients from the device adjoint (deme_adjoint_forward / deme_adjoint_backward) behind torch.autograd.

    runner = TrajectoryRunner(scene, n_steps=200, window=20)
    v0, E, nu, CoR, mu = runner.parameters(requires_grad=True)
    X = final_positions(runner, v0, E, nu, CoR, mu)          # (n, 3) float64, voxel-grid frame
    loss = ((X - X_target) ** 2).sum(1).mean()
    loss.backward()                                           # v0.grad, E.grad, ... from the HIP adjoint kernels

The torch tensors live on the host; the simulation and its adjoint run on the GPU inside the library.
"""
from __future__ import annotations

import numpy as np
import torch

from .abi import DeviceSystem, positions_f64


class TrajectoryRunner:
    """One device context for a scene of single-sphere owners; runs the checkpointed forward for given parameters."""

    def __init__(self, scene, n_steps: int, window: int):
        self.scene = scene
        self.n_steps = int(n_steps)
        self.window = int(window)
        self.system = DeviceSystem(scene)
        self.n = scene.n_clumps

    def parameters(self, requires_grad=True):
        st, m = self.scene.state, self.scene.materials
        v0 = np.stack([st["vX"], st["vY"], st["vZ"]], 1)[:self.n].astype(np.float64)
        out = [torch.tensor(v0)] + [torch.tensor(m[k].astype(np.float64)) for k in ("E", "nu", "CoR", "mu")]
        for t in out:
            t.requires_grad_(requires_grad)
        return out

    def run(self, v0, E, nu, CoR, mu) -> np.ndarray:
        mats = self.scene.materials.copy()
        for k, t in (("E", E), ("nu", nu), ("CoR", CoR), ("mu", mu)):
            mats[k] = np.asarray(t, np.float64).astype(np.float32)
        st = {k: v.copy() for k, v in self.scene.state.items()}
        v = np.asarray(v0, np.float64).astype(np.float32)
        st["vX"][:self.n], st["vY"][:self.n], st["vZ"][:self.n] = v[:, 0], v[:, 1], v[:, 2]
        self.system.upload_materials(mats)
        self.system.set_state(st)
        self.system.clear_contacts()
        self.system.adjoint_forward(self.n_steps, self.window)
        s = self.system.get_state()
        return positions_f64(self.scene.params, s["voxelID"], s["locX"], s["locY"], s["locZ"])[:self.n]


class _FinalPositions(torch.autograd.Function):
    @staticmethod
    def forward(ctx, v0, E, nu, CoR, mu, runner):
        ctx.runner = runner
        X = runner.run(v0.detach().numpy(), E.detach().numpy(), nu.detach().numpy(), CoR.detach().numpy(),
                       mu.detach().numpy())
        return torch.from_numpy(X)

    @staticmethod
    def backward(ctx, gX):
        r = ctx.runner
        gV, gM = r.system.adjoint_backward(gX.detach().to(torch.float64).numpy(), None)
        gV = torch.from_numpy(gV[:r.n].astype(np.float64))
        gM = torch.from_numpy(gM.astype(np.float64))
        return gV, gM[:, 0], gM[:, 1], gM[:, 2], gM[:, 3], None


def final_positions(runner: TrajectoryRunner, v0, E, nu, CoR, mu) -> torch.Tensor:
    """Final sphere positions after runner.n_steps steps, differentiable in v0 (n x 3) and the per-material E, nu, CoR,
    mu through the device adjoint."""
    return _FinalPositions.apply(v0, E, nu, CoR, mu, runner)


def mse_to_target(runner: TrajectoryRunner, target: np.ndarray, v0, E, nu, CoR, mu) -> torch.Tensor:
    """Mean squared distance of the final positions to a target state (n x 3)."""
    X = final_positions(runner, v0, E, nu, CoR, mu)
    return ((X - torch.as_tensor(target, dtype=torch.float64)) ** 2).sum(1).mean()
