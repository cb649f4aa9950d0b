"""torch.autograd binding of the sticking adjoint.

    sim = GpuSimulation(); sim.load(model); sim.rebuild_accel()
    sim.adjoint_setup(group_of_facet, targets)        # low-flux model
    y = sticking_response(s, sim, n_des=10**7, seed=42)   # y[t], differentiable in s
    loss = (y * w).sum(); loss.backward()              # s.grad from the backward kernel

Forward: sets the group sticking values, resets the unit (seed), traces
n_des desorptions while the trace kernel accumulates, per target facet t,
J0[t] = sum of hit weights and J[t][g] = dJ0[t]/ds_g; returns
y[t] = scale[t] * J0[t] / n_des (hit-equivalents per desorbed particle,
times an optional per-target factor such as MoleculesPerTP / area).
Backward: grad[g] = sum_t (dL/dy[t] * scale[t]) * J[t][g] / n_des, contracted
on the device by mfg_adjoint_backward.

The Jacobian lives on the device until the next run of the same unit, so
backward must come before the next forward on that unit (checked).
"""
from __future__ import annotations

import numpy as np
import torch

from .simulation import GpuSimulation, MolflowGpuError


class StickingResponse(torch.autograd.Function):
    @staticmethod
    def forward(ctx, sticking: torch.Tensor, sim: GpuSimulation, n_des: int, seed: int, max_events: int,
                scale: torch.Tensor | None):
        s = sticking.detach().to("cpu", torch.float64).numpy()
        sim.adjoint_set_sticking(s)
        sim.reset(seed=seed)
        r = sim.run(int(n_des), max_events=int(max_events))
        if r.desorbed != n_des:
            raise MolflowGpuError(f"adjoint forward desorbed {r.desorbed} of {n_des}")
        j0, _ = sim.adjoint_read()
        sc = np.ones(len(j0)) if scale is None else scale.detach().to("cpu", torch.float64).numpy()
        ctx.sim = sim
        ctx.generation = sim.generation
        ctx.scale = sc
        y = torch.from_numpy(sc * j0 / float(n_des))
        return y.to(device=sticking.device, dtype=sticking.dtype)

    @staticmethod
    def backward(ctx, grad_out: torch.Tensor):
        sim = ctx.sim
        if sim.generation != ctx.generation:
            raise MolflowGpuError("the unit ran again after this forward: its Jacobian is gone")
        up = grad_out.detach().to("cpu", torch.float64).numpy() * ctx.scale
        g = sim.adjoint_backward(up)
        return torch.from_numpy(g).to(device=grad_out.device, dtype=grad_out.dtype), None, None, None, None, None


def sticking_response(sticking: torch.Tensor, sim: GpuSimulation, n_des: int, seed: int = 42,
                      max_events: int = 0, scale: torch.Tensor | None = None) -> torch.Tensor:
    """Differentiable per-target response to the group sticking values."""
    return StickingResponse.apply(sticking, sim, n_des, seed, max_events, scale)
