"""torch.autograd.Function over the device's checkpointed adjoint (kwh_sim_vjp).

`KSpaceFirstOrderPressure.apply(p0, c, grid, rho0, sensor, options, checkpoint_every)` runs the
lossless linear k-space loop on the HIP device (KSpaceFirstOrder.swift:557-761 with p0 and a
sound-speed map) and returns the recorded pressure series [nSensor, nt]; `backward` returns
dL/dp0 and dL/dc from the hand-written reverse-time kernels (checkpoints every
`checkpoint_every` steps, recomputation between them). Tensors only carry the values across the
boundary: the simulation reads host copies of p0 and c and the gradients come back as tensors on
p0's device and dtype. The forward's simulation stays alive until backward has run.
"""
from __future__ import annotations

import numpy as np
import torch

from .api import KWaveMedium, KWaveSource, RecordField, Simulation


def _host(t: torch.Tensor) -> np.ndarray:
    return t.detach().to("cpu").numpy()


class KSpaceFirstOrderPressure(torch.autograd.Function):
    """y = sensor pressure series of the p0 problem as a function of (p0, c)."""

    @staticmethod
    def forward(ctx, p0, c, grid, rho0, sensor, options, checkpoint_every=32):
        if not (sensor.record & RecordField.p):
            raise ValueError("the sensor must record p")
        sim = Simulation(grid, KWaveMedium(_host(c).astype(np.float32), rho0),
                         KWaveSource(p0=_host(p0).astype(np.float32)), sensor, options)
        sim.set_checkpointing(int(checkpoint_every))
        sim.run()
        y = sim.outputs().p
        ctx.sim = sim
        ctx.every = int(checkpoint_every)
        return torch.from_numpy(np.ascontiguousarray(y)).to(device=p0.device, dtype=p0.dtype)

    @staticmethod
    def backward(ctx, grad_y):
        sim = ctx.sim
        try:
            gp0, gc = sim.vjp(np.ascontiguousarray(_host(grad_y)), checkpoint_every=ctx.every)
        finally:
            sim.close()
            ctx.sim = None
        dev, dt = grad_y.device, grad_y.dtype
        return (torch.from_numpy(np.ascontiguousarray(gp0)).to(device=dev, dtype=dt),
                torch.from_numpy(np.ascontiguousarray(gc)).to(device=dev, dtype=dt),
                None, None, None, None, None)
