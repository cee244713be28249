"""The torch.autograd.Function over the device adjoint (kwavehip.autograd): its gradients are the
C ABI's kwh_sim_vjp results, and in float64 they match central differences of the loss."""
import numpy as np
import pytest
import torch

from kwavehip import (DTypePrecision, KWaveGrid, KWaveMedium, KWaveSensor, KWaveSource, RecordField, Simulation,
                      SimulationOptions, makeDisc, random_field)
from kwavehip.autograd import KSpaceFirstOrderPressure

from test_gpu_parity import _device  # noqa: F401  (module fixture: the library must load, a device must exist)

pytestmark = pytest.mark.gpu


def _case(n=48, nt=60, seed=21):
    grid = KWaveGrid(n, 1e-4, n, 1e-4)
    c = random_field((n, n), seed, 4.0, 1450.0, 1650.0)
    grid.setTime(nt, 0.3 * 1e-4 / float(c.max()))
    p0 = (makeDisc(n, n, radius=4) + 0.5 * makeDisc(n, n, cx=14, cy=30, radius=3)).astype(np.float32)
    mask = np.zeros((n, n), np.float32)
    mask[n - 10, 6:n - 6:3] = 1
    return grid, c, p0, KWaveSensor(mask=mask, record=RecordField.p)


def test_autograd_equals_vjp_fp32():
    grid, c, p0, sens = _case()
    opts = SimulationOptions(pmlSize=8, smoothP0=False)
    print("\n[case] autograd vs kwh_sim_vjp FP32 48^2", flush=True)
    tp0 = torch.tensor(p0, dtype=torch.float32, requires_grad=True)
    tc = torch.tensor(c, dtype=torch.float32, requires_grad=True)
    y = KSpaceFirstOrderPressure.apply(tp0, tc, grid, np.float32(1000.0), sens, opts, 16)
    w = torch.from_numpy(np.random.default_rng(5).standard_normal(tuple(y.shape)).astype(np.float32))
    (y * w).sum().backward()
    sim = Simulation(grid, KWaveMedium(c, np.float32(1000.0)), KWaveSource(p0=p0), sens, opts)
    sim.set_checkpointing(16)
    sim.run()
    assert np.array_equal(sim.outputs().p, y.detach().numpy())
    gp0, gc = sim.vjp(w.numpy(), checkpoint_every=16)
    sim.close()
    np.testing.assert_allclose(tp0.grad.numpy(), gp0, rtol=1e-6, atol=1e-6 * np.abs(gp0).max())
    np.testing.assert_allclose(tc.grad.numpy(), gc, rtol=1e-6, atol=1e-6 * np.abs(gc).max())


def test_autograd_c_gradient_central_differences_fp64():
    grid, c, p0, sens = _case(seed=23)
    opts = SimulationOptions(pmlSize=8, smoothP0=False, dtype=DTypePrecision.float64)
    rho0 = np.float32(1000.0)
    y_obs = torch.from_numpy(np.asarray(KSpaceFirstOrderPressure.apply(
        torch.tensor(p0, dtype=torch.float64), torch.tensor(c * np.float32(1.003), dtype=torch.float64), grid, rho0,
        sens, opts, 16).detach().numpy()))

    def loss(cc):
        y = KSpaceFirstOrderPressure.apply(torch.tensor(p0, dtype=torch.float64), cc, grid, rho0, sens, opts, 16)
        return 0.5 * ((y - y_obs) ** 2).sum()

    tc = torch.tensor(c, dtype=torch.float64, requires_grad=True)
    loss(tc).backward()
    g = tc.grad.numpy()
    keep = c < c.max() - 20.0                     # c_ref = max(c) is held fixed
    # <g, v> for random signs has the scale ||g|| on the kept cells: the error is measured against it
    scale = float(np.sqrt(np.sum(np.where(keep, g, 0.0) ** 2)))
    rng = np.random.default_rng(8)
    print("\n[case] autograd FP64 c-gradient vs central differences", flush=True)
    for _ in range(3):
        v = np.where(keep, np.sign(rng.standard_normal(c.shape)), 0.0)
        h = 0.125                                 # exact in float32: the descriptor's c is float32
        lp = float(loss(torch.tensor(c + h * v, dtype=torch.float64)))
        lm = float(loss(torch.tensor(c - h * v, dtype=torch.float64)))
        fd = (lp - lm) / (2 * h)
        an = float(np.sum(g * v))
        print(f"  fd={fd:.6e} adjoint={an:.6e} |fd-an|/||g||={abs(fd - an) / scale:.2e}", flush=True)
        assert abs(fd - an) <= 1e-5 * scale
