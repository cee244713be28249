"""Pins of the float64 adjoint oracle (oracle/kwave_adjoint.py) on the CPU: its forward equals the
main oracle's loop, the p0 part passes the dot-product test, and the sound-speed gradient
matches central finite differences (c_ref held fixed, as the gradient is defined)."""
import numpy as np
import pytest

from oracle import kwave_adjoint as A
from oracle import kwave_oracle as O
from kwavehip import KWaveGrid, KWaveMedium, KWaveSensor, KWaveSource, RecordField, SimulationOptions, makeDisc


def problem(cart=False, rho_map=False, nx=40, ny=48, nt=60, seed=3):
    rng = np.random.default_rng(seed)
    grid = KWaveGrid(nx, 1e-4, ny, 1e-4)
    c = (1500.0 + 80.0 * rng.standard_normal((nx, ny))).astype(np.float32)
    grid.setTime(nt, 0.25 * 1e-4 / float(c.max()))
    rho = (1000.0 + 30.0 * rng.standard_normal((nx, ny))).astype(np.float32) if rho_map else np.float32(1000.0)
    p0 = (makeDisc(nx, ny, radius=5) + 0.5 * makeDisc(nx, ny, radius=3, cx=12, cy=30)).astype(np.float32)
    if cart:
        ang = np.linspace(0, 2 * np.pi, 12, endpoint=False)
        mask = np.stack([13e-4 * np.cos(ang), 15e-4 * np.sin(ang)]).astype(np.float32)
    else:
        mask = np.zeros((nx, ny), np.float32)
        mask[30, 6:42:3] = 1
    opts = SimulationOptions(pmlSize=8, smoothP0=False)
    ops = A.LosslessOperators(grid, rho, mask, opts, float(c.max()))
    return grid, c, rho, p0, mask, opts, ops


@pytest.mark.parametrize("cart,rho_map", [(False, False), (True, True)])
def test_forward_equals_main_oracle(cart, rho_map):
    grid, c, rho, p0, mask, opts, ops = problem(cart, rho_map)
    y = A.forward(ops, c.astype(np.float64), p0.astype(np.float64))
    ref = O.kspace_first_order(grid, KWaveMedium(c, rho), KWaveSource(p0=p0), KWaveSensor(mask=mask), opts)["p"]
    assert y.shape == ref.shape
    assert np.abs(y - ref).max() <= 1e-12 * np.abs(ref).max()


@pytest.mark.parametrize("cart,rho_map", [(False, False), (True, True)])
def test_p0_dot_product(cart, rho_map):
    grid, c, rho, p0, mask, opts, ops = problem(cart, rho_map)
    rng = np.random.default_rng(11)
    v = rng.standard_normal(ops.shape)
    w = rng.standard_normal((ops.sampler.n_points, ops.nt))
    Jv = A.forward(ops, c, v)
    gp0, _ = A.vjp(ops, c, np.zeros(ops.shape), w)
    lhs, rhs = float(np.sum(Jv * w)), float(np.sum(v * gp0))
    assert abs(lhs - rhs) <= 1e-10 * abs(lhs)


@pytest.mark.parametrize("cart,rho_map", [(False, False), (True, True)])
def test_c_gradient_central_differences(cart, rho_map):
    grid, c, rho, p0, mask, opts, ops = problem(cart, rho_map)
    c = c.astype(np.float64)
    y_obs = A.forward(ops, c * 1.01, p0)
    loss = lambda cc: 0.5 * float(np.sum((A.forward(ops, cc, p0) - y_obs) ** 2))
    y = A.forward(ops, c, p0)
    gp0, gc = A.vjp(ops, c, p0, y - y_obs)
    rng = np.random.default_rng(5)
    for k in range(3):
        v = rng.standard_normal(ops.shape)
        h = 1e-2
        fd = (loss(c + h * v) - loss(c - h * v)) / (2 * h)
        an = float(np.sum(gc * v))
        assert abs(fd - an) <= 1e-6 * abs(an), (k, fd, an)
    v = rng.standard_normal(ops.shape)
    fd = (0.5 * np.sum((A.forward(ops, c, p0 + v) - y_obs) ** 2) -
          0.5 * np.sum((A.forward(ops, c, p0 - v) - y_obs) ** 2)) / 2.0
    an = float(np.sum(gp0 * v))
    assert abs(fd - an) <= 1e-9 * abs(an)
