"""GPU adjoint (kwh_sim_vjp) against the float64 adjoint oracle and central finite differences.

Loss L = 1/2 ||y - y_obs||^2 of the recorded pressure series; gradients with respect to p0 and the
sound speed at fixed c_ref = max(c). FP64 on the GPU matches the oracle to 1e-9; FP32 to the
tolerances stated per test.
"""
import ctypes as C
import math

import numpy as np
import pytest

from oracle import kwave_adjoint as A
from kwavehip import (AbsorptionMode, DTypePrecision, KWaveGrid, KWaveMedium, KWaveSensor, KWaveSource, RecordField,
                      Simulation, SimulationOptions, makeDisc, random_field)
from kwavehip import _lib as L

pytestmark = pytest.mark.gpu
F32, F64 = DTypePrecision.float32, DTypePrecision.float64


@pytest.fixture(scope="module", autouse=True)
def _device():
    lib = L.load()
    h = C.c_void_p()
    if lib.kwh_ctx_create(0, C.byref(h)) != 0:
        pytest.skip("no HIP device")
    lib.kwh_ctx_destroy(h)


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return math.sqrt(np.mean((a - b) ** 2)) / math.sqrt(np.mean(b * b))


def case_2d(n=(48, 40), nt=70, cart=False, rho_map=False, seed=2):
    nx, ny = n
    grid = KWaveGrid(nx, 1e-4, ny, 1e-4)
    c = random_field((nx, ny), seed, 4.0, 1450.0, 1650.0)
    grid.setTime(nt, 0.3 * 1e-4 / float(c.max()))
    rho = random_field((nx, ny), seed + 1, 4.0, 950.0, 1050.0) if rho_map else np.float32(1000.0)
    p0 = (makeDisc(nx, ny, radius=4) + 0.6 * makeDisc(nx, ny, cx=14, cy=12, radius=3)).astype(np.float32)
    if cart:
        ang = np.linspace(0, 2 * np.pi, 16, endpoint=False)
        mask = np.stack([15e-4 * np.cos(ang), 13e-4 * np.sin(ang)]).astype(np.float32)
    else:
        mask = np.zeros((nx, ny), np.float32)
        mask[nx - 12, 4:ny - 4:2] = 1
    opts = SimulationOptions(pmlSize=8, smoothP0=False)
    return grid, c, rho, p0, mask, opts


def gpu_forward(grid, c, rho, p0, mask, opts, dtype):
    sim = Simulation(grid, KWaveMedium(c, rho), KWaveSource(p0=p0), KWaveSensor(mask=mask, record=RecordField.p),
                     SimulationOptions(pmlSize=opts.pmlSize, smoothP0=False, dtype=dtype))
    sim.run()
    y = sim.outputs().p.astype(np.float64)
    return sim, y


def oracle_grads(grid, c, rho, p0, mask, opts, dldy):
    ops = A.LosslessOperators(grid, rho, mask, opts, float(np.max(c)))
    return A.vjp(ops, np.asarray(c, np.float64), np.asarray(p0, np.float64), dldy)


@pytest.mark.parametrize("cart,rho_map", [(False, False), (True, True)])
def test_vjp_fp64_matches_oracle_2d(cart, rho_map):
    grid, c, rho, p0, mask, opts = case_2d(cart=cart, rho_map=rho_map)
    print(f"\n[case] vjp FP64 2D cart={cart} rho_map={rho_map}", flush=True)
    sim, y = gpu_forward(grid, c, rho, p0, mask, opts, F64)
    rng = np.random.default_rng(7)
    dldy = rng.standard_normal(y.shape)
    gp0, gc = sim.vjp(dldy, checkpoint_every=16)
    sim.close()
    rp0, rc = oracle_grads(grid, c, rho, p0, mask, opts, dldy)
    e0, ec = rel_l2(gp0, rp0), rel_l2(gc, rc)
    print(f"  grad_p0 relL2={e0:.3e} grad_c relL2={ec:.3e}", flush=True)
    assert e0 <= 1e-9 and ec <= 1e-9


def test_vjp_fp32_matches_oracle_2d():
    grid, c, rho, p0, mask, opts = case_2d(n=(64, 64), nt=120)
    print("\n[case] vjp FP32 2D 64^2 (register-resident kernels)", flush=True)
    sim, y = gpu_forward(grid, c, rho, p0, mask, opts, F32)
    y_obs = y * 0.9
    dldy = y - y_obs
    gp0, gc = sim.vjp(dldy, checkpoint_every=32)
    sim.close()
    rp0, rc = oracle_grads(grid, c, rho, p0, mask, opts, dldy)
    e0, ec = rel_l2(gp0, rp0), rel_l2(gc, rc)
    print(f"  grad_p0 relL2={e0:.3e} grad_c relL2={ec:.3e}", flush=True)
    assert e0 <= 1e-4 and ec <= 1e-4


def test_vjp_fp64_matches_oracle_3d():
    n = 24
    grid = KWaveGrid(n, 1e-4, n, 1e-4, n, 1e-4)
    c = random_field((n, n, n), 9, 3.0, 1450.0, 1650.0)
    grid.setTime(40, 0.3 * 1e-4 / float(c.max()))
    p0 = np.zeros((n, n, n), np.float32)
    p0[10:14, 9:15, 11:13] = 1.0
    mask = np.zeros((n, n, n), np.float32)
    mask[18, 4:20:3, 4:20:3] = 1
    opts = SimulationOptions(pmlSize=5, smoothP0=False)
    print("\n[case] vjp FP64 3D 24^3", flush=True)
    sim, y = gpu_forward(grid, c, np.float32(1000.0), p0, mask, opts, F64)
    dldy = np.random.default_rng(3).standard_normal(y.shape)
    gp0, gc = sim.vjp(dldy, checkpoint_every=8)
    sim.close()
    rp0, rc = oracle_grads(grid, c, np.float32(1000.0), p0, mask, opts, dldy)
    e0, ec = rel_l2(gp0, rp0), rel_l2(gc, rc)
    print(f"  grad_p0 relL2={e0:.3e} grad_c relL2={ec:.3e}", flush=True)
    assert e0 <= 1e-9 and ec <= 1e-9


def test_vjp_fp64_register_path_2d():
    """The fused adjoint row kernels (register-resident 64-point rows, FP64) against the oracle,
    with a Cartesian sensor and a density map (staggered dt / rho maps in the fused update)."""
    grid, c, rho, p0, mask, opts = case_2d(n=(64, 64), nt=80, cart=True, rho_map=True, seed=5)
    print("\n[case] vjp FP64 2D 64^2 register-resident", flush=True)
    sim, y = gpu_forward(grid, c, rho, p0, mask, opts, F64)
    dldy = np.random.default_rng(11).standard_normal(y.shape)
    gp0, gc = sim.vjp(dldy, checkpoint_every=16)
    sim.close()
    rp0, rc = oracle_grads(grid, c, rho, p0, mask, opts, dldy)
    e0, ec = rel_l2(gp0, rp0), rel_l2(gc, rc)
    print(f"  grad_p0 relL2={e0:.3e} grad_c relL2={ec:.3e}", flush=True)
    assert e0 <= 1e-9 and ec <= 1e-9


@pytest.mark.parametrize("n,dtype", [(32, F32), (64, F64)])
def test_vjp_register_path_3d(n, dtype):
    """3D adjoint on the register-resident rows (the z-axis PML varies along the row in the fused
    kernels) against the oracle."""
    grid = KWaveGrid(n, 1e-4, n, 1e-4, n, 1e-4)
    c = random_field((n, n, n), 19, 3.0, 1450.0, 1650.0)
    rho = random_field((n, n, n), 20, 3.0, 950.0, 1050.0)
    grid.setTime(30, 0.3 * 1e-4 / float(c.max()))
    p0 = np.zeros((n, n, n), np.float32)
    h = n // 2
    p0[h - 2:h + 2, h - 3:h + 3, h - 1:h + 1] = 1.0
    mask = np.zeros((n, n, n), np.float32)
    mask[h + 5, 4:n - 4:3, 4:n - 4:3] = 1
    opts = SimulationOptions(pmlSize=6, smoothP0=False)
    print(f"\n[case] vjp {dtype.value} 3D {n}^3 register-resident", flush=True)
    sim, y = gpu_forward(grid, c, rho, p0, mask, opts, dtype)
    dldy = np.random.default_rng(4).standard_normal(y.shape)
    gp0, gc = sim.vjp(dldy, checkpoint_every=8)
    sim.close()
    rp0, rc = oracle_grads(grid, c, rho, p0, mask, opts, dldy)
    e0, ec = rel_l2(gp0, rp0), rel_l2(gc, rc)
    print(f"  grad_p0 relL2={e0:.3e} grad_c relL2={ec:.3e}", flush=True)
    tol = 1e-4 if dtype == F32 else 1e-9
    assert e0 <= tol and ec <= tol


def test_checkpointed_forward_reused_and_outputs_kept():
    """set_checkpointing + run + vjp (no forward inside vjp) equals vjp from scratch; the recorded
    outputs and the final field are unchanged by the reverse sweep."""
    grid, c, rho, p0, mask, opts = case_2d(nt=50)
    print("\n[case] vjp checkpoint reuse", flush=True)
    a, y = gpu_forward(grid, c, rho, p0, mask, opts, F64)
    dldy = np.random.default_rng(1).standard_normal(y.shape)
    ga = a.vjp(dldy, checkpoint_every=7)
    a.close()
    b = Simulation(grid, KWaveMedium(c, rho), KWaveSource(p0=p0), KWaveSensor(mask=mask, record=RecordField.p),
                   SimulationOptions(pmlSize=opts.pmlSize, smoothP0=False, dtype=F64))
    b.set_checkpointing(7)
    b.run()
    before = b.outputs()
    gb = b.vjp(dldy, checkpoint_every=7)
    after = b.outputs()
    b.close()
    assert np.array_equal(ga[0], gb[0]) and np.array_equal(ga[1], gb[1])
    assert np.array_equal(before.p, after.p) and np.array_equal(before.pFinal, after.pFinal)


def test_c_gradient_central_differences_fp64():
    """Central differences of the GPU FP64 loss along random directions that leave max(c) alone."""
    grid, c, rho, p0, mask, opts = case_2d(nt=60)
    print("\n[case] vjp vs central differences (FP64)", flush=True)
    c64 = c.astype(np.float64)
    _, y_obs = gpu_forward(grid, (c64 * 1.005).astype(np.float32), rho, p0, mask, opts, F64)
    sim, y = gpu_forward(grid, c, rho, p0, mask, opts, F64)
    _, gc = sim.vjp(y - y_obs, checkpoint_every=10)
    sim.close()
    # c enters the sim as float32: perturb by an exactly representable step on cells away from the max
    rng = np.random.default_rng(21)
    keep = c64 < c64.max() - 20.0
    # <gc, v> for random signs has the scale ||gc||: the error is measured against that scale, so a
    # direction along which the derivative nearly cancels is not held to a relative bound
    scale = float(np.sqrt(np.sum(np.where(keep, gc, 0.0) ** 2)))
    for k in range(3):
        v = np.where(keep, np.sign(rng.standard_normal(c.shape)), 0.0)
        h = 0.125
        lp = 0.5 * np.sum((gpu_forward(grid, (c64 + h * v).astype(np.float32), rho, p0, mask, opts, F64)[1] - y_obs) ** 2)
        lm = 0.5 * np.sum((gpu_forward(grid, (c64 - h * v).astype(np.float32), rho, p0, mask, opts, F64)[1] - y_obs) ** 2)
        fd = (lp - lm) / (2 * h)
        an = float(np.sum(gc * v))
        print(f"  dir {k}: fd={fd:.6e} adjoint={an:.6e} |fd-an|/||gc||={abs(fd - an) / scale:.2e}", flush=True)
        assert abs(fd - an) <= 1e-5 * scale


def test_vjp_rejects_unsupported():
    grid, c, rho, p0, mask, opts = case_2d(nt=10)
    sim = Simulation(grid, KWaveMedium(c, rho, alphaCoeff=0.5, alphaPower=1.5, alphaMode=AbsorptionMode.powerLaw),
                     KWaveSource(p0=p0),
                     KWaveSensor(mask=mask, record=RecordField.p), SimulationOptions(pmlSize=8, smoothP0=False))
    with pytest.raises(NotImplementedError):
        sim.vjp(np.zeros((sim.n_sensor, grid.nt), np.float32))
    sim.close()
