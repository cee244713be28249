"""GPU parity of the axisymmetric solver (kwh_sim_create_axisymmetric, kspaceFirstOrderAS) against
the float64 axisymmetric oracle on identical float32 inputs.

Cases follow the reference's ParityAxisymmetricTests.swift and its generator
Scripts/parity/generate_reference_as.py (64 x 32 grid, dx = dy = 1e-4 m, c0 = 1500 m/s,
rho0 = 1000 kg/m^3, PML 10 with alpha 2, 120 steps at CFL 0.3, smoothed Gaussian p0 on the axis,
sensor column y = 8, a 3-point source on the axis driven by a windowed 2 MHz burst): p0 IVP,
additive / dirichlet pressure sources, Stokes absorption, B/A, additive ux / uy sources. Metrics
are the reference test's (ParityAxisymmetricTests.swift:60-69): rms and max of the difference over
max |ref|, both below 1e-4 in float32 (1e-9 in float64).
"""
import math

import numpy as np
import pytest

from oracle import kwave_oracle as O
from kwavehip import (AbsorptionMode, DTypePrecision, KWaveGrid, KWaveMedium, KWaveSensor, KWaveSource,
                      RecordField, SimulationOptions, SourceMode, kspaceFirstOrderAS, makeCartCircle, random_field)

from test_gpu_parity import _device  # noqa: F401  (module fixture: the library must load, a device must exist)

pytestmark = pytest.mark.gpu

F32 = DTypePrecision.float32
F64 = DTypePrecision.float64
NX, NY, STEPS, DX, C0, RHO0 = 64, 32, 120, 1e-4, 1500.0, 1000.0


def _grid(nx=NX, ny=NY, steps=STEPS):
    g = KWaveGrid(nx, DX, ny, DX)
    g.setTime(steps, 0.3 * DX / C0)
    return g


def _p0(nx=NX, ny=NY, amp=1.0):
    x = (np.arange(nx) - nx // 2)[:, None]
    y = np.arange(ny)[None, :]
    return (amp * np.exp(-(x ** 2 + y ** 2) / 16.0)).astype(np.float32)


def _burst(steps=STEPS):
    t = np.arange(steps)
    dt = 0.3 * DX / C0
    return (np.sin(2 * np.pi * 2e6 * t * dt) * np.exp(-((t - 30.0) ** 2) / 200.0)).astype(np.float32)


def _src_mask(nx=NX, ny=NY):
    m = np.zeros((nx, ny), np.float32)
    m[nx // 4, 0:3] = 1
    return m


def _sensor(nx=NX, ny=NY):
    m = np.zeros((nx, ny), np.float32)
    m[:, 8] = 1
    return KWaveSensor(mask=m, record=RecordField.p | RecordField.pFinal)


def _check(name, grid, medium, source, sensor, dtype=F32, **opt):
    opt.setdefault("pmlSize", 10)
    opt.setdefault("pmlAlpha", 2.0)
    print(f"\n[case] as_{name} dtype={dtype.value} grid={grid.size} nt={grid.nt}", flush=True)
    gpu = kspaceFirstOrderAS(grid, medium, source, sensor, SimulationOptions(dtype=dtype, **opt))
    ref = O.kspace_first_order_as(grid, medium, source, sensor, SimulationOptions(**opt))
    tol = 1e-4 if dtype == F32 else 1e-9
    for f in ("p", "pFinal"):
        g = np.asarray(getattr(gpu, f), np.float64)
        r = ref[f]
        assert g.shape == r.shape, (f, g.shape, r.shape)
        scale = np.abs(r).max()
        d = g - r
        rl2 = math.sqrt(np.mean(d * d)) / scale
        rmax = np.abs(d).max() / scale
        print(f"  {f:7s} rms/max={rl2:.3e} max/max={rmax:.3e} max|ref|={scale:.3e}", flush=True)
        assert rl2 < tol and rmax < tol, (name, f, rl2, rmax)


@pytest.mark.parametrize("dtype", [F32, F64])
def test_as_ivp(dtype):
    _check("ivp", _grid(), KWaveMedium(C0, RHO0), KWaveSource(p0=_p0()), _sensor(), dtype, smoothP0=True)


@pytest.mark.parametrize("mode", [SourceMode.additive, SourceMode.dirichlet, SourceMode.additiveNoCorrection])
def test_as_pressure_source(mode):
    _check(f"p_{mode.value}", _grid(), KWaveMedium(C0, RHO0), KWaveSource(pMask=_src_mask(), p=_burst(), pMode=mode),
           _sensor())


def test_as_stokes_absorption():
    med = KWaveMedium(C0, RHO0, alphaCoeff=5.0, alphaPower=2.0, alphaMode=AbsorptionMode.stokes)
    _check("stokes", _grid(), med, KWaveSource(p0=_p0()), _sensor(), smoothP0=True)


def test_as_nonlinear():
    _check("nl", _grid(), KWaveMedium(C0, RHO0, bOnA=10.0), KWaveSource(p0=_p0(amp=5e6)), _sensor(), smoothP0=True)


@pytest.mark.parametrize("axis,mode", [("x", SourceMode.additive), ("y", SourceMode.additive),
                                       ("x", SourceMode.dirichlet), ("y", SourceMode.additiveNoCorrection)])
def test_as_velocity_source(axis, mode):
    kw = {"ux": _burst()} if axis == "x" else {"uy": _burst()}
    _check(f"u{axis}_{mode.value}", _grid(), KWaveMedium(C0, RHO0),
           KWaveSource(uMask=_src_mask(), uMode=mode, **kw), _sensor())


@pytest.mark.parametrize("dtype", [F32, F64])
def test_as_heterogeneous_nonlinear_stokes_cartesian(dtype):
    """Heterogeneous c0 / rho0 maps, Stokes absorption and B/A together with an additive uy source
    and a Cartesian sensor, on 48 x 20 (expanded radial length 80: the generic transform kernels)."""
    nx, ny = 48, 20
    shape = (nx, ny)
    c = random_field(shape, 5, 4.0, 1450.0, 1600.0)
    rho = random_field(shape, 6, 4.0, 950.0, 1050.0)
    med = KWaveMedium(c, rho, alphaCoeff=2.0, alphaPower=2.0, alphaMode=AbsorptionMode.stokes, bOnA=5.0)
    src = KWaveSource(p0=_p0(nx, ny, amp=2e6), uMask=_src_mask(nx, ny), uy=(0.5 * _burst(90)).astype(np.float32))
    pts = makeCartCircle(0.5e-3, 24)
    pts = np.stack([pts[0], np.abs(pts[1]) + 0.2e-3]).astype(np.float32)   # inside the grid's y extent
    sens = KWaveSensor(mask=pts, record=RecordField.p | RecordField.pFinal)
    _check("hetero_cart", _grid(nx, ny, 90), med, src, sens, dtype, pmlSize=6)
