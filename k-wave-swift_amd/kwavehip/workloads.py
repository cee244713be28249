"""Seeded synthetic workloads at their real sizes, shared by the benchmark and the GPU tests.

Each function returns the reference's API values (KWaveGrid, KWaveMedium, KWaveSource, KWaveSensor,
SimulationOptions). Media are Gaussian random fields (`random_field`: SplitMix64 normals, periodic
Gaussian filter of the given correlation length, min-max mapped) or uniform speckle; every input is
rounded to float32 once, so float32 and float64 runs see identical values.
"""
from __future__ import annotations

import numpy as np

from .api import (AbsorptionMode, DTypePrecision, KWaveGrid, KWaveMedium, KWaveSensor, KWaveSource, RecordField,
                  SimulationOptions, SourceMode)
from .inputs import makeBall, makeCartCircle, makeDisc, random_field, toneBurst
from .transducer import KWaveTransducer, KWaveTransducerGeometry


def heterogeneous_nonlinear_2d(n=512, steps=2000, dtype=DTypePrecision.float32):
    """2D heterogeneous, absorbing, nonlinear run: n^2 (512), dx = 5e-5 m, PML 20; c and rho from
    random fields (seeds 1 / 2, correlation 20 cells) in [1400, 1600] m/s and [900, 1100] kg/m^3;
    power law 0.75 dB/(MHz^y cm), y = 1.5, with dispersion; B/A = 6; additive pressure source on
    a 64-point line (row x = 40, y = 224..287 at 512) playing a 1 MHz 3-cycle Hann-windowed burst
    of 1 MPa; CFL 0.3, `steps` steps. Two sensors, as the reference needs two runs for them: the
    full grid recording pMax / pRms, and a 128-point line (row x = 472, y = 192..319) recording p.
    Returns (grid, medium, source, full_sensor, line_sensor, options)."""
    dx = 5e-5
    s = n / 512.0
    c = random_field((n, n), 1, 20.0, 1400.0, 1600.0)
    rho = random_field((n, n), 2, 20.0, 900.0, 1100.0)
    med = KWaveMedium(c, rho, alphaCoeff=0.75, alphaPower=1.5, alphaMode=AbsorptionMode.powerLaw, bOnA=6.0)
    grid = KWaveGrid(n, dx, n, dx)
    grid.setTime(steps, 0.3 * dx / float(c.max()))
    mask = np.zeros((n, n), np.float32)
    mask[int(40 * s), int(224 * s):int(288 * s)] = 1
    sig = (1e6 * toneBurst(1 / grid.dt, 1e6, 3, envelope="hann")).astype(np.float32)
    src = KWaveSource(pMask=mask, p=sig, pMode=SourceMode.additive)
    full = KWaveSensor(mask=np.ones((n, n), np.float32), record=RecordField.pMax | RecordField.pRms)
    line = np.zeros((n, n), np.float32)
    line[int(472 * s), int(192 * s):int(320 * s)] = 1
    opts = SimulationOptions(pmlSize=20, smoothP0=False, dtype=dtype)
    return grid, med, src, full, KWaveSensor(mask=line, record=RecordField.p), opts


def throughput_3d(n=256, steps=1000, dtype=DTypePrecision.float32, plane=168, shape=None):
    """3D heterogeneous power-law run: n^3 (256) or `shape`, dx = 1e-4 m, PML 10, smoothP0; c and
    rho from random fields (seeds 3 / 4, correlation 20 cells) in [1400, 1600] / [900, 1100]; 0.5
    dB/(MHz^y cm), y = 1.1, with dispersion, linear; p0 = ball of radius 8 cells at the centre,
    1 Pa; CFL 0.3; a sensor plane x = `plane` recording p and uz."""
    dx = 1e-4
    shape = tuple(shape) if shape is not None else (n, n, n)
    c = random_field(shape, 3, 20.0, 1400.0, 1600.0)
    rho = random_field(shape, 4, 20.0, 900.0, 1100.0)
    med = KWaveMedium(c, rho, alphaCoeff=0.5, alphaPower=1.1, alphaMode=AbsorptionMode.powerLaw)
    grid = KWaveGrid(shape[0], dx, shape[1], dx, shape[2], dx)
    grid.setTime(steps, 0.3 * dx / float(c.max()))
    src = KWaveSource(p0=makeBall(*shape, radius=8))
    mask = np.zeros(shape, np.float32)
    mask[plane, :, :] = 1
    sens = KWaveSensor(mask=mask, record=RecordField.p | RecordField.uz)
    return grid, med, src, sens, SimulationOptions(pmlSize=10, smoothP0=True, dtype=dtype)


SCAN_LINE_SHAPES = [(96, 96, 96), (128, 96, 96), (160, 128, 128)]


def scan_line_transmit(run, shape, steps=800, dx=1e-4):
    """One transmit of a per-scan-line batch: speckle medium c = 1540 (1 + U[-0.02, 0.02]) m/s and
    rho = 1000 (1 + U[-0.02, 0.02]) kg/m^3 (seed 100 + run); a 32-element linear array (element
    width 2, length 24, kerf 0) at x = 10, centred in y and z, driven by a 2 MHz 2-cycle burst as an
    additive velocity source, steering U[-20, 20] degrees, focus U[15, 40] mm, elevation focus
    20 mm (seed 200 + run); the active voxels record p for the element-combined receive traces.
    Returns ((grid, medium, source, sensor, options), transducer)."""
    rng = np.random.default_rng(100 + run)
    c = (1540.0 * (1 + rng.uniform(-0.02, 0.02, shape))).astype(np.float32)
    rho = (1000.0 * (1 + rng.uniform(-0.02, 0.02, shape))).astype(np.float32)
    grid = KWaveGrid(shape[0], dx, shape[1], dx, shape[2], dx)
    grid.setTime(steps, 0.3 * dx / float(c.max()))
    beam = np.random.default_rng(200 + run)
    steer, focus = beam.uniform(-20.0, 20.0), beam.uniform(15e-3, 40e-3)
    geom = KWaveTransducerGeometry(numberElements=32, elementWidth=2, elementLength=24, elementSpacing=0,
                                   position=(10, (shape[1] - 64) // 2, (shape[2] - 24) // 2), dy=dx, dz=dx)
    burst = toneBurst(1 / grid.dt, 2e6, 2).astype(np.float32)
    tr = KWaveTransducer(geom, shape, grid.dt, 1540.0, focusDistance=focus, elevationFocusDistance=20e-3,
                         steeringAngle=steer, inputSignal=burst)
    mask, row, delays, scales = tr.velocity_source()
    src = KWaveSource(uMask=mask, ux=row, uDelay=delays, uScale=scales, uMode=SourceMode.additive)
    sens = KWaveSensor(mask=mask, record=RecordField.p)
    return (grid, KWaveMedium(c, rho), src, sens, SimulationOptions(pmlSize=10, smoothP0=False)), tr


def scan_line_batch(n_runs=48, steps=800, seed=7):
    """`n_runs` transmits with grid sizes drawn (seed 7) from 96^3, 128x96x96 and 160x128x128."""
    pick = np.random.default_rng(seed).integers(0, 3, n_runs)
    return [scan_line_transmit(r, SCAN_LINE_SHAPES[pick[r]], steps) for r in range(n_runs)]


def inversion_2d(n=512, steps=1500):
    """2D sound-speed / photoacoustic inversion: n^2 (512), dx = 1e-4 m, PML 20; c from a random
    field (seed 11) in [1450, 1650] m/s, rho = 1000; p0 = 20 random discs (seed 12, radius 3..12
    cells, amplitude U[0.2, 1]); smoothP0 off; 256-point Cartesian circle of radius 22 mm recording
    p; CFL 0.3. The observation comes from c perturbed by 1% of a random field (seed 13).
    Returns (grid, c, c_obs, p0, sensor, options)."""
    dx = 1e-4
    grid = KWaveGrid(n, dx, n, dx)
    c = random_field((n, n), 11, 20.0, 1450.0, 1650.0)
    grid.setTime(steps, 0.3 * dx / float(c.max()))
    rng = np.random.default_rng(12)
    p0 = np.zeros((n, n), np.float32)
    lo, hi = int(120 * n / 512), n - int(120 * n / 512)
    for _ in range(20):
        cx, cy = rng.integers(lo, hi, 2)
        p0 += float(rng.uniform(0.2, 1.0)) * makeDisc(n, n, cx=int(cx), cy=int(cy), radius=int(rng.integers(3, 13)))
    pert = random_field((n, n), 13, 20.0, -1.0, 1.0)
    c_obs = (c.astype(np.float64) * (1.0 + 0.01 * pert)).astype(np.float32)
    sens = KWaveSensor(mask=makeCartCircle(22e-3 * n / 512, 256), record=RecordField.p)
    return grid, c, c_obs, p0, sens, SimulationOptions(pmlSize=20, smoothP0=False)
