"""GPU parity of the register-resident kernels for the axis lengths L = J*E beyond the square
ones: 32 (4 x 8), 96 (4 x 24), 128 (8 x 16), 160 (4 x 40), 192 (8 x 24), 512 (16 x 32),
1024 (32 x 32). Each
length is run as the contiguous axis (row kernels), the outermost axis (middle pass) and, in 3D,
the strided axis, against the float64 oracle on identical float32 inputs; plus agreement with
the generic Stockham kernels on the same problem and bitwise neutrality of the launch schedule.

Tolerances are the reference's parity tolerances (ParityTests.swift:54-58): float32 relL2 <= 1e-4
and relMax <= 1e-3.
"""
import os

import numpy as np
import pytest

from kwavehip import (AbsorptionMode, KWaveGrid, KWaveMedium, KWaveSensor, KWaveSource, RecordField, Simulation,
                      SimulationOptions, SourceMode, kspaceFirstOrder, makeBall, makeDisc, random_field, toneBurst)

from test_gpu_parity import F32, _device, check, metrics  # noqa: F401  (_device: module fixture)

pytestmark = pytest.mark.gpu


def _medium(shape, seed, absorbing=True, nonlinear=True, mode=AbsorptionMode.powerLaw, y=1.5):
    c = random_field(shape, seed, 4.0, 1400.0, 1600.0)
    rho = random_field(shape, seed + 1, 4.0, 900.0, 1100.0)
    return KWaveMedium(c, rho, alphaCoeff=0.75 if absorbing else None, alphaPower=y,
                       alphaMode=mode if absorbing else AbsorptionMode.noAbsorption,
                       bOnA=6.0 if nonlinear else None)


@pytest.mark.parametrize("shape", [(96, 160), (160, 96), (128, 192), (192, 128), (512, 32), (32, 512), (1024, 48),
                                   (48, 1024)])
def test_2d_register_lengths(shape):
    """Heterogeneous absorbing nonlinear medium, p0 disc plus an additive (k-space corrected)
    velocity line source; line sensor p / ux / uy and pFinal."""
    nx, ny = shape
    dx = 1e-4
    med = _medium(shape, seed=61)
    grid = KWaveGrid(nx, dx, ny, dx)
    grid.setTime(90, 0.3 * dx / float(np.max(med.soundSpeed)))
    umask = np.zeros(shape, np.float32)
    umask[nx // 2 - 3, ny // 2 - 4: ny // 2 + 4] = 1
    sig = (0.2 * toneBurst(1 / grid.dt, 3e6, 2)).astype(np.float32)
    sens = np.zeros(shape, np.float32)
    sens[nx // 2 + 6, :] = 1
    check(f"2d_reg_{nx}x{ny}", grid, med,
          KWaveSource(p0=makeDisc(nx, ny, radius=5), uMask=umask, ux=sig, uMode=SourceMode.additive),
          KWaveSensor(mask=sens, record=RecordField.p | RecordField.ux | RecordField.uy | RecordField.pFinal),
          ["p", "ux", "uy", "pFinal"], pmlSize=10)


@pytest.mark.parametrize("absorbing,mode,nonlinear", [
    (False, AbsorptionMode.powerLaw, False), (False, AbsorptionMode.powerLaw, True),
    (True, AbsorptionMode.stokes, False), (True, AbsorptionMode.stokes, True),
    (True, AbsorptionMode.noDispersion, False), (True, AbsorptionMode.powerLaw, True)])
def test_2d_register_density_variants(absorbing, mode, nonlinear):
    """Every density-kernel variant (EOS, Stokes EOS, FFT absorption with / without dispersion;
    linear and B/A) with an additive k-space corrected pressure source, on a 96 x 160 grid."""
    shape = (96, 160)
    med = _medium(shape, seed=71, absorbing=absorbing, nonlinear=nonlinear, mode=mode, y=1.4)
    grid = KWaveGrid(shape[0], 1e-4, shape[1], 1e-4)
    grid.setTime(120, 0.3 * 1e-4 / float(np.max(med.soundSpeed)))
    pmask = np.zeros(shape, np.float32)
    pmask[40, 70:90] = 1
    sig = (1e5 * toneBurst(1 / grid.dt, 2e6, 2)).astype(np.float32)
    check(f"2d_reg_density_{absorbing}_{mode.value}_{nonlinear}", grid, med,
          KWaveSource(p0=makeDisc(*shape, radius=6), pMask=pmask, p=sig, pMode=SourceMode.additive),
          KWaveSensor(), ["pFinal"], pmlSize=10)


@pytest.mark.parametrize("mode,nonlinear", [(AbsorptionMode.powerLaw, False), (AbsorptionMode.noDispersion, False),
                                            (AbsorptionMode.powerLaw, True), (AbsorptionMode.noDispersion, True)])
def test_2d_512_rows_fft_absorption(mode, nonlinear):
    """Length-512 rows with FFT power-law absorption, linear and B/A, with and without dispersion,
    against the oracle (pFinal and the p series of a line sensor): the half-form density kernel
    forms rhoT and R2C_z(rhoT) itself (the row-pair form leaves them to reg_z_rhot)."""
    shape = (48, 512)
    med = _medium(shape, seed=75, absorbing=True, nonlinear=nonlinear, mode=mode, y=1.3)
    grid = KWaveGrid(shape[0], 1e-4, shape[1], 1e-4)
    grid.setTime(100, 0.3 * 1e-4 / float(np.max(med.soundSpeed)))
    sens = np.zeros(shape, np.float32)
    sens[30, :] = 1
    check(f"2d_512_rows_{mode.value}_{nonlinear}", grid, med, KWaveSource(p0=makeDisc(*shape, radius=6)),
          KWaveSensor(mask=sens, record=RecordField.p | RecordField.pFinal), ["p", "pFinal"], pmlSize=10)


@pytest.mark.parametrize("absorbing,mode,nonlinear", [
    (False, AbsorptionMode.powerLaw, False), (False, AbsorptionMode.powerLaw, True),
    (True, AbsorptionMode.stokes, False), (True, AbsorptionMode.stokes, True)])
def test_2d_512_rows_density_variants(absorbing, mode, nonlinear):
    """Length-512 rows in the half form (one real row per 256-point transform) through the EOS and
    Stokes density kernels, linear and B/A, with an additive (k-space corrected) pressure source."""
    shape = (48, 512)
    med = _medium(shape, seed=73, absorbing=absorbing, nonlinear=nonlinear, mode=mode, y=1.4)
    grid = KWaveGrid(shape[0], 1e-4, shape[1], 1e-4)
    grid.setTime(100, 0.3 * 1e-4 / float(np.max(med.soundSpeed)))
    pmask = np.zeros(shape, np.float32)
    pmask[20, 200:260] = 1
    sig = (1e5 * toneBurst(1 / grid.dt, 2e6, 2)).astype(np.float32)
    sens = np.zeros(shape, np.float32)
    sens[30, :] = 1
    check(f"2d_512_density_{absorbing}_{mode.value}_{nonlinear}", grid, med,
          KWaveSource(p0=makeDisc(*shape, radius=6), pMask=pmask, p=sig, pMode=SourceMode.additive),
          KWaveSensor(mask=sens, record=RecordField.p | RecordField.pFinal), ["p", "pFinal"], pmlSize=10)


@pytest.mark.parametrize("mode", [SourceMode.dirichlet, SourceMode.additiveNoCorrection])
def test_2d_512_rows_point_sources(mode):
    """Length-512 rows in the half form with per-point pressure and velocity sources (applied in the
    row kernels' LDS, sample n of a row in slot n / 2): points at even and odd positions along the
    row, against the oracle."""
    shape = (40, 512)
    grid = KWaveGrid(shape[0], 1e-4, shape[1], 1e-4)
    grid.setTime(80, 2e-8)
    pmask = np.zeros(shape, np.float32)
    pmask[15, 250:254] = 1
    umask = np.zeros(shape, np.float32)
    umask[25, 101:106] = 1
    sig = toneBurst(1 / grid.dt, 1e6, 3).astype(np.float32)
    sens = np.zeros(shape, np.float32)
    sens[20, :] = 1
    src = KWaveSource(pMask=pmask, p=1e4 * sig, pMode=mode, uMask=umask, ux=0.01 * sig, uy=0.005 * sig, uMode=mode)
    check(f"2d_512_points_{mode.value}", grid, KWaveMedium(random_field(shape, 83, 4.0, 1450.0, 1550.0), 1000.0),
          src, KWaveSensor(mask=sens, record=RecordField.p | RecordField.ux | RecordField.pFinal),
          ["p", "ux", "pFinal"], pmlSize=8, smoothP0=False)


def test_3d_512_columns():
    """Length-512 y columns (the strided passes, 16 x 32, with the two-output gradient inverse) on
    an 8 x 512 x 32 heterogeneous power-law B/A medium with a p0 ball and a line sensor along y
    (p, uy, pFinal), against the oracle."""
    shape = (8, 512, 32)
    c = random_field(shape, 93, 4.0, 1400.0, 1600.0)
    rho = random_field(shape, 94, 4.0, 900.0, 1100.0)
    med = KWaveMedium(c, rho, alphaCoeff=0.6, alphaPower=1.2, alphaMode=AbsorptionMode.powerLaw, bOnA=5.0)
    grid = KWaveGrid(shape[0], 1e-4, shape[1], 1e-4, shape[2], 1e-4)
    grid.setTime(40, 0.3 * 1e-4 / float(c.max()))
    sens = np.zeros(shape, np.float32)
    sens[4, :, 16] = 1
    check("3d_512_columns", grid, med, KWaveSource(p0=makeBall(*shape, radius=3)),
          KWaveSensor(mask=sens, record=RecordField.p | RecordField.uy | RecordField.pFinal), ["p", "uy", "pFinal"],
          pmlSize=3)


def test_1d_512_half_form():
    """1-D grid of 512 points (one row, in a workgroup of 16 rows with 15 masked) with power-law
    absorption and a dirichlet point source at an odd position, against the oracle."""
    n = 512
    grid = KWaveGrid(n, 1e-4)
    grid.setTime(300, 2e-8)
    mask = np.zeros(n, np.float32)
    mask[201] = 1
    sens = np.zeros(n, np.float32)
    sens[[170, 240]] = 1   # within reach of the source's wave in 300 steps (0.3 cells per step)
    sig = toneBurst(1 / grid.dt, 1e6, 3).astype(np.float32)
    med = KWaveMedium(1500.0, 1000.0, alphaCoeff=0.75, alphaPower=1.5, alphaMode=AbsorptionMode.powerLaw)
    check("1d_512_half", grid, med, KWaveSource(pMask=mask, p=sig, pMode=SourceMode.dirichlet),
          KWaveSensor(mask=sens), ["p", "pFinal"], pmlSize=20, smoothP0=False)


@pytest.mark.parametrize("shape", [(48, 512), (12, 16, 512)])
def test_512_rows_half_form_matches_row_pairs(shape):
    """KWH_ROW_HALF=0 runs the velocity, density and EOS kernels of length-512 rows as row pairs
    (512-point transforms, 16 x 32) with the summed-density pass: the two forms differ only in FP32
    rounding (power-law medium with B/A, p0, line sensor)."""
    dim = len(shape)
    args = [v for n in shape for v in (n, 1e-4)]
    grid = KWaveGrid(*args)
    grid.setTime(40, 2e-8)
    p0 = makeDisc(*shape, radius=5) if dim == 2 else makeBall(*shape, radius=3)
    c = random_field(shape, 87, 4.0, 1400.0, 1600.0)
    med = KWaveMedium(c, 1000.0, alphaCoeff=0.6, alphaPower=1.2, alphaMode=AbsorptionMode.powerLaw, bOnA=5.0)
    sens = np.zeros(shape, np.float32)
    sens[shape[0] // 2 + 2] = 1
    run = lambda: kspaceFirstOrder(grid, med, KWaveSource(p0=p0),  # noqa: E731
                                   KWaveSensor(mask=sens, record=RecordField.p | RecordField.pFinal),
                                   SimulationOptions(pmlSize=4, dtype=F32))
    print(f"\n[case] half form vs row pairs {shape}", flush=True)
    a = run()
    os.environ["KWH_ROW_HALF"] = "0"
    try:
        b = run()
    finally:
        del os.environ["KWH_ROW_HALF"]
    for f in ("p", "pFinal"):
        rl2, rmax = metrics(getattr(a, f), getattr(b, f))
        print(f"  {f} relL2={rl2:.3e} relMax={rmax:.3e}", flush=True)
        assert rl2 < 1e-5 and rmax < 1e-4


def test_3d_1024_rows_fft_absorption():
    """3D grid with length-1024 rows (32 lanes x 32 elements per transform) and FFT power-law
    absorption with B/A, against the oracle."""
    shape = (8, 12, 1024)
    c = random_field(shape, 79, 4.0, 1400.0, 1600.0)
    rho = random_field(shape, 80, 4.0, 900.0, 1100.0)
    med = KWaveMedium(c, rho, alphaCoeff=0.6, alphaPower=1.2, alphaMode=AbsorptionMode.powerLaw, bOnA=5.0)
    grid = KWaveGrid(shape[0], 1e-4, shape[1], 1e-4, shape[2], 1e-4)
    grid.setTime(40, 0.3 * 1e-4 / float(c.max()))
    sens = np.zeros(shape, np.float32)
    sens[4, 6, :] = 1
    check("3d_1024_rows_powerlaw", grid, med, KWaveSource(p0=makeBall(*shape, radius=3)),
          KWaveSensor(mask=sens, record=RecordField.p | RecordField.pFinal), ["p", "pFinal"], pmlSize=3)


def test_3d_512_rows_fft_absorption():
    """3D grid with length-512 rows and FFT power-law absorption (the summed-density pass over three
    axes), heterogeneous medium with B/A, against the oracle."""
    shape = (12, 16, 512)
    c = random_field(shape, 77, 4.0, 1400.0, 1600.0)
    rho = random_field(shape, 78, 4.0, 900.0, 1100.0)
    med = KWaveMedium(c, rho, alphaCoeff=0.6, alphaPower=1.2, alphaMode=AbsorptionMode.powerLaw, bOnA=5.0)
    grid = KWaveGrid(shape[0], 1e-4, shape[1], 1e-4, shape[2], 1e-4)
    grid.setTime(40, 0.3 * 1e-4 / float(c.max()))
    sens = np.zeros(shape, np.float32)
    sens[6, 8, :] = 1
    check("3d_512_rows_powerlaw", grid, med, KWaveSource(p0=makeBall(*shape, radius=3)),
          KWaveSensor(mask=sens, record=RecordField.p | RecordField.pFinal), ["p", "pFinal"], pmlSize=4)


@pytest.mark.parametrize("shape", [(32, 96, 160), (160, 32, 96), (96, 128, 32), (128, 192, 32)])
def test_3d_register_lengths(shape):
    """3D heterogeneous power-law medium with B/A, p0 ball plus a velocity source patch; plane
    sensor recording p, ux and uz (the contiguous component is collocated on the sensor rows)."""
    nx, ny, nz = shape
    c = random_field(shape, 81, 4.0, 1400.0, 1600.0)
    rho = random_field(shape, 82, 4.0, 900.0, 1100.0)
    med = KWaveMedium(c, rho, alphaCoeff=0.5, alphaPower=1.1, alphaMode=AbsorptionMode.powerLaw, bOnA=5.0)
    grid = KWaveGrid(nx, 1e-4, ny, 1e-4, nz, 1e-4)
    grid.setTime(40, 0.3 * 1e-4 / float(c.max()))
    umask = np.zeros(shape, np.float32)
    umask[nx // 2 - 4, ny // 2 - 3: ny // 2 + 3, nz // 2 - 3: nz // 2 + 3] = 1
    sig = (0.2 * toneBurst(1 / grid.dt, 3e6, 2)).astype(np.float32)
    sens = np.zeros(shape, np.float32)
    sens[nx // 2 + 3] = 1
    check(f"3d_reg_{nx}x{ny}x{nz}", grid, med,
          KWaveSource(p0=makeBall(nx, ny, nz, radius=4), uMask=umask, ux=sig),
          KWaveSensor(mask=sens, record=RecordField.p | RecordField.ux | RecordField.uz),
          ["p", "ux", "uz"], pmlSize=6, workers=8)


@pytest.mark.parametrize("shape", [(96, 128, 160), (160, 128, 128), (512, 512)])
def test_register_and_generic_paths_agree(shape):
    """The register-resident kernels and the generic Stockham kernels on the same float32 problem
    (the scan-line grid shapes and the 512^2 grid): the two differ only in FP32 rounding."""
    dim = len(shape)
    args = [v for n in shape for v in (n, 1e-4)]
    grid = KWaveGrid(*args)
    grid.setTime(40, 2e-8)
    p0 = makeDisc(*shape, radius=6) if dim == 2 else makeBall(*shape, radius=6)
    med = KWaveMedium(random_field(shape, 91, 6.0, 1450.0, 1600.0), 1000.0)
    sens = np.zeros(shape, np.float32)
    sens[shape[0] // 2 + 8] = 1
    run = lambda: kspaceFirstOrder(grid, med, KWaveSource(p0=p0), KWaveSensor(mask=sens, record=RecordField.p),
                                   SimulationOptions(pmlSize=10))
    print(f"\n[case] register vs generic {shape}", flush=True)
    a = run()
    os.environ["KWH_GENERIC_FFT"] = "1"
    try:
        b = run()
    finally:
        del os.environ["KWH_GENERIC_FFT"]
    for f in ("p", "pFinal"):
        rl2, rmax = metrics(getattr(a, f), getattr(b, f))
        print(f"  {f} relL2={rl2:.3e} relMax={rmax:.3e}", flush=True)
        assert rl2 < 1e-5 and rmax < 1e-4


@pytest.mark.parametrize("shape", [(96, 160), (32, 96, 160), (160, 128, 128)])
def test_register_lengths_schedule_is_bitwise_neutral(shape):
    """Autotuned launch schedules change launch geometry only: results equal the default
    schedule's bit for bit."""
    dim = len(shape)
    args = [v for n in shape for v in (n, 1e-4)]
    grid = KWaveGrid(*args)
    grid.setTime(25, 2e-8)
    p0 = makeDisc(*shape, radius=4) if dim == 2 else makeBall(*shape, radius=4)
    mask = np.zeros(shape, np.float32)
    mask[(shape[0] * 3) // 4] = 1
    med = KWaveMedium(random_field(shape, 19, 4.0, 1450.0, 1600.0), 1000.0, alphaCoeff=0.5, alphaPower=1.3,
                      alphaMode=AbsorptionMode.powerLaw)
    make = lambda: Simulation(grid, med, KWaveSource(p0=p0), KWaveSensor(mask=mask, record=RecordField.p),
                              SimulationOptions(pmlSize=6))
    print(f"\n[case] autotune {shape}", flush=True)
    a = make()
    a.run()
    ref = a.outputs()
    a.close()
    b = make()
    report = b.autotune()
    print("  " + report.replace("\n", "\n  "), flush=True)
    b.run()
    out = b.outputs()
    b.close()
    assert np.array_equal(out.p, ref.p) and np.array_equal(out.pFinal, ref.pFinal)
