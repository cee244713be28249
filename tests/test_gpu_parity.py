"""GPU parity: the HIP library (through the C ABI) against the float64 oracle on identical
float32 inputs. Cases follow the reference's own tests (Tests/KWaveTests/*.swift) and parity
generator configurations (Scripts/parity/generate_reference_*.py).

Metrics are the reference's (ParityTests.swift:54-58): relL2 = rms(diff)/rms(ref),
relMax = max|diff|/max|ref|. Tolerances: float32 relL2 <= 1e-4 and relMax <= 1e-3;
float64 relL2 <= 1e-9.
"""
import ctypes as C
import math
import sys

import numpy as np
import pytest

from oracle import kwave_oracle as O
from kwavehip import (AbsorptionMode, DirectivityPattern, DTypePrecision, KWaveGrid, KWaveMedium,
                      KWaveSensor, KWaveSource, RecordField, Simulation, SimulationOptions, SourceMode,
                      kspaceFirstOrder, makeBall, makeCartCircle, makeCircle, makeDisc, random_field,
                      timeReversal, toneBurst)
from kwavehip import _lib as L

pytestmark = pytest.mark.gpu

F32 = DTypePrecision.float32
F64 = DTypePrecision.float64


@pytest.fixture(scope="module", autouse=True)
def _device():
    lib = L.load()   # the library must be built: a missing .so fails, it is never skipped
    h = C.c_void_p()
    if lib.kwh_ctx_create(0, C.byref(h)) != 0:
        pytest.skip("no HIP device")
    lib.kwh_ctx_destroy(h)


def metrics(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    d = a - b
    rl2 = math.sqrt(np.mean(d * d)) / max(math.sqrt(np.mean(b * b)), 1e-300)
    rmax = np.abs(d).max() / max(np.abs(b).max(), 1e-300)
    return rl2, rmax


def check(name, grid, medium, source, sensor, fields, dtype=F32, l2=None, mx=None, final_scale=None, workers=1,
          **opt):
    print(f"\n[case] {name} dtype={dtype.value} grid={grid.size} nt={grid.nt}", flush=True)
    gpu = kspaceFirstOrder(grid, medium, source, sensor, SimulationOptions(dtype=dtype, **opt))
    ref = O.kspace_first_order(grid, medium, source, sensor, SimulationOptions(**opt), workers=workers)
    if l2 is None:
        l2 = 1e-4 if dtype == F32 else 1e-9
    if mx is None:
        mx = 1e-3 if dtype == F32 else 1e-8
    for f in fields:
        g = getattr(gpu, f)
        r = ref[f]
        assert g is not None, f"{name}: {f} missing"
        assert g.shape == r.shape, (f, g.shape, r.shape)
        assert np.all(np.isfinite(g)), f
        rl2, rmax = metrics(g, r)
        if final_scale is not None and f.endswith("Final"):
            # residual field after the pulse has passed: normalise by max over t of |p|
            d = np.asarray(g, np.float64) - r
            rl2 = math.sqrt(np.mean(d * d)) / final_scale
            rmax = np.abs(d).max() / final_scale
        print(f"  {f:8s} relL2={rl2:.3e} relMax={rmax:.3e} max|ref|={np.abs(r).max():.3e}", flush=True)
        assert rl2 <= l2 and rmax <= mx, (name, f, rl2, rmax)
    return gpu, ref


# ------------------------------------------------------------------------------------------ 1D
@pytest.mark.parametrize("dtype", [F32, F64])
def test_1d_symmetric_ivp(dtype):
    """SolverTests.swift:18-47."""
    n = 64
    grid = KWaveGrid(n, 1e-4)
    grid.makeTime(1500, cfl=0.3, tEnd=4e-7)
    p0 = np.zeros(n, np.float32)
    p0[n // 2 - 3:n // 2 + 4] = 1
    check("1d_ivp", grid, KWaveMedium(1500.0, 1000.0), KWaveSource(p0=p0), KWaveSensor(), ["pFinal"], dtype)


def test_1d_time_varying_source():
    """SolverTests.swift:51-74."""
    n = 96
    grid = KWaveGrid(n, 1e-4)
    grid.makeTime(1500, cfl=0.3, tEnd=5e-6)
    mask = np.zeros(n, np.float32)
    mask[16] = 1
    sens = np.zeros(n, np.float32)
    sens[64] = 1
    sig = toneBurst(1 / grid.dt, 1e6, 3).astype(np.float32)
    gpu, _ = check("1d_source", grid, KWaveMedium(1500.0, 1000.0), KWaveSource(pMask=mask, p=sig),
                   KWaveSensor(mask=sens), ["p", "pFinal"])
    assert gpu.p.shape == (1, grid.nt)


# ------------------------------------------------------------------------------------------ 2D
def test_2d_ivp_parity_config():
    """ParityTests.swift:21-66 (64^2, disc r=5, PML 20, smoothed, nt = 302)."""
    grid = KWaveGrid(64, 1e-4, 64, 1e-4)
    grid.makeTime(1500, cfl=0.3)
    assert grid.nt == 302
    check("2d_ivp_302", grid, KWaveMedium(1500.0, 1000.0), KWaveSource(p0=makeDisc(64, 64, radius=5)),
          KWaveSensor(), ["pFinal"], pmlSize=20)


def test_2d_line_sensor_gpu_smoke_config():
    """GPUSmokeTests.swift:58-75."""
    n = 48
    grid = KWaveGrid(n, 1e-4, n, 1e-4)
    grid.makeTime(1500, cfl=0.3, tEnd=1.2e-6)
    mask = np.zeros((n, n), np.float32)
    mask[8, :] = 1
    check("2d_line", grid, KWaveMedium(1500.0, 1000.0), KWaveSource(p0=makeDisc(n, n, radius=4)),
          KWaveSensor(mask=mask), ["pFinal", "p"])


def test_2d_c0_config_128():
    """128^2 IVP: disc r=5, PML 20, smoothP0, tEnd 4e-6 -> nt 201, circle sensor r=40, p/pFinal/pMax."""
    n = 128
    grid = KWaveGrid(n, 1e-4, n, 1e-4)
    grid.makeTime(1500, cfl=0.3, tEnd=4e-6)
    assert grid.nt == 201
    sens = KWaveSensor(mask=makeCircle(n, n, radius=40), record=RecordField.p | RecordField.pFinal | RecordField.pMax)
    check("2d_c0", grid, KWaveMedium(1500.0, 1000.0), KWaveSource(p0=makeDisc(n, n, radius=5)), sens,
          ["p", "pFinal", "pMax"])


def _hetero_2d(n, seed=1, absorbing=True, nonlinear=True, mode=AbsorptionMode.powerLaw, y=1.5):
    c = random_field((n, n), seed, 4.0, 1400.0, 1600.0)
    rho = random_field((n, n), seed + 1, 4.0, 900.0, 1100.0)
    return KWaveMedium(c, rho, alphaCoeff=0.75 if absorbing else None, alphaPower=y,
                       alphaMode=mode if absorbing else AbsorptionMode.noAbsorption,
                       bOnA=6.0 if nonlinear else None)


@pytest.mark.parametrize("dtype", [F32, F64])
def test_2d_hetero_absorbing_nonlinear_line_source(dtype):
    """Heterogeneous c/rho, power-law y=1.5 with dispersion, B/A=6, additive p line source with a
    Hann tone burst, full-grid pMax/pRms + line p (the 512^2 workload at 64^2)."""
    n = 64
    dx = 5e-5
    med = _hetero_2d(n)
    grid = KWaveGrid(n, dx, n, dx)
    grid.setTime(300, 0.3 * dx / float(np.max(med.soundSpeed)))
    mask = np.zeros((n, n), np.float32)
    mask[8, 24:40] = 1
    sig = (1e6 * toneBurst(1 / grid.dt, 1e6 * 4, 3, envelope="hann")).astype(np.float32)
    src = KWaveSource(pMask=mask, p=sig, pMode=SourceMode.additive)
    full = KWaveSensor(mask=np.ones((n, n), np.float32), record=RecordField.pMax | RecordField.pRms)
    _, ref = check("2d_hetero_full", grid, med, src, full, ["pMax", "pRms"], dtype, pmlSize=10, smoothP0=False)
    check("2d_hetero_final", grid, med, src, full, ["pFinal"], dtype, final_scale=float(np.abs(ref["pMax"]).max()),
          pmlSize=10, smoothP0=False)
    line = np.zeros((n, n), np.float32)
    line[50, 16:48] = 1
    check("2d_hetero_line", grid, med, src, KWaveSensor(mask=line), ["p"], dtype, pmlSize=10, smoothP0=False)


@pytest.mark.parametrize("mode", [AbsorptionMode.stokes, AbsorptionMode.noDispersion])
def test_2d_absorption_modes(mode):
    n = 48
    med = _hetero_2d(n, seed=5, nonlinear=False, mode=mode, y=1.3)
    grid = KWaveGrid(n, 1e-4, n, 1e-4)
    grid.setTime(150, 0.3 * 1e-4 / float(np.max(med.soundSpeed)))
    check(f"2d_{mode.value}", grid, med, KWaveSource(p0=makeDisc(n, n, radius=5)), KWaveSensor(), ["pFinal"])


@pytest.mark.parametrize("mode", [SourceMode.dirichlet, SourceMode.additive, SourceMode.additiveNoCorrection])
def test_2d_pressure_source_modes(mode):
    """ParitySourceTests.swift:22-101 (64^2, PML 10, 60 steps, centre point, 1 MHz 3-cycle tone)."""
    n = 64
    grid = KWaveGrid(n, 1e-4, n, 1e-4)
    grid.makeTime(1500, cfl=0.3)
    grid.setTime(60, grid.dt)
    mask = np.zeros((n, n), np.float32)
    mask[n // 2, n // 2] = 1
    sig = toneBurst(1 / grid.dt, 1e6, 3).astype(np.float32)
    sens = np.zeros((n, n), np.float32)
    sens[20, 16:48] = 1
    check(f"2d_psrc_{mode.value}", grid, KWaveMedium(1500.0, 1000.0), KWaveSource(pMask=mask, p=sig, pMode=mode),
          KWaveSensor(mask=sens), ["p", "pFinal"], pmlSize=10, smoothP0=False)


@pytest.mark.parametrize("mode", [SourceMode.dirichlet, SourceMode.additive, SourceMode.additiveNoCorrection])
def test_2d_velocity_source_modes(mode):
    n = 64
    grid = KWaveGrid(n, 1e-4, n, 1e-4)
    grid.setTime(60, 2e-8)
    mask = np.zeros((n, n), np.float32)
    mask[n // 2, n // 2 - 3:n // 2 + 3] = 1
    sig = toneBurst(1 / grid.dt, 1e6, 3).astype(np.float32)
    sens = np.zeros((n, n), np.float32)
    sens[20, 16:48] = 1
    check(f"2d_usrc_{mode.value}", grid, KWaveMedium(1500.0, 1000.0),
          KWaveSource(uMask=mask, ux=sig, uy=0.5 * sig, uMode=mode),
          KWaveSensor(mask=sens), ["p", "pFinal"], pmlSize=10, smoothP0=False)


def test_2d_per_point_signals():
    n = 48
    grid = KWaveGrid(n, 1e-4, n, 1e-4)
    grid.setTime(50, 2e-8)
    mask = np.zeros((n, n), np.float32)
    mask[24, 20:28] = 1
    base = toneBurst(1 / grid.dt, 1e6, 2).astype(np.float32)
    sig = np.stack([np.roll(np.pad(base, (0, 10)), d) for d in range(8)]).astype(np.float32)
    sens = np.zeros((n, n), np.float32)
    sens[10, :] = 1
    check("2d_rows", grid, KWaveMedium(1500.0, 1000.0), KWaveSource(uMask=mask, ux=sig, uMode=SourceMode.additive),
          KWaveSensor(mask=sens), ["p"], pmlSize=10, smoothP0=False)


def test_2d_record_all_fields():
    """ParityRecordTests.swift:35-94 (120 steps, line mask[20,16:48], full record list)."""
    n = 64
    grid = KWaveGrid(n, 1e-4, n, 1e-4)
    grid.makeTime(1500, cfl=0.3)
    grid.setTime(120, grid.dt)
    mask = np.zeros((n, n), np.float32)
    mask[n // 2, n // 2] = 1
    sig = toneBurst(1 / grid.dt, 1e6, 3).astype(np.float32)
    sm = np.zeros((n, n), np.float32)
    sm[20, 16:48] = 1
    rec = (RecordField.p | RecordField.ux | RecordField.uy | RecordField.pMax | RecordField.pMin | RecordField.pRms
           | RecordField.uMax | RecordField.uRms | RecordField.iAvg | RecordField.uFinal)
    fields = ["p", "ux", "uy", "pMax", "pMin", "pRms", "uxMax", "uyMax", "uxRms", "uyRms", "uxFinal", "uyFinal"]
    check("2d_record", grid, KWaveMedium(1500.0, 1000.0), KWaveSource(pMask=mask, p=sig),
          KWaveSensor(mask=sm, record=rec), fields, pmlSize=10, smoothP0=False)
    # intensity: 1e-3 in the reference (ParityRecordTests: Ix/Iy avg)
    check("2d_record_iavg", grid, KWaveMedium(1500.0, 1000.0), KWaveSource(pMask=mask, p=sig),
          KWaveSensor(mask=sm, record=rec), ["ixAvg", "iyAvg"], l2=1e-3, mx=1e-2, pmlSize=10, smoothP0=False)


def test_2d_cartesian_sensor():
    """ParityCartesianTests.swift:69-100: multilinear gather on a Cartesian circle."""
    n = 64
    dx = 1e-4
    grid = KWaveGrid(n, dx, n, dx)
    grid.setTime(150, 2e-8)
    pts = makeCartCircle(20 * dx, 40)
    check("2d_cart", grid, KWaveMedium(1500.0, 1000.0), KWaveSource(p0=makeDisc(n, n, radius=5)),
          KWaveSensor(mask=pts, record=RecordField.p | RecordField.pMax), ["p", "pMax"])


@pytest.mark.parametrize("pattern", [DirectivityPattern.pressure, DirectivityPattern.gradient])
def test_2d_directivity(pattern):
    """ParityDirectivityTests.swift:20-81."""
    n = 64
    grid = KWaveGrid(n, 1e-4, n, 1e-4)
    grid.setTime(120, 2e-8)
    sens = makeCircle(n, n, radius=20)
    ang = np.zeros((n, n), np.float32)
    idx = np.flatnonzero(sens.reshape(-1))
    ii, jj = np.unravel_index(idx, (n, n))
    ang.reshape(-1)[idx] = np.round(np.arctan2(jj - n // 2, ii - n // 2) * 4) / 4   # a handful of unique angles
    src = KWaveSource(p0=makeDisc(n, n, radius=5))
    sensor = KWaveSensor(mask=sens, directivityAngle=ang, directivityPattern=pattern)
    if pattern == DirectivityPattern.pressure:
        check("2d_dir_pressure", grid, KWaveMedium(1500.0, 1000.0), src, sensor, ["p"])
        return
    # The gradient weight k_n/|k| is odd in k, so Re(ifft2(W fft2 p)) keeps only the Nyquist-row
    # asymmetry: compare absolutely, against the pressure scale.
    print("\n[case] 2d_dir_gradient", flush=True)
    gpu = kspaceFirstOrder(grid, KWaveMedium(1500.0, 1000.0), src, sensor, SimulationOptions())
    ref = O.kspace_first_order(grid, KWaveMedium(1500.0, 1000.0), src, sensor, SimulationOptions())
    plain = O.kspace_first_order(grid, KWaveMedium(1500.0, 1000.0), src, KWaveSensor(mask=sens), SimulationOptions())
    scale = np.abs(plain["p"]).max()
    err = np.abs(gpu.p - ref["p"]).max() / scale
    print(f"  p abs err / max|p| = {err:.3e}", flush=True)
    assert err < 1e-5


@pytest.mark.parametrize("shape", [(45, 36), (33, 31)])
def test_2d_non_power_of_two_and_odd(shape):
    """Non-power-of-two and odd sizes (radix 3/5 and the odd-size smoothing placement)."""
    nx, ny = shape
    grid = KWaveGrid(nx, 1e-4, ny, 1e-4)
    grid.setTime(80, 2e-8)
    p0 = makeDisc(nx, ny, radius=4)
    check(f"2d_{nx}x{ny}", grid, KWaveMedium(1500.0, 1000.0), KWaveSource(p0=p0), KWaveSensor(), ["pFinal"],
          pmlSize=8, smoothP0=True)


def test_2d_smooth_media_maps():
    n = 48
    med = KWaveMedium(random_field((n, n), 21, 3.0, 1450.0, 1550.0), random_field((n, n), 22, 3.0, 950.0, 1050.0))
    grid = KWaveGrid(n, 1e-4, n, 1e-4)
    grid.setTime(100, 1.8e-8)
    check("2d_smoothc0rho0", grid, med, KWaveSource(p0=makeDisc(n, n, radius=4)), KWaveSensor(), ["pFinal"],
          smoothC0=True, smoothRho0=True, pmlSize=10)


# ------------------------------------------------------------------------------------------ 3D
def test_3d_ball_gpu_smoke_config():
    """GPUSmokeTests.swift:77-89 (24^3, ball r=3, tEnd 3e-7)."""
    n = 24
    grid = KWaveGrid(n, 1e-4, n, 1e-4, n, 1e-4)
    grid.makeTime(1500, cfl=0.3, tEnd=3e-7)
    check("3d_24", grid, KWaveMedium(1500.0, 1000.0), KWaveSource(p0=makeBall(n, n, n, radius=3)), KWaveSensor(),
          ["pFinal"], pmlSize=10)


@pytest.mark.parametrize("dtype", [F32, F64])
def test_3d_parity_config_32(dtype):
    """Parity3DTests.swift:24-81 (32^3, ball r=4, PML 10), smoothed, 60 steps with a plane sensor."""
    n = 32
    grid = KWaveGrid(n, 1e-4, n, 1e-4, n, 1e-4)
    grid.makeTime(1500, cfl=0.3)
    grid.setTime(60, grid.dt)
    sens = np.zeros((n, n, n), np.float32)
    sens[20, :, :] = 1
    check("3d_32", grid, KWaveMedium(1500.0, 1000.0), KWaveSource(p0=makeBall(n, n, n, radius=4)),
          KWaveSensor(mask=sens), ["pFinal", "p"], dtype, pmlSize=10)


@pytest.mark.parametrize("dtype", [F32, F64])
def test_3d_hetero_absorbing_plane_p_uz(dtype):
    """The 256^3 throughput workload at 32^3: GRF c/rho, power law y=1.1, ball p0, plane sensor p+uz."""
    n = 32
    c = random_field((n, n, n), 3, 4.0, 1400.0, 1600.0)
    rho = random_field((n, n, n), 4, 4.0, 900.0, 1100.0)
    med = KWaveMedium(c, rho, alphaCoeff=0.5, alphaPower=1.1, alphaMode=AbsorptionMode.powerLaw)
    grid = KWaveGrid(n, 1e-4, n, 1e-4, n, 1e-4)
    grid.setTime(80, 0.3 * 1e-4 / float(c.max()))
    sens = np.zeros((n, n, n), np.float32)
    sens[21, :, :] = 1
    check("3d_hetero_c2like", grid, med, KWaveSource(p0=makeBall(n, n, n, radius=4)),
          KWaveSensor(mask=sens, record=RecordField.p | RecordField.uz), ["p", "uz", "pFinal"], dtype, pmlSize=6)


def test_3d_non_cubic_mixed_radix():
    grid = KWaveGrid(20, 1e-4, 18, 1e-4, 15, 1e-4)
    grid.setTime(40, 2e-8)
    check("3d_20x18x15", grid, KWaveMedium(1500.0, 1000.0), KWaveSource(p0=makeBall(20, 18, 15, radius=3)),
          KWaveSensor(), ["pFinal"], pmlSize=4)


def test_3d_velocity_source_and_nonlinear():
    n = 24
    grid = KWaveGrid(n, 1e-4, n, 1e-4, n, 1e-4)
    grid.setTime(50, 2e-8)
    mask = np.zeros((n, n, n), np.float32)
    mask[6, 10:14, 10:14] = 1
    sig = (0.5 * toneBurst(1 / grid.dt, 2e6, 2)).astype(np.float32)   # 0.5 m/s: ~1 MPa, B/A visible
    sens = np.zeros((n, n, n), np.float32)
    sens[16, :, 12] = 1
    check("3d_usrc_nl", grid, KWaveMedium(1500.0, 1000.0, bOnA=8.0), KWaveSource(uMask=mask, ux=sig),
          KWaveSensor(mask=sens), ["p", "pFinal"], pmlSize=6, smoothP0=False)


def test_3d_corrected_sources_in_few_slabs():
    """k-space-corrected additive p and ux sources in different x planes (12 and 20) of a 64x32x32
    grid: the correction transforms only the x slabs 12..20 along y and z; against the oracle."""
    nx, n = 64, 32
    grid = KWaveGrid(nx, 1e-4, n, 1e-4, n, 1e-4)
    grid.setTime(60, 2e-8)   # 0.3 cells per step: both wavefronts pass the sensor plane
    pm = np.zeros((nx, n, n), np.float32)
    pm[12, 14:18, 14:18] = 1
    um = np.zeros((nx, n, n), np.float32)
    um[20, 10:22, 10:22] = 1
    sig = toneBurst(1 / grid.dt, 2e6, 2).astype(np.float32)
    src = KWaveSource(pMask=pm, p=(1e5 * sig).astype(np.float32), pMode=SourceMode.additive,
                      uMask=um, ux=(0.1 * sig).astype(np.float32), uMode=SourceMode.additive)
    sens = np.zeros((nx, n, n), np.float32)
    sens[26, :, 16] = 1
    check("3d_slab_sources", grid, KWaveMedium(1500.0, 1000.0), src, KWaveSensor(mask=sens), ["p", "pFinal"],
          pmlSize=8, smoothP0=False)


def test_corrected_source_slab_path_matches_full_path():
    """The same additive ux source once within one x row (the correction transforms only that row)
    and once with one more point, driven by a zero signal, 40 rows away (the source then spans more
    than half the x axis and the correction transforms the whole grid). In float64 both agree with
    each other and with the oracle to rounding. They are not bitwise equal: the full pass also
    stores the spectrum of the source row's partner in the two-rows-per-transform packing, whose
    exact value 0 comes out as rounding residue."""
    n = 64
    grid = KWaveGrid(n, 1e-4, n, 1e-4)
    grid.setTime(50, 2e-8)
    base = toneBurst(1 / grid.dt, 1e6, 2).astype(np.float32)
    near = np.zeros((n, n), np.float32)
    near[20, 28:36] = 1
    rows = np.stack([np.roll(np.pad(base, (0, 8)), d) for d in range(8)]).astype(np.float32)
    far = near.copy()
    far[60, 5] = 1
    rows_far = np.concatenate([rows, np.zeros((1, rows.shape[1]), np.float32)])
    sens = KWaveSensor(mask=np.ones((n, n), np.float32), record=RecordField.p | RecordField.pFinal)
    med = KWaveMedium(1500.0, 1000.0)
    src_near = KWaveSource(uMask=near, ux=rows, uMode=SourceMode.additive)
    print("\n[case] corrected source: slab path vs full path, 2D 64^2 float64", flush=True)
    a = kspaceFirstOrder(grid, med, src_near, sens, SimulationOptions(pmlSize=8, smoothP0=False, dtype=F64))
    b = kspaceFirstOrder(grid, med, KWaveSource(uMask=far, ux=rows_far, uMode=SourceMode.additive), sens,
                         SimulationOptions(pmlSize=8, smoothP0=False, dtype=F64))
    ref = O.kspace_first_order(grid, med, src_near, sens, SimulationOptions(pmlSize=8, smoothP0=False))
    for name, x, y in (("slab vs full", a.p, b.p), ("slab vs oracle", a.p, ref["p"]), ("full vs oracle", b.p, ref["p"])):
        rl2, rmax = metrics(x, y)
        print(f"  {name}: relL2={rl2:.3e} relMax={rmax:.3e}", flush=True)
        assert rl2 <= 1e-12 and rmax <= 1e-11, (name, rl2, rmax)


def test_lossless_closed_form_3d():
    """pmlAlpha = 0, homogeneous: the device loop against IFFT(P0 cos(c|k|t)) directly."""
    n, dx, c = 32, 1e-4, 1500.0
    p0 = makeBall(n, n, n, radius=5)
    grid = KWaveGrid(n, dx, n, dx, n, dx)
    grid.setTime(50, 0.3 * dx / c)
    out = kspaceFirstOrder(grid, KWaveMedium(c, 1000.0), KWaveSource(p0=p0), KWaveSensor(),
                           SimulationOptions(pmlAlpha=0.0, smoothP0=False, dtype=F64))
    ref = O.closed_form_lossless(p0.astype(np.float64), [n] * 3, [dx] * 3, c, grid.dt, [grid.nt - 1])[0]
    rl2, rmax = metrics(out.pFinal, ref)
    print(f"\n[case] closed_form_3d relL2={rl2:.3e} relMax={rmax:.3e}", flush=True)
    assert rl2 < 1e-10 and rmax < 1e-9


# ------------------------------------------------------------------------------------------ runtime
def test_chunked_run_and_state_roundtrip():
    n = 32
    grid = KWaveGrid(n, 1e-4, n, 1e-4)
    grid.setTime(60, 2e-8)
    sens = np.zeros((n, n), np.float32)
    sens[5, :] = 1
    args = (grid, KWaveMedium(1500.0, 1000.0), KWaveSource(p0=makeDisc(n, n, radius=4)),
            KWaveSensor(mask=sens, record=RecordField.p | RecordField.pMax))
    print("\n[case] chunked/state", flush=True)
    full = Simulation(*args)
    full.run()
    a = Simulation(*args)
    a.run(0, 25)
    a.run(25, 60)
    assert np.array_equal(full.output(L.OUT_P_FINAL), a.output(L.OUT_P_FINAL))
    assert np.array_equal(full.output(L.OUT_P), a.output(L.OUT_P))
    b = Simulation(*args)
    b.run(0, 25)
    blob = b.state()
    c = Simulation(*args)
    c.load_state(blob)
    c.run(25, 60)
    assert np.array_equal(full.output(L.OUT_P_FINAL), c.output(L.OUT_P_FINAL))
    assert np.array_equal(full.output(L.OUT_P_MAX), c.output(L.OUT_P_MAX))
    with pytest.raises(RuntimeError):
        c.run(10, 20)   # non-contiguous


def test_progress_and_field_monitor_cadence():
    """VizTests.swift:41-60 / SolverTests progress count: fieldMonitor at t % 5 == 0 with p, progress every step."""
    n = 32
    grid = KWaveGrid(n, 1e-4, n, 1e-4)
    grid.setTime(16, 2e-8)
    seen, prog = [], []
    opts = SimulationOptions(fieldMonitor=lambda t, p: seen.append((t, p.shape)), fieldMonitorInterval=5,
                             progress=lambda t, nt: prog.append(t))
    kspaceFirstOrder(grid, KWaveMedium(1500.0, 1000.0), KWaveSource(p0=makeDisc(n, n, radius=3)), KWaveSensor(), opts)
    assert [t for t, _ in seen] == [0, 5, 10, 15] and all(s == (n, n) for _, s in seen)
    assert prog == list(range(16))


def test_callback_fields_equal_the_state_after_their_step():
    """kwh_step_cb: one call per step in order (steps run in chunks between syncs), the pressure
    after step t delivered when t % interval == 0, none with interval 0 (progress only)."""
    n = 32
    grid = KWaveGrid(n, 1e-4, n, 1e-4)
    grid.setTime(150, 2e-8)
    args = (grid, KWaveMedium(1500.0, 1000.0), KWaveSource(p0=makeDisc(n, n, radius=3)), KWaveSensor())
    seen = {}
    sim = Simulation(*args)
    sim.run(callback=lambda t, nt, p: seen.__setitem__(t, None if p is None else p.copy()), interval=7)
    assert list(seen) == list(range(150))
    assert [t for t, v in seen.items() if v is not None] == list(range(0, 150, 7))
    for t in (0, 7, 98, 147):
        ref = Simulation(*args)
        ref.run(0, t + 1)
        assert np.array_equal(seen[t], ref.output(L.OUT_P_FINAL)), t
    prog = []
    quiet = Simulation(*args)
    quiet.run(callback=lambda t, nt, p: prog.append((t, nt, p is None)), interval=0)
    assert prog == [(t, 150, True) for t in range(150)]
    assert np.array_equal(quiet.output(L.OUT_P_FINAL), sim.output(L.OUT_P_FINAL))


def test_deterministic_replay():
    n = 48
    grid = KWaveGrid(n, 1e-4, n, 1e-4)
    grid.setTime(40, 2e-8)
    args = (grid, _hetero_2d(n, seed=9), KWaveSource(p0=makeDisc(n, n, radius=4)), KWaveSensor())
    a = kspaceFirstOrder(*args)
    b = kspaceFirstOrder(*args)
    assert np.array_equal(a.pFinal, b.pFinal)


def _outputs_equal(a, b, fields):
    for f in fields:
        x, y = getattr(a, f), getattr(b, f)
        assert x is not None and y is not None, f
        assert np.array_equal(x, y), (f, float(np.abs(np.asarray(x, np.float64) - y).max()))


def _run_chunks(make, chunks):
    """Run a fresh Simulation in [t0, t1) chunks, graph replay on or off per chunk."""
    sim = make()
    try:
        for t0, t1, graph in chunks:
            sim.set_graph(graph)
            sim.run(t0, t1)
        return sim.outputs()
    finally:
        sim.close()


def test_graph_replay_matches_eager_2d_sources():
    """Captured step graphs (one per set of active k-space-corrected sources) replayed for every
    step, alternating with eager chunks: bitwise equal to an all-eager run. The additive source
    ends at step 40, so the replay switches graphs mid-run; the 64^2 grid takes the
    register-resident kernels."""
    n = 64
    grid = KWaveGrid(n, 1e-4, n, 1e-4)
    grid.setTime(120, 0.3 * 1e-4 / 1600.0)
    med = _hetero_2d(n, seed=21)
    mask = np.zeros((n, n), np.float32)
    mask[10, 20:44] = 1
    sig = (1e5 * toneBurst(1 / grid.dt, 2e6, 2)[:40]).astype(np.float32)
    src = KWaveSource(pMask=mask, p=sig, pMode=SourceMode.additive, p0=makeDisc(n, n, radius=5))
    line = np.zeros((n, n), np.float32)
    line[44, 8:56] = 1
    sens = KWaveSensor(mask=line, record=RecordField.p | RecordField.pMax | RecordField.pRms | RecordField.ux |
                       RecordField.uy | RecordField.pFinal)
    make = lambda: Simulation(grid, med, src, sens, SimulationOptions(pmlSize=10))
    print("\n[case] graph replay vs eager, 2D 64^2 with an additive source", flush=True)
    eager = _run_chunks(make, [(0, 120, False)])
    mixed = _run_chunks(make, [(0, 30, True), (30, 55, False), (55, 90, True), (90, 120, True)])
    _outputs_equal(eager, mixed, ["p", "pMax", "pRms", "ux", "uy", "pFinal"])


def test_graph_replay_matches_eager_3d_velocity_source():
    """3D 32^3 (generic kernels): additive and dirichlet velocity sources, graph replay from step 0
    (no p0, so step 0 is a captured step too) against eager."""
    n = 32
    grid = KWaveGrid(n, 1e-4, n, 1e-4, n, 1e-4)
    grid.setTime(60, 2e-8)
    mask = np.zeros((n, n, n), np.float32)
    mask[8, 12:20, 12:20] = 1
    sig = (0.3 * toneBurst(1 / grid.dt, 3e6, 2)).astype(np.float32)
    src = KWaveSource(uMask=mask, ux=sig, uz=sig, uMode=SourceMode.additive)
    sens_mask = np.zeros((n, n, n), np.float32)
    sens_mask[20, :, 16] = 1
    sens = KWaveSensor(mask=sens_mask, record=RecordField.p | RecordField.ux | RecordField.uz | RecordField.uMax)
    make = lambda: Simulation(grid, KWaveMedium(1500.0, 1000.0), src, sens, SimulationOptions(pmlSize=6))
    print("\n[case] graph replay vs eager, 3D 32^3 velocity source", flush=True)
    eager = _run_chunks(make, [(0, 60, False)])
    graph = _run_chunks(make, [(0, 60, True)])
    _outputs_equal(eager, graph, ["p", "ux", "uz", "uxMax", "uzMax"])


@pytest.mark.parametrize("shape", [(64, 64), (256, 256), (32, 32, 32), (64, 64, 64)])
def test_autotuned_schedule_is_bitwise_neutral(shape):
    """The autotuner only changes launch geometry (tile columns, row pairs per workgroup, LDS
    budget); every transform's arithmetic is unchanged, so results are bitwise those of the
    default schedule."""
    dim = len(shape)
    args = [v for n in shape for v in (n, 1e-4)]
    grid = KWaveGrid(*args)
    grid.setTime(30, 2e-8)
    p0 = makeDisc(*shape, radius=4) if dim == 2 else makeBall(*shape, radius=4)
    mask = np.zeros(shape, np.float32)
    mask[(shape[0] * 3) // 4] = 1
    med = KWaveMedium(random_field(shape, 17, 4.0, 1450.0, 1600.0), 1000.0, alphaCoeff=0.5, alphaPower=1.3)
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


@pytest.mark.parametrize("dtype", [F32, F64])
def test_time_reversal(dtype):
    """timeReversal (TimeReversal.swift:20-51) on the device, flip and positivity included,
    against the oracle's loop run with the same recording flipped on the host as a Dirichlet
    source, then max(0, 2 pFinal)."""
    n = 64
    grid = KWaveGrid(n, 1e-4, n, 1e-4)
    grid.makeTime(1500.0, cfl=0.3)
    med = KWaveMedium(1500.0, 1000.0)
    mask = makeCircle(n, n, radius=26)
    opts = SimulationOptions(pmlSize=10)
    rec = O.kspace_first_order(grid, med, KWaveSource(p0=makeDisc(n, n, radius=6)), KWaveSensor(mask=mask), opts)["p"]
    rec = rec.astype(np.float32)
    print(f"\n[case] time reversal {dtype.value} 64^2, {rec.shape[0]} sensor points x {rec.shape[1]} steps", flush=True)
    est = timeReversal(grid, med, KWaveSensor(mask=mask), rec, 2.0, SimulationOptions(pmlSize=10, dtype=dtype))
    src = KWaveSource(pMask=mask, p=np.ascontiguousarray(rec[:, ::-1]), pMode=SourceMode.dirichlet)
    ref = np.maximum(2.0 * O.kspace_first_order(grid, med, src, KWaveSensor(), opts)["pFinal"], 0.0)
    rl2, rmax = metrics(est, ref)
    print(f"  relL2={rl2:.3e} relMax={rmax:.3e}", flush=True)
    assert est.shape == (n, n) and np.all(est >= 0)
    assert rl2 <= (1e-4 if dtype == F32 else 1e-9) and rmax <= (1e-3 if dtype == F32 else 1e-8)


# ------------------------------------------------------------------------------------------ register-resident path
@pytest.mark.parametrize("absorbing,mode,nonlinear", [
    (False, AbsorptionMode.powerLaw, False), (False, AbsorptionMode.powerLaw, True),
    (True, AbsorptionMode.stokes, False), (True, AbsorptionMode.stokes, True),
    (True, AbsorptionMode.noDispersion, False), (True, AbsorptionMode.powerLaw, False)])
def test_2d_64_register_resident_density_variants(absorbing, mode, nonlinear):
    """Every density-kernel variant on the register-resident 64^2 path: EOS, EOS with Stokes
    absorption, FFT absorption with and without dispersion; linear and B/A."""
    n = 64
    med = _hetero_2d(n, seed=11, absorbing=absorbing, nonlinear=nonlinear, mode=mode, y=1.4)
    grid = KWaveGrid(n, 1e-4, n, 1e-4)
    grid.setTime(160, 0.3 * 1e-4 / float(np.max(med.soundSpeed)))
    check(f"2d_64_density_{absorbing}_{mode.value}_{nonlinear}", grid, med,
          KWaveSource(p0=makeDisc(n, n, radius=6)), KWaveSensor(), ["pFinal"], pmlSize=10)


def test_3d_64_register_resident_stokes():
    """3D Stokes absorption (y = 2) with B/A on the radix-8 x 8 path."""
    n = 64
    c = random_field((n, n, n), 51, 6.0, 1400.0, 1600.0)
    med = KWaveMedium(c, 1000.0, alphaCoeff=0.5, alphaPower=2.0, alphaMode=AbsorptionMode.stokes, bOnA=4.0)
    grid = KWaveGrid(n, 1e-4, n, 1e-4, n, 1e-4)
    grid.setTime(50, 0.3 * 1e-4 / float(c.max()))
    check("3d_64_stokes", grid, med, KWaveSource(p0=makeBall(n, n, n, radius=6)), KWaveSensor(), ["pFinal"],
          pmlSize=8, workers=8)


def test_2d_256_register_resident_path():
    """256^2 runs the register-resident radix-16 x 16 kernels on both axes."""
    n = 256
    grid = KWaveGrid(n, 1e-4, n, 1e-4)
    grid.setTime(120, 2e-8)
    med = _hetero_2d(n, seed=31)
    sens = np.zeros((n, n), np.float32)
    sens[100, :] = 1     # 28 cells from the disc: reached within the run
    check("2d_256_sq", grid, med, KWaveSource(p0=makeDisc(n, n, radius=6)),
          KWaveSensor(mask=sens, record=RecordField.p | RecordField.uy), ["p", "uy"], pmlSize=20)


@pytest.mark.parametrize("dtype", [F32, F64])
def test_3d_64_register_resident_path(dtype):
    """64^3 runs the radix-8 x 8 register-resident kernels on every axis."""
    n = 64
    c = random_field((n, n, n), 41, 6.0, 1400.0, 1600.0)
    rho = random_field((n, n, n), 42, 6.0, 900.0, 1100.0)
    med = KWaveMedium(c, rho, alphaCoeff=0.5, alphaPower=1.1, alphaMode=AbsorptionMode.powerLaw, bOnA=5.0)
    grid = KWaveGrid(n, 1e-4, n, 1e-4, n, 1e-4)
    grid.setTime(60, 0.3 * 1e-4 / float(c.max()))
    sens = np.zeros((n, n, n), np.float32)
    sens[16, :, :] = 1   # 6 cells downstream of the velocity source: the burst dominates the series
    mask = np.zeros((n, n, n), np.float32)
    mask[10, 28:36, 28:36] = 1
    sig = (0.2 * toneBurst(1 / grid.dt, 2e6, 2)).astype(np.float32)
    check("3d_64_sq", grid, med, KWaveSource(p0=makeBall(n, n, n, radius=5), uMask=mask, ux=sig),
          KWaveSensor(mask=sens, record=RecordField.p | RecordField.ux | RecordField.uz), ["p", "ux", "uz"], dtype,
          pmlSize=8)


def test_generic_and_register_paths_agree():
    """The generic Stockham engine and the register-resident kernels on the same 64^3 problem."""
    import os
    n = 64
    grid = KWaveGrid(n, 1e-4, n, 1e-4, n, 1e-4)
    grid.setTime(30, 2e-8)
    args = (grid, KWaveMedium(1500.0, 1000.0), KWaveSource(p0=makeBall(n, n, n, radius=6)), KWaveSensor())
    print("\n[case] generic vs register-resident 64^3", flush=True)
    a = kspaceFirstOrder(*args, SimulationOptions(dtype=F64))
    os.environ["KWH_GENERIC_FFT"] = "1"
    try:
        b = kspaceFirstOrder(*args, SimulationOptions(dtype=F64))
    finally:
        del os.environ["KWH_GENERIC_FFT"]
    rl2, rmax = metrics(a.pFinal, b.pFinal)
    print(f"  relL2={rl2:.3e} relMax={rmax:.3e}", flush=True)
    assert rl2 < 1e-12


def test_3d_256_fp32_vs_fp64_100_steps():
    """FP32 against FP64 on the device at 256^3 over 100 steps: relative L2 of the plane p / uz
    series and of pFinal normalised by max_t |p| (target <= 1e-4)."""
    n = 256
    c = random_field((n, n, n), 3, 20.0, 1400.0, 1600.0)
    rho = random_field((n, n, n), 4, 20.0, 900.0, 1100.0)
    med = KWaveMedium(c, rho, alphaCoeff=0.5, alphaPower=1.1, alphaMode=AbsorptionMode.powerLaw)
    grid = KWaveGrid(n, 1e-4, n, 1e-4, n, 1e-4)
    grid.setTime(100, 0.3 * 1e-4 / float(c.max()))
    sens = np.zeros((n, n, n), np.float32)
    sens[150, :, :] = 1
    args = (grid, med, KWaveSource(p0=makeBall(n, n, n, radius=8)),
            KWaveSensor(mask=sens, record=RecordField.p | RecordField.uz))
    print("\n[case] 3d_256 fp32 vs fp64, 100 steps", flush=True)
    a = kspaceFirstOrder(*args, SimulationOptions(pmlSize=10, dtype=F32))
    b = kspaceFirstOrder(*args, SimulationOptions(pmlSize=10, dtype=F64))
    for f in ("p", "uz"):
        rl2, rmax = metrics(getattr(a, f), getattr(b, f))
        print(f"  {f} relL2={rl2:.3e} relMax={rmax:.3e}", flush=True)
        assert rl2 <= 1e-4
    scale = np.abs(b.p).max()
    d = a.pFinal.astype(np.float64) - b.pFinal
    rl2 = math.sqrt(np.mean(d * d)) / scale
    print(f"  pFinal/max|p| relL2={rl2:.3e}", flush=True)
    assert rl2 <= 1e-4
