"""GPU: transducer drive synthesised on the device, receive combination on the device, and the
batched runner (kwh_batch_*) against single runs, through the C ABI."""
import ctypes as C
import math

import numpy as np
import pytest

from oracle import kwave_oracle as O
from kwavehip import (KWaveGrid, KWaveMedium, KWaveSensor, KWaveSource, KWaveTransducer, KWaveTransducerGeometry,
                      RecordField, Simulation, SimulationBatch, SimulationOptions, SourceMode, toneBurst)
from kwavehip import _lib as L

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _device():
    lib = L.load()
    h = C.c_void_p()
    if lib.kwh_ctx_create(0, C.byref(h)) != 0:
        pytest.skip("no HIP device")
    lib.kwh_ctx_destroy(h)


def scan_line_case(shape, seed, steer, focus, nt=140, dx=1e-4):
    """One transmit of a 16-element array (width 2, length 8) at x = 4: speckle medium, delayed
    velocity drive, sensor on the active voxels recording p."""
    rng = np.random.default_rng(seed)
    nx, ny, nz = shape
    c = (1540.0 * (1 + rng.uniform(-0.02, 0.02, shape))).astype(np.float32)
    rho = (1000.0 * (1 + rng.uniform(-0.02, 0.02, shape))).astype(np.float32)
    grid = KWaveGrid(nx, dx, ny, dx, nz, dx)
    grid.setTime(nt, 0.3 * dx / float(c.max()))
    geom = KWaveTransducerGeometry(numberElements=16, elementWidth=2, elementLength=8, elementSpacing=0,
                                   position=(4, (ny - 32) // 2, (nz - 8) // 2), dy=dx, dz=dx)
    burst = (0.5 * toneBurst(1 / grid.dt, 2e6, 2)).astype(np.float32)
    tr = KWaveTransducer(geom, shape, grid.dt, 1540.0, focusDistance=focus, elevationFocusDistance=6e-3,
                         steeringAngle=steer, transmitApodization="hanning", inputSignal=burst)
    mask, row, delays, scales = tr.velocity_source()
    src = KWaveSource(uMask=mask, ux=row, uDelay=delays, uScale=scales, uMode=SourceMode.additive)
    sens = KWaveSensor(mask=mask, record=RecordField.p)
    return grid, KWaveMedium(c, rho), src, sens, SimulationOptions(pmlSize=6, smoothP0=False), tr


def expanded(src, grid):
    """The same drive with the per-point rows built on the host."""
    row = np.asarray(src.ux, np.float32)
    d = np.asarray(src.uDelay)
    s = np.asarray(src.uScale, np.float32)
    L_ = row.size + int(d.max())
    rows = np.zeros((d.size, L_), np.float32)
    for k in range(d.size):
        rows[k, d[k]:d[k] + row.size] = s[k] * row
    return KWaveSource(uMask=src.uMask, ux=rows, uMode=src.uMode)


def test_device_drive_synthesis_matches_host_rows_and_oracle():
    grid, med, src, sens, opts, tr = scan_line_case((32, 48, 24), 1, 12.0, 8e-3, nt=90)
    print("\n[case] delayed-copy drive, 32x48x24", flush=True)
    a = Simulation(grid, med, src, sens, opts)
    a.run()
    pa = a.outputs().p
    a.close()
    host = expanded(src, grid)
    b = Simulation(grid, med, host, sens, opts)
    b.run()
    pb = b.outputs().p
    b.close()
    assert np.array_equal(pa, pb)
    ref = O.kspace_first_order(grid, med, host, sens, opts, workers=8)["p"]
    d = np.asarray(pa, np.float64) - ref
    rl2 = math.sqrt(np.mean(d * d)) / math.sqrt(np.mean(ref * ref))
    print(f"  p relL2={rl2:.3e}", flush=True)
    assert rl2 < 1e-4


def test_element_traces_match_host_combination():
    grid, med, src, sens, opts, tr = scan_line_case((32, 48, 24), 2, -8.0, np.inf, nt=80)
    print("\n[case] device receive combination", flush=True)
    sim = Simulation(grid, med, src, sens, opts)
    sim.run()
    p = sim.outputs().p
    dev = sim.transducer_traces(tr)
    sim.close()
    host = tr.combineSensorData(p)
    scale = float(np.abs(host).max())
    assert dev.shape == host.shape
    assert np.abs(dev - host).max() <= 1e-6 * scale


@pytest.mark.parametrize("use_graph", [True, False])
def test_batch_of_uneven_sizes_equals_single_runs(use_graph):
    """Three sims of different sizes and drives (their sources switch off at different steps,
    so the batched graph is rebuilt mid-run): bitwise equal to separate runs."""
    cases = [scan_line_case((32, 48, 24), 3, 0.0, 8e-3, nt=70), scan_line_case((40, 40, 24), 4, 15.0, np.inf, nt=90),
             scan_line_case((36, 48, 32), 5, -10.0, 6e-3, nt=60)]
    print(f"\n[case] batch of 3 uneven sims, graph={use_graph}", flush=True)
    singles = []
    for grid, med, src, sens, opts, _ in cases:
        s = Simulation(grid, med, src, sens, opts)
        s.run()
        singles.append(s.outputs())
        s.close()
    batch = SimulationBatch([c[:5] for c in cases])
    batch.run(use_graph=use_graph)
    for i, (sim, ref) in enumerate(zip(batch.sims, singles)):
        out = sim.outputs()
        assert out.p.shape == ref.p.shape, i
        assert np.array_equal(out.p, ref.p), (i, float(np.abs(out.p - ref.p).max()))
        assert np.array_equal(out.pFinal, ref.pFinal), i
        tr_out = sim.transducer_traces(cases[i][5])
        assert tr_out.shape == (16, cases[i][0].nt)
    batch.close()
