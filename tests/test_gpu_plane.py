"""GPU parity of plane mode (kwh_sim_set_plane): the y and z transforms of each field in one
kernel per x plane. Checked against the float64 oracle (the reference's algorithm) at the
reference's tolerances, against the split passes (same physics, other factorisation), and for
bitwise equality of graph replay, batched runs and mode switches with single eager runs.
"""
import ctypes as C
import math

import numpy as np
import pytest

from oracle import kwave_oracle as O
from kwavehip import (AbsorptionMode, DTypePrecision, KWaveGrid, KWaveMedium, KWaveSensor, KWaveSource, RecordField,
                      Simulation, SimulationBatch, SimulationOptions, SourceMode, makeBall, random_field, toneBurst)
from kwavehip import _lib as L
from kwavehip import workloads

pytestmark = pytest.mark.gpu

F32 = DTypePrecision.float32
F64 = DTypePrecision.float64


@pytest.fixture(scope="module", autouse=True)
def _device():
    lib = L.load()
    h = C.c_void_p()
    if lib.kwh_ctx_create(0, C.byref(h)) != 0:
        pytest.skip("no HIP device")
    lib.kwh_ctx_destroy(h)


def metrics(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    d = a - b
    return (math.sqrt(np.mean(d * d)) / max(math.sqrt(np.mean(b * b)), 1e-300),
            np.abs(d).max() / max(np.abs(b).max(), 1e-300))


def run(args, plane, graph=False, chunks=None):
    sim = Simulation(*args)
    try:
        sim.set_plane(bool(plane))   # plane mode is the default where it applies
        if graph:
            sim.set_graph(True)
        if chunks:
            t = 0
            for t1, mode in chunks:
                sim.set_plane(mode)
                sim.run(t, t1)
                t = t1
            sim.run(t, args[0].nt)
        else:
            sim.run()
        return sim.outputs()
    finally:
        sim.close()


def ivp_case(shape, dtype, nt=30):
    n = shape
    c = random_field(n, 21, 6.0, 1450.0, 1600.0)
    rho = random_field(n, 22, 6.0, 950.0, 1100.0)
    grid = KWaveGrid(n[0], 1e-4, n[1], 1e-4, n[2], 1e-4)
    grid.setTime(nt, 0.3 * 1e-4 / float(c.max()))
    p0 = makeBall(*n, cx=n[0] // 2, cy=n[1] // 2 + 5, cz=n[2] // 2 - 7, radius=4)
    mask = np.zeros(n, np.float32)
    mask[n[0] // 2, ::3, ::5] = 1
    mask[2, 10:40, 7] = 1
    sens = KWaveSensor(mask=mask, record=RecordField.p | RecordField.pMax | RecordField.pRms | RecordField.pFinal)
    opts = SimulationOptions(pmlSize=4, smoothP0=True, dtype=dtype)
    return grid, KWaveMedium(c, rho), KWaveSource(p0=p0), sens, opts


def sources_case(shape, dtype, nt=30):
    """Stokes absorption, B/A, a k-space corrected additive pressure source on a line and a
    dirichlet ux source on a patch."""
    n = shape
    c = random_field(n, 31, 6.0, 1480.0, 1560.0)
    grid = KWaveGrid(n[0], 1e-4, n[1], 1e-4, n[2], 1e-4)
    grid.setTime(nt, 0.3 * 1e-4 / float(c.max()))
    med = KWaveMedium(c, 1000.0, alphaCoeff=0.8, alphaPower=2.0, alphaMode=AbsorptionMode.stokes, bOnA=5.0)
    pm = np.zeros(n, np.float32)
    pm[3, n[1] // 2, 10:60] = 1
    sig = (2e5 * toneBurst(1 / grid.dt, 1.5e6, 2)).astype(np.float32)
    um = np.zeros(n, np.float32)
    um[n[0] - 4, 20:24, 30:36] = 1
    usig = (0.05 * toneBurst(1 / grid.dt, 2e6, 2)).astype(np.float32)
    src = KWaveSource(pMask=pm, p=sig, pMode=SourceMode.additive, uMask=um, ux=usig, uMode=SourceMode.dirichlet)
    mask = np.zeros(n, np.float32)
    mask[n[0] // 2, :, n[2] // 3] = 1
    sens = KWaveSensor(mask=mask, record=RecordField.p | RecordField.pFinal)
    return grid, med, src, sens, SimulationOptions(pmlSize=4, smoothP0=False, dtype=dtype)


SHAPES = [(12, 96, 96), (10, 128, 128), (16, 64, 64), (12, 32, 32)]


def absorbing_case(shape, dtype, alpha_map=False, dispersion=True, nonlinear=False, nt=30):
    """Power-law absorption (the FFT term of the EOS), optionally with an alpha map (tau / eta
    maps instead of evaluating them from c), without dispersion, or with B/A."""
    n = shape
    c = random_field(n, 41, 6.0, 1450.0, 1600.0)
    rho = random_field(n, 42, 6.0, 950.0, 1100.0)
    grid = KWaveGrid(n[0], 1e-4, n[1], 1e-4, n[2], 1e-4)
    grid.setTime(nt, 0.3 * 1e-4 / float(c.max()))
    alpha = random_field(n, 43, 6.0, 0.4, 1.2) if alpha_map else 0.75
    med = KWaveMedium(c, rho, alphaCoeff=alpha, alphaPower=1.5, alphaMode=AbsorptionMode.powerLaw,
                      bOnA=5.0 if nonlinear else 0.0)
    if not dispersion:
        med = KWaveMedium(c, rho, alphaCoeff=alpha, alphaPower=1.5, alphaMode=AbsorptionMode.noDispersion,
                          bOnA=5.0 if nonlinear else 0.0)
    p0 = makeBall(*n, cx=n[0] // 2, cy=n[1] // 2 + 3, cz=n[2] // 2 - 4, radius=4)
    mask = np.zeros(n, np.float32)
    mask[n[0] // 2, ::3, ::5] = 1
    sens = KWaveSensor(mask=mask, record=RecordField.p | RecordField.pFinal)
    return grid, med, KWaveSource(p0=p0), sens, SimulationOptions(pmlSize=4, smoothP0=True, dtype=dtype)


@pytest.mark.parametrize("dtype", [F32, F64])
@pytest.mark.parametrize("shape,kind", [((12, 96, 96), "c"), ((10, 128, 128), "c"), ((16, 64, 64), "map_nl"),
                                        ((12, 32, 32), "nodisp")])
def test_plane_power_law_absorption(shape, kind, dtype):
    args = absorbing_case(shape, dtype, alpha_map=kind == "map_nl", nonlinear=kind == "map_nl",
                          dispersion=kind != "nodisp")
    print(f"\n[case] plane power law {shape} {kind} {dtype.value}", flush=True)
    gp = run(args, plane=True)
    gs = run(args, plane=False)
    ref = O.kspace_first_order(*args[:4], SimulationOptions(pmlSize=4, smoothP0=True), workers=4)
    for f in ("p", "pFinal"):
        rl2, rmax = metrics(getattr(gp, f), ref[f])
        sl2, _ = metrics(getattr(gp, f), getattr(gs, f))
        print(f"  {f:6s} vs oracle relL2={rl2:.2e} relMax={rmax:.2e}; vs split passes relL2={sl2:.2e}", flush=True)
        if dtype == F32:
            assert rl2 <= 1e-4 and rmax <= 1e-3, (f, rl2, rmax)
            assert sl2 <= 1e-5, (f, sl2)
        else:
            assert rl2 <= 1e-9 and rmax <= 1e-8, (f, rl2, rmax)
            assert sl2 <= 1e-12, (f, sl2)


@pytest.mark.parametrize("dtype", [F32, F64])
@pytest.mark.parametrize("shape", SHAPES)
def test_plane_ivp_heterogeneous(shape, dtype):
    args = ivp_case(shape, dtype)
    print(f"\n[case] plane ivp {shape} {dtype.value}", flush=True)
    gp = run(args, plane=True)
    gs = run(args, plane=False)
    ref = O.kspace_first_order(*args[:4], SimulationOptions(pmlSize=4, smoothP0=True), workers=4)
    for f in ("p", "pMax", "pRms", "pFinal"):
        rl2, rmax = metrics(getattr(gp, f), ref[f])
        sl2, _ = metrics(getattr(gp, f), getattr(gs, f))
        print(f"  {f:6s} vs oracle relL2={rl2:.2e} relMax={rmax:.2e}; vs split passes relL2={sl2:.2e}", flush=True)
        if dtype == F32:
            assert rl2 <= 1e-4 and rmax <= 1e-3, (f, rl2, rmax)
            assert sl2 <= 1e-5, (f, sl2)
        else:
            assert rl2 <= 1e-9 and rmax <= 1e-8, (f, rl2, rmax)
            assert sl2 <= 1e-12, (f, sl2)


@pytest.mark.parametrize("dtype", [F32, F64])
@pytest.mark.parametrize("shape", SHAPES)
def test_plane_sources_stokes_nonlinear(shape, dtype):
    args = sources_case(shape, dtype)
    print(f"\n[case] plane sources {shape} {dtype.value}", flush=True)
    gp = run(args, plane=True)
    ref = O.kspace_first_order(*args[:4], SimulationOptions(pmlSize=4, smoothP0=False), workers=4)
    for f in ("p", "pFinal"):
        rl2, rmax = metrics(getattr(gp, f), ref[f])
        print(f"  {f:6s} vs oracle relL2={rl2:.2e} relMax={rmax:.2e}", flush=True)
        if dtype == F32:
            assert rl2 <= 1e-4 and rmax <= 1e-3, (f, rl2, rmax)
        else:
            assert rl2 <= 1e-9 and rmax <= 1e-8, (f, rl2, rmax)


def scan_cases(nt=40):
    """Scan-line transmits at the plane shapes: the transducer's k-space corrected additive ux
    source with per-voxel delays and scales, receive on the active voxels."""
    return [workloads.scan_line_transmit(r, shape, nt) for r, shape in
            enumerate([(24, 96, 96), (24, 96, 96), (20, 128, 128)])]


def host_rows(src):
    """The transducer drive with its per-point delayed rows built on the host (the oracle's input
    form: each active voxel plays the padded signal delayed by its delayMask entry and scaled by its
    element's apodization, KWaveTransducer.swift:182-212, 246-260)."""
    row = np.asarray(src.ux, np.float32)
    d = np.asarray(src.uDelay)
    s = np.asarray(src.uScale, np.float32)
    rows = np.zeros((d.size, row.size + int(d.max())), np.float32)
    for k in range(d.size):
        rows[k, d[k]:d[k] + row.size] = s[k] * row
    return KWaveSource(uMask=src.uMask, ux=rows, uMode=src.uMode)


def test_plane_scan_line_vs_oracle():
    (args, tr) = scan_cases()[2]
    print("\n[case] plane scan line 20x128x128 f32", flush=True)
    gp = run(args, plane=True)
    ref = O.kspace_first_order(args[0], args[1], host_rows(args[2]), args[3], args[4], workers=4)
    rl2, rmax = metrics(gp.p, ref["p"])
    print(f"  p vs oracle relL2={rl2:.2e} relMax={rmax:.2e}", flush=True)
    assert rl2 <= 1e-4 and rmax <= 1e-3, (rl2, rmax)


def test_plane_graph_batch_bitwise():
    cases = scan_cases()
    print("\n[case] plane batch (2 x 24x96x96, 20x128x128) vs single eager runs in plane mode", flush=True)
    ref = [run(a, plane=True).p for a, _ in cases]
    graph = [run(a, plane=True, graph=True).p for a, _ in cases]
    for a, b in zip(ref, graph):
        assert np.array_equal(a, b)
    for use_graph, fused in ((True, True), (False, True), (True, False)):
        batch = SimulationBatch([a for a, _ in cases])
        try:
            batch.run(use_graph=use_graph, fused=fused, plane=True)
            for sim, r in zip(batch.sims, ref):
                assert np.array_equal(sim.output(L.OUT_P), r), (use_graph, fused)
        finally:
            batch.close()


def test_plane_mode_switch_and_state():
    """Switching modes between steps converts the carried pressure spectrum; a state blob saved in
    plane mode reloads into either mode."""
    args = ivp_case((12, 96, 96), F64, nt=24)
    full = run(args, plane=True)
    mixed = run(args, plane=True, chunks=[(5, False), (11, True), (17, False)])
    for f in ("p", "pFinal"):
        rl2, _ = metrics(getattr(mixed, f), getattr(full, f))
        print(f"  switch {f}: relL2 vs plane run {rl2:.2e}", flush=True)
        assert rl2 <= 1e-12, (f, rl2)
    sim = Simulation(*args)
    try:
        sim.set_plane(True)
        sim.run(0, 9)
        blob = sim.state()
        sim.run(9, args[0].nt)
        a = sim.outputs().pFinal
        sim.load_state(blob)
        sim.set_plane(False)
        sim.run(9, args[0].nt)
        b = sim.outputs().pFinal
    finally:
        sim.close()
    rl2, _ = metrics(b, a)
    print(f"  reload into split mode: relL2 {rl2:.2e}", flush=True)
    assert rl2 <= 1e-12


def test_plane_unsupported_configurations():
    n = (8, 96, 96)
    grid = KWaveGrid(n[0], 1e-4, n[1], 1e-4, n[2], 1e-4)
    grid.setTime(4, 2e-8)
    p0 = makeBall(*n, radius=3)
    sens = KWaveSensor(mask=np.ones(n, np.float32), record=RecordField.p)
    colloc = KWaveSensor(mask=np.ones(n, np.float32), record=RecordField.p | RecordField.ux)
    for med, sen in ((KWaveMedium(1500.0, 1000.0), colloc),):
        sim = Simulation(grid, med, KWaveSource(p0=p0), sen, SimulationOptions(pmlSize=2))
        try:
            with pytest.raises(NotImplementedError):   # KWH_EUNSUPPORTED
                sim.set_plane(True)
        finally:
            sim.close()
    g2 = KWaveGrid(48, 1e-4, 40, 1e-4, 40, 1e-4)   # no plane kernels for a 40 x 40 plane
    g2.setTime(4, 2e-8)
    sim = Simulation(g2, KWaveMedium(1500.0, 1000.0), KWaveSource(p0=makeBall(48, 40, 40, radius=3)),
                     KWaveSensor(mask=np.ones((48, 40, 40), np.float32)), SimulationOptions(pmlSize=2))
    try:
        with pytest.raises(NotImplementedError):
            sim.set_plane(True)
    finally:
        sim.close()

