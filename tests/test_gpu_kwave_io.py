"""GPU: k-Wave input files run end to end on the device (kwh_h5_read_input -> kwh_sim_run ->
kwh_h5_write_output, and the kwh_cli program), bitwise equal to kspaceFirstOrder on the inputs the
file holds, and within float32 rounding of the float64 oracle. The 2D case is the reference's
C++ parity scenario (CppParityTests.swift:32-66: 64^2, dx 1e-4, c 1500, rho 1000, PML 10, a disc
of radius 5, every grid point a sensor)."""
import ctypes as C
import json
import math
import os
import subprocess

import numpy as np
import pytest

from oracle import kwave_oracle as O
from kwavehip import (DTypePrecision, KWaveGrid, KWaveMedium, KWaveSensor, KWaveSource, RecordField,
                      SimulationOptions, kspaceFirstOrder, makeDisc, random_field)
from kwavehip import _lib as L
from kwavehip import kwave_io as IO

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not IO.available(), reason="libhdf5 not loadable")]

CLI = os.path.join(os.path.dirname(L.LIB_PATH), "kwh_cli")


@pytest.fixture(scope="module", autouse=True)
def _device():
    lib = L.load()
    h = C.c_void_p()
    if lib.kwh_ctx_create(0, C.byref(h)) != 0:
        pytest.skip("no HIP device")
    lib.kwh_ctx_destroy(h)


def _cpp_parity_case():
    n, dx, c = 64, 1e-4, 1500.0
    g = KWaveGrid(n, dx, n, dx)
    g.makeTime(c, cfl=0.3)
    g.setTime(80, g.dt)
    return g, KWaveMedium(c, 1000.0), KWaveSource(p0=makeDisc(n, n, radius=5)), \
        KWaveSensor(mask=np.ones((n, n), np.float32)), 10


def _hetero_nonlinear_3d():
    n = (40, 36, 32)
    g = KWaveGrid(n[0], 1e-4, n[1], 1e-4, n[2], 1e-4)
    c = random_field(n, 21, 4.0, 1450.0, 1600.0)
    g.setTime(60, 0.3 * 1e-4 / float(c.max()))
    med = KWaveMedium(c, random_field(n, 22, 4.0, 950.0, 1050.0), bOnA=np.float32(6.0))
    p0 = 2e5 * random_field(n, 23, 2.0, 0.0, 1.0)
    mask = np.zeros(n, np.float32)
    mask[30, 4:32:3, 2:30:5] = 1
    return g, med, KWaveSource(p0=p0.astype(np.float32)), KWaveSensor(mask=mask), 6


def _column_major_rows(mask):
    """Rows of the API's C-order sensor series in the file's column-major (x fastest) order."""
    pts = np.argwhere(np.asarray(mask) > 0)
    shape = np.asarray(mask).shape
    colmaj = [int(np.ravel_multi_index(tuple(p), shape, order="F")) for p in pts]
    return np.argsort(colmaj, kind="stable")


@pytest.mark.parametrize("case,dtype", [("cpp_parity_2d", DTypePrecision.float32),
                                        ("hetero_nonlinear_3d", DTypePrecision.float32),
                                        ("hetero_nonlinear_3d", DTypePrecision.float64)])
def test_input_file_runs_like_the_api(tmp_path, case, dtype):
    g, med, src, sens, pml = _cpp_parity_case() if case == "cpp_parity_2d" else _hetero_nonlinear_3d()
    print(f"\n[case] k-Wave file {case} {dtype.value} grid={g.size} nt={g.nt}", flush=True)
    inp, out = str(tmp_path / "in.h5"), str(tmp_path / "out.h5")
    IO.writeInput(inp, g, med, src, sens, pmlSize=pml, pmlAlpha=2.0)
    rec = RecordField.p | RecordField.pMax | RecordField.pRms | RecordField.pFinal
    IO.runFile(inp, out, record=rec, dtype=dtype)
    # the same inputs through the API (the file stores dx, dt as float32: read them back)
    g2, med2, src2, sens2, opts2 = IO.readInput(inp)
    sens2.record = rec
    api = kspaceFirstOrder(g2, med2, src2, sens2, SimulationOptions(pmlSize=opts2.pmlSize,
                                                                     pmlAlpha=opts2.pmlAlpha, smoothP0=False,
                                                                     dtype=dtype))
    pf = IO.readPFinal(out, *g.size) if g.dim == 3 else IO.readPFinal2D(out, *g.size)
    assert np.array_equal(pf, api.pFinal.astype(np.float64))
    order = _column_major_rows(sens.mask)
    p = IO.readFloatDataset(out, "p")
    assert p.shape == (1, g.nt, order.size)
    assert np.array_equal(p[0].T, api.p[order].astype(np.float64))
    for name, field in (("p_max", api.pMax), ("p_rms", api.pRms)):
        assert np.array_equal(IO.readFloatDataset(out, name)[0, 0], field[order].astype(np.float64)), name
    assert IO.readRootStringAttribute(out, "file_type") == "output"
    # and against the float64 oracle on the same inputs
    ref = O.kspace_first_order(g2, med2, src2, sens2, SimulationOptions(pmlSize=opts2.pmlSize,
                                                                        pmlAlpha=opts2.pmlAlpha, smoothP0=False),
                               workers=8)
    d = pf - ref["pFinal"]
    rl2 = math.sqrt(np.mean(d * d)) / math.sqrt(np.mean(ref["pFinal"] ** 2))
    print(f"  pFinal relL2 vs float64 oracle = {rl2:.3e}", flush=True)
    assert rl2 <= (1e-4 if dtype == DTypePrecision.float32 else 1e-9)


def test_cli_run_info_validate(tmp_path):
    g, med, src, sens, pml = _cpp_parity_case()
    inp, out = str(tmp_path / "in.h5"), str(tmp_path / "out.h5")
    IO.writeInput2D(inp, g, med, src, sens, pml, 2.0)
    print("\n[case] kwh_cli info / validate / run", flush=True)
    info = json.loads(subprocess.run([CLI, "info", "--input", inp], capture_output=True, text=True, check=True,
                                     timeout=60).stdout)
    assert info["n"][:2] == [64, 64] and info["nt"] == 80 and info["n_sensor"] == 64 * 64
    assert subprocess.run([CLI, "validate", "--input", inp], capture_output=True, timeout=60).returncode == 0
    res = subprocess.run([CLI, "run", "--input", inp, "--output", out, "--p_max"], capture_output=True, text=True,
                         timeout=120)
    assert res.returncode == 0, res.stderr
    assert json.loads(res.stdout)["steps"] == 80
    g2, med2, src2, sens2, opts2 = IO.readInput(inp)
    sens2.record = RecordField.pMax | RecordField.pFinal
    api = kspaceFirstOrder(g2, med2, src2, sens2, SimulationOptions(pmlSize=pml, smoothP0=False))
    assert np.array_equal(IO.readPFinal2D(out, 64, 64), api.pFinal.astype(np.float64))
    order = _column_major_rows(sens.mask)
    assert np.array_equal(IO.readFloatDataset(out, "p_max")[0, 0], api.pMax[order].astype(np.float64))
    # a missing input file fails with a message, not a crash
    bad = subprocess.run([CLI, "run", "--input", str(tmp_path / "none.h5"), "--output", out], capture_output=True,
                         text=True, timeout=60)
    assert bad.returncode == 1 and "could not open" in bad.stderr
