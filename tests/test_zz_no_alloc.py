"""GPU: kwh_sim_read_output_noalloc (outputs staged in the library's static device scratch) against
kwh_sim_read_output (staging allocated per call), through the C ABI, bit for bit, on calls in a row
whose second and third inputs are smaller than the first: the first read's series span several
scratch chunks (65 536 sensors x 100 steps of float32 = 25 MiB through 8 MiB of scratch)."""
import ctypes as C

import numpy as np
import pytest

from kwavehip import (DTypePrecision, KWaveGrid, KWaveMedium, KWaveSensor, KWaveSource, RecordField, Simulation,
                      SimulationOptions, makeDisc)
from kwavehip import _lib as L

pytestmark = pytest.mark.gpu

FIELDS = {
    "p": L.OUT_P, "pMax": L.OUT_P_MAX, "pMin": L.OUT_P_MIN, "pRms": L.OUT_P_RMS, "pFinal": L.OUT_P_FINAL,
    "ux": L.OUT_UX, "uy": L.OUT_UY, "uxMax": L.OUT_UX_MAX, "uyRms": L.OUT_UY_RMS, "uxFinal": L.OUT_UX_FINAL,
    "ixAvg": L.OUT_IX_AVG, "iyAvg": L.OUT_IY_AVG,
}
RECORD = (RecordField.p | RecordField.pMax | RecordField.pMin | RecordField.pRms | RecordField.pFinal |
          RecordField.ux | RecordField.uy | RecordField.uMax | RecordField.uRms | RecordField.uFinal |
          RecordField.iAvg)


@pytest.fixture(scope="module", autouse=True)
def _device():
    lib = L.load()
    h = C.c_void_p()
    if lib.kwh_ctx_create(0, C.byref(h)) != 0:
        pytest.skip("no HIP device")
    lib.kwh_ctx_destroy(h)


def _sim(n, steps, mask, dtype):
    grid = KWaveGrid(n, 1e-4, n, 1e-4)
    grid.setTime(steps, 2e-8)
    sim = Simulation(grid, KWaveMedium(1500.0, 1000.0), KWaveSource(p0=makeDisc(n, n, radius=4)),
                     KWaveSensor(mask=mask, record=RECORD), SimulationOptions(dtype=dtype, pmlSize=8))
    sim.run()
    return sim


def _read(sim, field, entry):
    shape = (C.c_int64 * 3)()
    nd = C.c_int32()
    L.check(sim.lib.kwh_sim_output_shape(sim.handle, field, shape, C.byref(nd)))
    out = np.full(tuple(shape[i] for i in range(nd.value)), np.nan, dtype=sim.dtype)
    L.check(entry(sim.handle, field, out.ctypes.data_as(C.c_void_p), out.nbytes))
    return out


def _compare(sim, label):
    lib = sim.lib
    for name, f in FIELDS.items():
        a = _read(sim, f, lib.kwh_sim_read_output)
        b = _read(sim, f, lib.kwh_sim_read_output_noalloc)
        assert a.shape == b.shape and a.size > 0, (label, name)
        assert np.all(np.isfinite(a)), (label, name)
        assert a.tobytes() == b.tobytes(), (label, name)


def test_noalloc_read_equals_allocating_read_on_successive_calls():
    n1 = 256
    print(f"\n[case] no-alloc read: {n1}^2 full-grid sensor, 100 steps, float32", flush=True)
    big = _sim(n1, 100, np.ones((n1, n1), np.float32), DTypePrecision.float32)
    assert big.n_sensor * 100 * 4 > 3 * (8 << 20)          # several scratch chunks per series
    _compare(big, "first")
    big.close()
    n2 = 48
    line = np.zeros((n2, n2), np.float32)
    line[6, 3:40] = 1
    print(f"[case] no-alloc read: {n2}^2 line sensor, 60 steps, float32 then float64", flush=True)
    small = _sim(n2, 60, line, DTypePrecision.float32)
    _compare(small, "second")
    small64 = _sim(n2, 40, line, DTypePrecision.float64)
    _compare(small64, "third")
    # wrong arguments fail like the allocating entry point
    buf = np.zeros(4, np.float32)
    assert small.lib.kwh_sim_read_output_noalloc(small.handle, L.OUT_P, buf.ctypes.data_as(C.c_void_p),
                                                 buf.nbytes) == L.EINVAL
    assert small.lib.kwh_sim_read_output_noalloc(small.handle, 99, buf.ctypes.data_as(C.c_void_p),
                                                 buf.nbytes) == L.EINVAL
