"""k-Wave HDF5 input / output through the C ABI (kwh_h5_*), mirroring the reference's `KWaveCpp`
and `HDF5File` (Sources/KWave/IO/KWaveInput.swift:12-111, Sources/KWave/IO/HDF5.swift:26-161).

    writeInput(path, grid, medium, source, sensor, pmlSize, pmlAlpha)   KWaveCpp.writeInput2D, 2D/3D
    readInput(path) -> (grid, medium, source, sensor, options)          the inverse, for the API
    readPFinal(path, nx, ny[, nz])                                       KWaveCpp.readPFinal2D
    runFile(input, output, record, dtype)                                `kwave-cli run --input --output`
    readFloatDataset(path, name) / readRootStringAttribute(path, name)   HDF5File readers

The library loads libhdf5 at run time; every call raises NotImplementedError when it cannot.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional

import numpy as np

from . import _lib as L
from .api import (DTypePrecision, KWaveGrid, KWaveMedium, KWaveSensor, KWaveSource, RecordField, SimulationOptions,
                  build_desc, _context)


def available() -> bool:
    """True when libhdf5 can be loaded."""
    return L.load().kwh_h5_available() == L.OK


def writeInput(path: str, grid: KWaveGrid, medium: KWaveMedium, source: KWaveSource, sensor: KWaveSensor,
               pmlSize: int = 20, pmlAlpha: float = 2.0) -> None:
    """k-Wave input file for a p0 problem with a binary sensor mask (KWaveInput.swift:12-99)."""
    d, keep, _ = build_desc(grid, medium, source, sensor, SimulationOptions(pmlSize=pmlSize, pmlAlpha=pmlAlpha,
                                                                            smoothP0=False))
    L.check(L.load().kwh_h5_write_input(str(path).encode(), C.byref(d)))
    del keep


def writeInput2D(path: str, grid: KWaveGrid, medium: KWaveMedium, source: KWaveSource, sensor: KWaveSensor,
                 pmlSize: int, pmlAlpha: float) -> None:
    """The reference's name for the 2D case (KWaveInput.swift:12-20)."""
    if grid.dim != 2:
        raise ValueError("writeInput2D requires a 2D grid")
    writeInput(path, grid, medium, source, sensor, pmlSize, pmlAlpha)


class InputFile:
    """A k-Wave input file read by the library (kwh_h5_read_input): `desc` is its kwh_sim_desc."""

    def __init__(self, path: str):
        self.lib = L.load()
        self.handle = C.c_void_p()
        L.check(self.lib.kwh_h5_read_input(str(path).encode(), C.byref(self.handle)))

    @property
    def desc(self) -> L.SimDesc:
        return self.lib.kwh_h5_input_desc(self.handle).contents

    def close(self):
        if getattr(self, "handle", None):
            self.lib.kwh_h5_input_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _arr(f: L.Field, shape):
    if not f.data:
        return None
    a = np.ctypeslib.as_array(f.data, shape=(f.count,)).copy()
    return np.float32(a[0]) if f.count == 1 else a.reshape(shape)


def readInput(path: str):
    """(grid, medium, source, sensor, options) of a k-Wave input file, as the API's values."""
    f = InputFile(path)
    try:
        d = f.desc
        shape = tuple(d.n[a] for a in range(d.dim))
        grid = KWaveGrid(*sum(([d.n[a], d.d[a]] for a in range(d.dim)), []))
        grid.setTime(d.nt, d.dt)
        medium = KWaveMedium(_arr(d.sound_speed, shape), _arr(d.density, shape), bOnA=_arr(d.bona, shape))
        source = KWaveSource(p0=_arr(d.p0, shape))
        mask = np.zeros(int(np.prod(shape)), np.float32)
        if d.n_sensor:
            mask[np.ctypeslib.as_array(d.sensor_index, shape=(d.n_sensor,))] = 1
        sensor = KWaveSensor(mask=mask.reshape(shape), record=RecordField.pFinal)
        options = SimulationOptions(pmlSize=d.pml_size, pmlAlpha=d.pml_alpha, smoothP0=False)
        return grid, medium, source, sensor, options
    finally:
        f.close()


def runFile(input_path: str, output_path: str, record: RecordField = RecordField.pFinal,
            dtype: DTypePrecision = DTypePrecision.float32, device: int = 0) -> None:
    """Run a k-Wave input file on the device and write the k-Wave output file (p_final, and the
    recorded p / pMax / pMin / pRms in the input's sensor order)."""
    f = InputFile(input_path)
    lib = f.lib
    try:
        d = L.SimDesc.from_buffer_copy(f.desc)
        d.record = int(record) | int(RecordField.pFinal)
        d.dtype = 1 if dtype == DTypePrecision.float64 else 0
        sim = C.c_void_p()
        L.check(lib.kwh_sim_create(_context(device), C.byref(d), C.byref(sim)))
        try:
            L.check(lib.kwh_sim_run(sim, 0, d.nt, L.STEP_CB(), 0, None))
            L.check(lib.kwh_h5_write_output(str(output_path).encode(), sim, f.handle))
        finally:
            lib.kwh_sim_destroy(sim)
    finally:
        f.close()


def readFloatDataset(path: str, name: str) -> np.ndarray:
    """A dataset as float64 in its stored shape (HDF5File.readFloatDataset, HDF5.swift:50-71)."""
    lib = L.load()
    dims = (C.c_int64 * 8)()
    nd = C.c_int32()
    L.check(lib.kwh_h5_read_dataset(str(path).encode(), name.encode(), dims, C.byref(nd), None, 0))
    shape = tuple(dims[i] for i in range(nd.value))
    out = np.empty(shape, np.float64)
    L.check(lib.kwh_h5_read_dataset(str(path).encode(), name.encode(), dims, C.byref(nd),
                                    out.ctypes.data_as(C.POINTER(C.c_double)), out.size))
    return out


def readRootStringAttribute(path: str, name: str) -> str:
    """HDF5File.readRootStringAttribute (HDF5.swift:149-161)."""
    buf = C.create_string_buffer(4096)
    L.check(L.load().kwh_h5_read_root_attribute(str(path).encode(), name.encode(), buf, len(buf)))
    return buf.value.decode()


def readPFinal(path: str, nx: int, ny: int, nz: Optional[int] = None) -> np.ndarray:
    """p_final of a k-Wave output file as the API's C-order [nx, ny(, nz)] array: the on-disk
    [Nz, Ny, Nx] transposed back (readPFinal2D, KWaveInput.swift:101-111)."""
    disk = readFloatDataset(path, "p_final")
    want = (nz or 1, ny, nx)
    if disk.shape[-2:] != (ny, nx) or disk.size != nx * ny * (nz or 1):
        raise ValueError(f"unexpected p_final shape {disk.shape}")
    t = disk.reshape(want).transpose(2, 1, 0)
    return t[:, :, 0] if nz is None else t


def readPFinal2D(path: str, nx: int, ny: int) -> np.ndarray:
    return readPFinal(path, nx, ny)
