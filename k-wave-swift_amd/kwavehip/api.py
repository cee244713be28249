"""Host-side mirror of KWaveSwift's solver API for the `.hip` device.

Same names, argument meaning and error behaviour as the reference's Swift types:
  KWaveGrid          Sources/KWave/Grid/KWaveGrid.swift:9-143
  KWaveMedium        Sources/KWave/Medium/KWaveMedium.swift:21-76 (AbsorptionMode :8-17)
  KWaveSource        Sources/KWave/Source/KWaveSource.swift:11-71 (SourceMode :4-8)
  KWaveSensor        Sources/KWave/Sensor/KWaveSensor.swift:32-62 (RecordField :4-21)
  SimulationOptions  Sources/KWave/Solver/KSpaceFirstOrder.swift:17-40
  SimulationOutput   Sources/KWave/Solver/KSpaceFirstOrder.swift:47-73
  kspaceFirstOrder   Sources/KWave/Solver/KSpaceFirstOrder.swift:79-107
Arrays are numpy (C-order over [nx, ny, nz]) where the reference uses MLXArray. The compute runs
in libkwavehip (HIP, gfx950) through the C ABI; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import enum
import math
from dataclasses import dataclass, field
from typing import Callable, Optional

import numpy as np

from . import _lib as L


class DTypePrecision(str, enum.Enum):
    """KWave.swift:4-9."""
    float32 = "float32"
    float64 = "float64"


class DeviceKind(str, enum.Enum):
    """KWave.swift:12-22 plus `.hip`, the device this library implements."""
    gpu = "gpu"
    cpu = "cpu"
    hip = "hip"


class AbsorptionMode(str, enum.Enum):
    noAbsorption = "noAbsorption"
    powerLaw = "powerLaw"
    noDispersion = "noDispersion"
    stokes = "stokes"


class SourceMode(str, enum.Enum):
    dirichlet = "dirichlet"
    additive = "additive"
    additiveNoCorrection = "additiveNoCorrection"


class DirectivityPattern(str, enum.Enum):
    pressure = "pressure"
    gradient = "gradient"


class RecordField(enum.IntFlag):
    """KWaveSensor.swift:8-20 (same bit values)."""
    p = 1 << 0
    pMax = 1 << 1
    pMin = 1 << 2
    pRms = 1 << 3
    pFinal = 1 << 4
    ux = 1 << 5
    uy = 1 << 6
    uz = 1 << 7
    uMax = 1 << 8
    uRms = 1 << 9
    uFinal = 1 << 10
    iAvg = 1 << 11
    iMax = 1 << 12


_ALPHA = {AbsorptionMode.noAbsorption: 0, AbsorptionMode.powerLaw: 1, AbsorptionMode.noDispersion: 2,
          AbsorptionMode.stokes: 3}
_SRC = {SourceMode.dirichlet: 0, SourceMode.additive: 1, SourceMode.additiveNoCorrection: 2}


def _fft_freq(n: int, d: float) -> np.ndarray:
    """numpy-style fftfreq (Internal/Support.swift:6-14)."""
    val = 1.0 / (n * d)
    out = np.zeros(n)
    cutoff = (n - 1) // 2 + 1
    out[:cutoff] = np.arange(cutoff) * val
    out[cutoff:] = np.arange(-(n // 2), 0) * val
    return out


class KWaveGrid:
    """Computational grid (KWaveGrid.swift:9-143). `KWaveGrid(nx, dx)`, `(nx, dx, ny, dy)`,
    `(nx, dx, ny, dy, nz, dz)`."""

    def __init__(self, nx: int, dx: float, ny: Optional[int] = None, dy: Optional[float] = None,
                 nz: Optional[int] = None, dz: Optional[float] = None):
        if ny is None:
            self.dim = 1
        elif nz is None:
            self.dim = 2
        else:
            self.dim = 3
        self.nx, self.ny, self.nz = int(nx), int(ny or 1), int(nz or 1)
        self.dx, self.dy, self.dz = float(dx), float(dy or 1.0), float(dz or 1.0)
        self.kxVec = 2 * math.pi * _fft_freq(self.nx, self.dx)
        self.kyVec = 2 * math.pi * _fft_freq(self.ny, self.dy)
        self.kzVec = 2 * math.pi * _fft_freq(self.nz, self.dz)
        self.dt = 0.0
        self.nt = 0
        self.tArray = np.zeros(0)

    @property
    def totalGridPoints(self) -> int:
        return self.nx * self.ny * self.nz

    @property
    def spacing(self):
        return [self.dx, self.dy, self.dz][: self.dim]

    @property
    def size(self):
        return [self.nx, self.ny, self.nz][: self.dim]

    @property
    def k(self) -> np.ndarray:
        """|k| in FFT order (KWaveGrid.swift:79-100)."""
        vecs = [self.kxVec, self.kyVec, self.kzVec][: self.dim]
        k2 = np.zeros(tuple(self.size))
        for a, v in enumerate(vecs):
            shape = [1] * self.dim
            shape[a] = v.size
            k2 = k2 + v.reshape(shape) ** 2
        return np.sqrt(k2)

    def makeTime(self, soundSpeedMax: float, soundSpeedMin: Optional[float] = None, cfl: float = 0.3,
                 tEnd: Optional[float] = None) -> None:
        """dt = cfl*min(d)/cmax; nt = floor(tEnd/dt)+1 (KWaveGrid.swift:112-135)."""
        cmin = soundSpeedMin if soundSpeedMin is not None else soundSpeedMax
        step = cfl * min(self.spacing) / soundSpeedMax
        if tEnd is None:
            tEnd = math.sqrt(sum((n * d) ** 2 for n, d in zip(self.size, self.spacing))) / cmin
        count = int(math.floor(tEnd / step)) + 1
        self.dt, self.nt = step, count
        self.tArray = np.arange(count) * step

    def setTime(self, nt: int, dt: float) -> None:
        """KWaveGrid.swift:138-142."""
        self.dt, self.nt = float(dt), int(nt)
        self.tArray = np.arange(self.nt) * self.dt


@dataclass
class KWaveMedium:
    """Acoustic medium (KWaveMedium.swift:21-76). Scalars or grid-shaped maps."""
    soundSpeed: object
    density: object
    alphaCoeff: object = None
    alphaPower: Optional[float] = None
    alphaMode: AbsorptionMode = AbsorptionMode.noAbsorption
    bOnA: object = None

    @property
    def isHomogeneous(self) -> bool:
        return np.size(self.soundSpeed) == 1 and np.size(self.density) == 1


@dataclass
class KWaveSource:
    """Sources (KWaveSource.swift:11-71); stress fields are elastic-only and not part of this path."""
    p0: object = None
    pMask: object = None
    p: object = None
    pMode: SourceMode = SourceMode.additive
    uMask: object = None
    ux: object = None
    uy: object = None
    uz: object = None
    uMode: SourceMode = SourceMode.additive
    # Delayed copies (a transducer drive; not in the reference's KWaveSource): with a delay per
    # mask point, p / ux / uy / uz is one row and point k plays scale[k] * row[t - delay[k]],
    # expanded on the device (KWaveTransducer.velocity_source() produces these).
    pDelay: object = None
    pScale: object = None
    uDelay: object = None
    uScale: object = None


@dataclass
class KWaveSensor:
    """Sensor (KWaveSensor.swift:32-62). `mask` is a binary grid mask or Cartesian [dim, n] points."""
    mask: object = None
    record: RecordField = RecordField.p
    timeReversalBoundaryData: object = None
    directivityAngle: object = None
    directivitySize: Optional[float] = None
    directivityPattern: DirectivityPattern = DirectivityPattern.pressure
    frequencyResponse: object = None


@dataclass
class SimulationOptions:
    """KSpaceFirstOrder.swift:17-40. `device` defaults to `.hip` (this library)."""
    pmlInside: bool = True
    pmlSize: int = 20
    pmlAlpha: float = 2.0
    smoothP0: bool = True
    smoothC0: bool = False
    smoothRho0: bool = False
    dtype: DTypePrecision = DTypePrecision.float32
    device: DeviceKind = DeviceKind.hip
    plotSim: bool = False
    recordMovie: Optional[str] = None
    saveToDisk: Optional[str] = None
    progress: Optional[Callable[[int, int], None]] = None
    fieldMonitor: Optional[Callable[[int, np.ndarray], None]] = None
    fieldMonitorInterval: int = 1
    # no counterpart in the reference: scan p for NaN / inf every this many steps (0 = off); the
    # run stops with kwavehip.KWaveNonFiniteError (KWH_ENONFINITE) at the first flagged check
    nonfiniteCheckEvery: int = 0


@dataclass
class SimulationOutput:
    """KSpaceFirstOrder.swift:47-73; series [nSensor, nt], aggregates [nSensor], finals grid-shaped."""
    p: Optional[np.ndarray] = None
    pMax: Optional[np.ndarray] = None
    pMin: Optional[np.ndarray] = None
    pRms: Optional[np.ndarray] = None
    pFinal: Optional[np.ndarray] = None
    ux: Optional[np.ndarray] = None
    uy: Optional[np.ndarray] = None
    uz: Optional[np.ndarray] = None
    uxMax: Optional[np.ndarray] = None
    uyMax: Optional[np.ndarray] = None
    uzMax: Optional[np.ndarray] = None
    uxRms: Optional[np.ndarray] = None
    uyRms: Optional[np.ndarray] = None
    uzRms: Optional[np.ndarray] = None
    uxFinal: Optional[np.ndarray] = None
    uyFinal: Optional[np.ndarray] = None
    uzFinal: Optional[np.ndarray] = None
    ixAvg: Optional[np.ndarray] = None
    iyAvg: Optional[np.ndarray] = None
    izAvg: Optional[np.ndarray] = None


def flatNonzeroIndices(mask) -> np.ndarray:
    """Ascending C-order flat indices of nonzero entries (KSpaceFirstOrder.swift:778-782)."""
    return np.flatnonzero(np.asarray(mask).reshape(-1) != 0).astype(np.int64)


_ctx = {}


def _context(device: int = 0):
    lib = L.load()
    if device not in _ctx:
        h = C.c_void_p()
        L.check(lib.kwh_ctx_create(device, C.byref(h)))
        _ctx[device] = h
    return _ctx[device]


def build_desc(grid: KWaveGrid, medium: KWaveMedium, source: KWaveSource, sensor: KWaveSensor,
               options: SimulationOptions):
    """Flatten the API values into the C descriptor (kwh_sim_desc). Returns (desc, keepalive, nSensor)."""
    keep = []  # host buffers the descriptor points into
    dtype64 = options.dtype == DTypePrecision.float64

    def fld(v):
        if v is None:
            return L.Field(None, 0)
        a = np.ascontiguousarray(np.asarray(v, dtype=np.float32).reshape(-1))
        keep.append(a)
        return L.Field(a.ctypes.data_as(C.POINTER(C.c_float)), a.size)

    def src_field(mask, signal, mode, delay=None, scale=None):
        if mask is None or signal is None:
            return L.SourceField(None, 0, None, 0, 0, 0, None, None)
        idx = flatNonzeroIndices(mask)
        sig = np.asarray(signal, dtype=np.float32)
        sig = sig.reshape(1, -1) if sig.ndim <= 1 else sig.reshape(sig.shape[0], -1)
        sig = np.ascontiguousarray(sig)
        keep.extend([idx, sig])
        if idx.size == 0:
            return L.SourceField(None, 0, None, 0, 0, 0, None, None)
        f = L.SourceField(idx.ctypes.data_as(C.POINTER(C.c_int64)), idx.size,
                          sig.ctypes.data_as(C.POINTER(C.c_float)), sig.shape[0], sig.shape[1],
                          _SRC[SourceMode(mode)], None, None)
        if delay is not None:
            dl = np.ascontiguousarray(np.asarray(delay, dtype=np.int32).reshape(-1))
            if dl.size != idx.size:
                raise ValueError("source delays must have one entry per mask point")
            keep.append(dl)
            f.delay = dl.ctypes.data_as(C.POINTER(C.c_int32))
            if scale is not None:
                sc = np.ascontiguousarray(np.asarray(scale, dtype=np.float32).reshape(-1))
                if sc.size != idx.size:
                    raise ValueError("source scales must have one entry per mask point")
                keep.append(sc)
                f.scale = sc.ctypes.data_as(C.POINTER(C.c_float))
        return f

    d = L.SimDesc()
    d.abi_version = L.ABI_VERSION
    d.dim = grid.dim
    for a in range(3):
        d.n[a] = [grid.nx, grid.ny, grid.nz][a] if a < grid.dim else 1
        d.d[a] = [grid.dx, grid.dy, grid.dz][a] if a < grid.dim else 1.0
    d.dt = grid.dt
    d.nt = grid.nt
    d.dtype = 1 if dtype64 else 0
    d.pml_size = int(options.pmlSize)
    d.pml_alpha = float(options.pmlAlpha)
    d.smooth_p0, d.smooth_c0, d.smooth_rho0 = int(options.smoothP0), int(options.smoothC0), int(options.smoothRho0)
    d.sound_speed = fld(medium.soundSpeed)
    d.density = fld(medium.density)
    d.alpha_coeff = fld(medium.alphaCoeff)
    d.alpha_power = float(medium.alphaPower) if medium.alphaPower is not None else 1.5
    d.alpha_mode = _ALPHA[AbsorptionMode(medium.alphaMode)]
    d.bona = fld(medium.bOnA)
    d.p0 = fld(source.p0)
    d.p_source = src_field(source.pMask, source.p, source.pMode, source.pDelay, source.pScale)
    for a, s in enumerate([source.ux, source.uy, source.uz]):
        d.u_source[a] = src_field(source.uMask, s, source.uMode, source.uDelay, source.uScale)
    n_sensor = 0
    if sensor is not None and sensor.mask is not None:
        m = np.asarray(sensor.mask)
        if m.ndim == 2 and m.shape[0] == grid.dim:          # Recording.swift:50-53
            pts = np.ascontiguousarray(m.astype(np.float32))
            keep.append(pts)
            d.sensor_cartesian = pts.ctypes.data_as(C.POINTER(C.c_float))
            d.n_cartesian = pts.shape[1]
            n_sensor = pts.shape[1]
        else:
            idx = flatNonzeroIndices(m)
            keep.append(idx)
            if idx.size:
                d.sensor_index = idx.ctypes.data_as(C.POINTER(C.c_int64))
                d.n_sensor = idx.size
            n_sensor = idx.size
        d.record = int(sensor.record)
        if sensor.directivityAngle is not None:
            ang = np.ascontiguousarray(np.asarray(sensor.directivityAngle, dtype=np.float32).reshape(-1))
            keep.append(ang)
            d.directivity_angle = ang.ctypes.data_as(C.POINTER(C.c_float))
            d.directivity_size = float(sensor.directivitySize) if sensor.directivitySize is not None else 0.0
            d.directivity_pattern = 1 if DirectivityPattern(sensor.directivityPattern) == DirectivityPattern.gradient else 0
    else:
        d.record = int(sensor.record) if sensor is not None else 0
    return d, keep, n_sensor


class Simulation:
    """One simulation on the device: the C ABI's kwh_sim handle with the inputs it borrows.

    Built from the reference's API values; `run()` advances the time loop, `output()` reads a
    recorded field. kspaceFirstOrder() is the one-call form.
    """

    def __init__(self, grid: KWaveGrid, medium: KWaveMedium, source: KWaveSource, sensor: KWaveSensor,
                 options: SimulationOptions = None, device: int = 0, axisymmetric: bool = False):
        options = options or SimulationOptions()
        if options.device != DeviceKind.hip:
            raise NotImplementedError("this library implements device .hip only (the MLX path is not part of it)")
        self.lib = L.load()
        self.grid = grid
        self.dtype = np.float64 if options.dtype == DTypePrecision.float64 else np.float32
        d, self._keep, self.n_sensor = build_desc(grid, medium, source, sensor, options)
        self.record = RecordField(d.record)
        self.handle = C.c_void_p()
        create = self.lib.kwh_sim_create_axisymmetric if axisymmetric else self.lib.kwh_sim_create
        L.check(create(_context(device), C.byref(d), C.byref(self.handle)))
        if getattr(options, "nonfiniteCheckEvery", 0):
            self.set_nonfinite_check(options.nonfiniteCheckEvery)

    @classmethod
    def _view(cls, lib, handle, grid, dtype, record, n_sensor):
        """A non-owning view of a sim that belongs to a SimulationBatch."""
        obj = cls.__new__(cls)
        obj.lib, obj.handle, obj.grid, obj.dtype = lib, handle, grid, dtype
        obj.record, obj.n_sensor, obj._keep, obj._owned = record, n_sensor, [], False
        return obj

    def close(self):
        if getattr(self, "handle", None):
            if getattr(self, "_owned", True):
                self.lib.kwh_sim_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def steps_done(self) -> int:
        s = C.c_int32()
        L.check(self.lib.kwh_sim_steps_done(self.handle, C.byref(s)))
        return s.value

    def run(self, t_begin: int = None, t_end: int = None, callback=None, interval: int = 1):
        """Advance steps [t_begin, t_end); callback(t, nt, p_or_None) once per step, with the pressure
        after step t when t % interval == 0 (interval <= 0: never)."""
        t0 = self.steps_done if t_begin is None else t_begin
        t1 = self.grid.nt if t_end is None else t_end
        if callback is None:
            L.check(self.lib.kwh_sim_run(self.handle, t0, t1, L.STEP_CB(), 0, None))
            return
        shape = tuple(self.grid.size)
        n = int(np.prod(shape))
        ctype = C.c_double if self.dtype == np.float64 else C.c_float

        def _cb(t, nt, ptr, user):
            arr = None
            if ptr:
                arr = np.ctypeslib.as_array(C.cast(ptr, C.POINTER(ctype)), shape=(n,)).copy().reshape(shape)
            callback(t, nt, arr)

        cb = L.STEP_CB(_cb)
        L.check(self.lib.kwh_sim_run(self.handle, t0, t1, cb, int(interval), None))

    def launch(self, t_begin: int, t_end: int):
        L.check(self.lib.kwh_sim_launch(self.handle, t_begin, t_end))

    def sync(self):
        L.check(self.lib.kwh_sim_sync(self.handle))

    def output(self, which: int) -> np.ndarray:
        shape = (C.c_int64 * 3)()
        nd = C.c_int32()
        L.check(self.lib.kwh_sim_output_shape(self.handle, which, shape, C.byref(nd)))
        shp = tuple(shape[i] for i in range(nd.value))
        out = np.empty(shp, dtype=self.dtype)
        L.check(self.lib.kwh_sim_read_output(self.handle, which, out.ctypes.data_as(C.c_void_p), out.nbytes))
        return out

    def has_output(self, which: int) -> bool:
        shape = (C.c_int64 * 3)()
        nd = C.c_int32()
        return self.lib.kwh_sim_output_shape(self.handle, which, shape, C.byref(nd)) == 0

    def state(self) -> bytes:
        n = C.c_size_t()
        L.check(self.lib.kwh_sim_state_bytes(self.handle, C.byref(n)))
        buf = (C.c_char * n.value)()
        L.check(self.lib.kwh_sim_save_state(self.handle, buf, n.value))
        return bytes(buf)

    def load_state(self, blob: bytes):
        buf = C.create_string_buffer(blob, len(blob))
        L.check(self.lib.kwh_sim_load_state(self.handle, buf, len(blob)))

    def set_kernel_timing(self, on: bool, only: Optional[str] = None):
        """HIP-event timing of every kernel launch, or of the launches of kernel `only`."""
        if on and only:
            L.check(self.lib.kwh_sim_set_kernel_timing_only(self.handle, only.encode()))
        else:
            L.check(self.lib.kwh_sim_set_kernel_timing(self.handle, 1 if on else 0))

    def kernel_stats(self):
        cnt = C.c_int32()
        L.check(self.lib.kwh_sim_kernel_stats(self.handle, None, 0, C.byref(cnt)))
        arr = (L.KernelStat * max(cnt.value, 1))()
        L.check(self.lib.kwh_sim_kernel_stats(self.handle, arr, cnt.value, C.byref(cnt)))
        return [{"name": arr[i].name.decode(), "launches": arr[i].launches, "total_ms": arr[i].total_ms,
                 "bytes_per_launch": arr[i].bytes_per_launch} for i in range(cnt.value)]

    def set_nonfinite_check(self, every: int):
        """Scan p for NaN / inf after every `every`-th step of later runs (0 = off); a run that finds
        one stops and raises KWaveNonFiniteError (KWH_ENONFINITE)."""
        L.check(self.lib.kwh_sim_set_nonfinite_check(self.handle, int(every)))

    def set_graph(self, on: bool):
        """Replay the time step from captured hipGraphs (bitwise equal to eager launches)."""
        L.check(self.lib.kwh_sim_set_graph(self.handle, 1 if on else 0))

    def set_plane(self, on: bool):
        """Plane mode (kwh_sim_set_plane): the y and z transforms of each field in one kernel per
        x plane; 3-D grids with a 32x32, 64x64, 96x96 or 128x128 (y, z) plane, any medium. On by
        default wherever it applies; set_plane(False) selects the split passes."""
        L.check(self.lib.kwh_sim_set_plane(self.handle, 1 if on else 0))

    def element_traces(self, field: int, element, delay, n_elements: int, scale: float) -> np.ndarray:
        """Receive combination on the device (combineSensorData, KWaveTransducer.swift:263-289):
        out[e, t] = scale * sum over sensor points k of element e of series[k, t + delay[k]]."""
        el = np.ascontiguousarray(np.asarray(element, np.int32).reshape(-1))
        dl = np.ascontiguousarray(np.asarray(delay, np.int32).reshape(-1))
        if el.size != self.n_sensor or dl.size != self.n_sensor:
            raise ValueError("element and delay need one entry per sensor point")
        shape = (C.c_int64 * 3)()
        nd = C.c_int32()
        L.check(self.lib.kwh_sim_output_shape(self.handle, field, shape, C.byref(nd)))
        out = np.empty((int(n_elements), int(shape[1])), dtype=self.dtype)
        L.check(self.lib.kwh_sim_element_traces(self.handle, field, el.ctypes.data_as(C.POINTER(C.c_int32)),
                                                dl.ctypes.data_as(C.POINTER(C.c_int32)), int(n_elements),
                                                float(scale), out.ctypes.data_as(C.c_void_p), out.nbytes))
        return out

    def autotune(self) -> str:
        """Pick the fastest launch schedule for this grid (before the first step); returns the
        candidates' measured step times."""
        buf = C.create_string_buffer(4096)
        L.check(self.lib.kwh_sim_autotune(self.handle, buf, len(buf)))
        return buf.value.decode()

    def set_checkpointing(self, every: int):
        """Save the state every `every` steps during later runs, for vjp() (0 = off)."""
        L.check(self.lib.kwh_sim_set_checkpointing(self.handle, int(every)))

    def vjp(self, dL_dp: np.ndarray, checkpoint_every: int = 32):
        """Gradient of a loss of the recorded pressure series y [nSensor, nt]: given dL/dy,
        returns (dL/dp0, dL/dc) on the grid (lossless linear p0 problem; c_ref held fixed).
        Runs the forward pass first unless the sim has just run all nt steps with checkpoints
        every `checkpoint_every` steps; the reverse sweep recomputes between checkpoints."""
        g = np.ascontiguousarray(np.asarray(dL_dp, dtype=self.dtype))
        if g.shape != (self.n_sensor, self.grid.nt):
            raise ValueError(f"dL/dp must be [nSensor, nt] = {(self.n_sensor, self.grid.nt)}")
        shape = tuple(self.grid.size)
        gp0 = np.empty(shape, dtype=self.dtype)
        gc = np.empty(shape, dtype=self.dtype)
        L.check(self.lib.kwh_sim_vjp(self.handle, g.ctypes.data_as(C.c_void_p), int(checkpoint_every),
                                     gp0.ctypes.data_as(C.c_void_p), gc.ctypes.data_as(C.c_void_p)))
        return gp0, gc

    def transducer_traces(self, transducer, field: int = L.OUT_P) -> np.ndarray:
        """Per-element receive traces of a sensor placed on the transducer's active voxels."""
        element, delay = transducer.receive_combination()
        g = transducer.geometry
        return self.element_traces(field, element, delay, transducer.numberActiveElements,
                                   1.0 / (g.elementWidth * g.elementLength))

    def outputs(self) -> SimulationOutput:
        """finalizeRecording (Recording.swift:129-168): only the recorded fields are set."""
        out = SimulationOutput()
        out.pFinal = self.output(L.OUT_P_FINAL)
        names = {L.OUT_P: "p", L.OUT_P_MAX: "pMax", L.OUT_P_MIN: "pMin", L.OUT_P_RMS: "pRms",
                 L.OUT_UX: "ux", L.OUT_UY: "uy", L.OUT_UZ: "uz",
                 L.OUT_UX_MAX: "uxMax", L.OUT_UY_MAX: "uyMax", L.OUT_UZ_MAX: "uzMax",
                 L.OUT_UX_RMS: "uxRms", L.OUT_UY_RMS: "uyRms", L.OUT_UZ_RMS: "uzRms",
                 L.OUT_UX_FINAL: "uxFinal", L.OUT_UY_FINAL: "uyFinal", L.OUT_UZ_FINAL: "uzFinal",
                 L.OUT_IX_AVG: "ixAvg", L.OUT_IY_AVG: "iyAvg", L.OUT_IZ_AVG: "izAvg"}
        for k, nm in names.items():
            if self.has_output(k):
                setattr(out, nm, self.output(k))
        return out


def timeReversal(grid: KWaveGrid, medium: KWaveMedium, sensor: KWaveSensor, recordedPressure,
                 compensationFactor: float = 2.0, options: SimulationOptions = None, device: int = 0) -> np.ndarray:
    """Time-reversal reconstruction of p0 (TimeReversal.swift:20-51): the recording re-injected,
    reversed in time, as a Dirichlet pressure source on sensor.mask; returns
    max(0, compensationFactor * pFinal). Flip and positivity run on the device."""
    options = options or SimulationOptions()
    if sensor is None or sensor.mask is None:
        raise ValueError("time reversal requires sensor.mask")
    if grid.nt <= 0:
        raise ValueError("set grid time (makeTime/setTime) before time reversal")
    rec = np.ascontiguousarray(np.asarray(recordedPressure, dtype=np.float32))
    if rec.ndim != 2:
        raise ValueError("recordedPressure must be [numSensorPoints, nt]")
    d, keep, n_sensor = build_desc(grid, medium, KWaveSource(), KWaveSensor(mask=sensor.mask), options)
    if rec.shape[0] != n_sensor:
        raise ValueError("recordedPressure rows must match the sensor mask points")
    lib = L.load()
    out = np.empty(tuple(grid.size), dtype=np.float64 if options.dtype == DTypePrecision.float64 else np.float32)
    L.check(lib.kwh_time_reversal(_context(device), C.byref(d), rec.ctypes.data_as(C.POINTER(C.c_float)),
                                  rec.shape[1], float(compensationFactor), out.ctypes.data_as(C.c_void_p)))
    return out


class SimulationBatch:
    """Independent simulations of uneven sizes advanced together on one device (kwh_batch_*):
    per-scan-line transmits, for instance. `sims[i]` reads outputs like a Simulation.

    run(fused=True): one launch per pass of the time step covers all sims of one grid shape, each
    shape on a stream of its own (overlapping the others); with use_graph each shape's batched
    step is captured once and replayed.
    run(fused=False): each sim on its own stream, replaying its own step graphs (use_graph) or
    enqueuing its steps eagerly.
    run(plane=True): switch every sim that plane mode covers into it (Simulation.set_plane; the
    default where it applies); with plane=False every sim keeps its own mode.
    All modes give results bitwise equal to single runs in the same plane mode."""

    def __init__(self, cases, device: int = 0):
        self.lib = L.load()
        n = len(cases)
        if n == 0:
            raise ValueError("empty batch")
        descs = (L.SimDesc * n)()
        self._keep = []
        meta = []
        for i, (grid, medium, source, sensor, options) in enumerate(cases):
            options = options or SimulationOptions()
            d, keep, n_sensor = build_desc(grid, medium, source, sensor, options)
            descs[i] = d
            self._keep.append(keep)
            meta.append((grid, np.float64 if options.dtype == DTypePrecision.float64 else np.float32,
                         RecordField(d.record), n_sensor))
        self.handle = C.c_void_p()
        L.check(self.lib.kwh_batch_create(_context(device), descs, n, C.byref(self.handle)))
        self.sims = []
        for i, (grid, dt, rec, ns) in enumerate(meta):
            h = C.c_void_p()
            L.check(self.lib.kwh_batch_sim(self.handle, i, C.byref(h)))
            self.sims.append(Simulation._view(self.lib, h, grid, dt, rec, ns))

    def run(self, use_graph: bool = True, fused: bool = True, plane: bool = False):
        mode = (L.BATCH_GRAPH if use_graph else 0) | (L.BATCH_FUSED if fused else 0) | (L.BATCH_PLANE if plane else 0)
        L.check(self.lib.kwh_batch_run(self.handle, mode))

    def close(self):
        if getattr(self, "handle", None):
            for s in self.sims:
                s.handle = None
            self.lib.kwh_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def kspaceFirstOrder(grid: KWaveGrid, medium: KWaveMedium, source: KWaveSource, sensor: KWaveSensor,
                     options: SimulationOptions = None) -> SimulationOutput:
    """k-space pseudospectral first-order solver on the HIP device (KSpaceFirstOrder.swift:79-107).

    progress(t, nt) is called once per step and fieldMonitor(t, p) every fieldMonitorInterval
    steps, as in the reference (:337-340); both are delivered between device step chunks.
    """
    options = options or SimulationOptions()
    if grid.dim not in (1, 2, 3):
        raise NotImplementedError(f"kspaceFirstOrder: dim {grid.dim} not yet implemented")
    if grid.nt <= 0:
        raise ValueError("call grid.makeTime(...) before running the solver")
    if source.p0 is None and source.pMask is None and source.uMask is None:
        raise ValueError(f"{grid.dim}D solver requires source.p0 or a time-varying p/u source")
    sim = Simulation(grid, medium, source, sensor if sensor is not None else KWaveSensor(), options)
    return _run_to_output(sim, options)


def kspaceFirstOrderAS(grid: KWaveGrid, medium: KWaveMedium, source: KWaveSource, sensor: KWaveSensor,
                       options: SimulationOptions = None) -> SimulationOutput:
    """Axisymmetric k-space solver on the HIP device (KSpaceFirstOrderAS.swift:17-309): a 2D grid
    with x axial and y radial (y index 0 on the axis), WSWA-FFT radial symmetry, Stokes absorption,
    B/A, sources p0 / p / ux / uy; records the pressure (p, pFinal)."""
    options = options or SimulationOptions()
    if grid.dim != 2:
        raise ValueError("axisymmetric solver requires a 2D grid (x axial, y radial)")
    if grid.nt <= 0:
        raise ValueError("call grid.makeTime(...) before running the solver")
    if source.p0 is None and source.pMask is None and source.uMask is None:
        raise ValueError("axisymmetric solver requires source.p0 or a time-varying p/u source")
    sim = Simulation(grid, medium, source, sensor if sensor is not None else KWaveSensor(), options,
                     axisymmetric=True)
    return _run_to_output(sim, options)


def _run_to_output(sim: "Simulation", options: SimulationOptions) -> SimulationOutput:
    try:
        if options.progress is None and options.fieldMonitor is None:
            sim.run()
        else:
            iv = max(int(options.fieldMonitorInterval), 1)

            def cb(t, nt, p):
                if options.fieldMonitor is not None and p is not None:
                    options.fieldMonitor(t, p)
                if options.progress is not None:
                    options.progress(t, nt)

            sim.run(callback=cb, interval=iv if options.fieldMonitor is not None else 0)
        return sim.outputs()
    finally:
        sim.close()
