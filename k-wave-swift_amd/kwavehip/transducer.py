"""Linear-array transducer (KWaveTransducer.swift:1-316, k-Wave's kWaveTransducerSimple /
NotATransducer): element masks, transmit / elevation delays, apodization, the padded drive
signal, and the receive-side element combination and scan-line beamforming.

Masks are C-order over the grid like every other mask of this package. The drive is handed to
the solver as a delayed velocity source (`velocity_source`): one signal row plus a delay and a
scale per source point, expanded into per-point rows on the device; receive traces are combined
per element on the device (`Simulation.element_traces`).
"""
from dataclasses import dataclass
from enum import Enum
from typing import Optional, Sequence, Union

import numpy as np


@dataclass(frozen=True)
class KWaveTransducerGeometry:
    """KWaveTransducer.swift:8-38. The array lies in the plane x = position[0]; elements are laid
    out along y (width + kerf per element) and extend along z (elevation). Grid points, 0-based."""
    numberElements: int
    dy: float
    dz: float
    elementWidth: int = 1
    elementLength: int = 20
    elementSpacing: int = 0
    position: tuple = (0, 0, 0)

    def __post_init__(self):
        if not (self.numberElements > 0 and self.elementWidth > 0 and self.elementLength > 0
                and self.elementSpacing >= 0):
            raise ValueError("invalid transducer geometry")

    @property
    def elementPitch(self) -> float:
        return (self.elementSpacing + self.elementWidth) * self.dy

    @property
    def transducerWidth(self) -> int:
        return self.numberElements * self.elementWidth + (self.numberElements - 1) * self.elementSpacing


class DelayMaskMode(Enum):
    both = "both"
    elevationOnly = "elevationOnly"
    azimuthOnly = "azimuthOnly"


Apodization = Union[str, Sequence[float]]   # "rectangular", "hanning" or explicit per-element values


class KWaveTransducer:
    """KWaveTransducer.swift:45-316."""

    def __init__(self, geometry: KWaveTransducerGeometry, gridSize: Sequence[int], dt: float,
                 soundSpeed: float = 1540.0, activeElements: Optional[Sequence[bool]] = None,
                 focusDistance: float = np.inf, elevationFocusDistance: float = np.inf,
                 steeringAngle: float = 0.0, transmitApodization: Apodization = "rectangular",
                 receiveApodization: Apodization = "rectangular", inputSignal: Optional[np.ndarray] = None):
        if len(gridSize) != 3:
            raise ValueError("transducers require a 3D grid")
        nx, ny, nz = (int(v) for v in gridSize)
        px, py, pz = geometry.position
        if px >= nx:
            raise ValueError("transducer outside the grid in x")
        if py + geometry.transducerWidth > ny:
            raise ValueError("transducer too large or outside the grid in y")
        if pz + geometry.elementLength > nz:
            raise ValueError("transducer too large or outside the grid in z")
        self.geometry = geometry
        self.gridSize = (nx, ny, nz)
        self.dt = float(dt)
        self.soundSpeed = float(soundSpeed)
        self.activeElements = (np.ones(geometry.numberElements, bool) if activeElements is None
                               else np.asarray(activeElements, bool))
        if self.activeElements.size != geometry.numberElements:
            raise ValueError("activeElements must have one entry per element")
        self.focusDistance = float(focusDistance)
        self.elevationFocusDistance = float(elevationFocusDistance)
        self.steeringAngle = float(steeringAngle)
        self.transmitApodization = transmitApodization
        self.receiveApodization = receiveApodization
        self.inputSignal = None if inputSignal is None else np.asarray(inputSignal, np.float32)
        # element number (1-based, 0 = empty) and position along the element length (1-based)
        e = np.arange(geometry.numberElements)
        y0 = py + (geometry.elementWidth + geometry.elementSpacing) * e
        ys = (y0[:, None] + np.arange(geometry.elementWidth)[None, :]).ravel()            # [E*W]
        el = np.repeat(e + 1, geometry.elementWidth)
        zs = pz + np.arange(geometry.elementLength)
        flat = ((px * ny + ys)[:, None] * nz + zs[None, :]).ravel()
        self._indexed = np.zeros(nx * ny * nz, np.int64)
        self._voxel = np.zeros(nx * ny * nz, np.int64)
        self._indexed[flat] = np.repeat(el, geometry.elementLength)
        self._voxel[flat] = np.tile(np.arange(1, geometry.elementLength + 1), ys.size)

    # ------------------------------------------------------------------ masks
    @property
    def numberActiveElements(self) -> int:
        return int(self.activeElements.sum())

    @property
    def allElementsMask(self) -> np.ndarray:
        return (self._indexed != 0).astype(np.float32).reshape(self.gridSize)

    @property
    def activeElementsMask(self) -> np.ndarray:
        act = np.concatenate([[False], self.activeElements])
        return act[self._indexed].astype(np.float32).reshape(self.gridSize)

    def _indexed_active(self) -> np.ndarray:
        """Active-element number renumbered from 1 at the lowest active element; 0 = empty."""
        if not self.activeElements.any():
            return np.zeros_like(self._indexed)
        lowest = int(np.argmax(self.activeElements))
        act = np.concatenate([[False], self.activeElements])
        return np.where(act[self._indexed], self._indexed - lowest, 0)

    # ------------------------------------------------------------------ beamforming
    @property
    def beamformingDelays(self) -> np.ndarray:
        """Per-active-element transmit delays in samples (KWaveTransducer.swift:146-161)."""
        n = self.numberActiveElements
        idx = np.arange(n) - (n - 1) / 2.0
        pitch = self.geometry.elementPitch
        th = np.deg2rad(self.steeringAngle)
        if np.isinf(self.focusDistance):
            sec = pitch * idx * np.sin(th) / self.soundSpeed
        else:
            q = idx * pitch / self.focusDistance
            sec = self.focusDistance / self.soundSpeed * (1.0 - np.sqrt(1.0 + q * q - 2.0 * q * np.sin(th)))
        return np.rint(sec / self.dt).astype(np.int64)

    @property
    def elevationBeamformingDelays(self) -> np.ndarray:
        """Per-voxel elevation delays along the element length (KWaveTransducer.swift:165-177)."""
        n = self.geometry.elementLength
        if np.isinf(self.elevationFocusDistance):
            return np.zeros(n, np.int64)
        idx = np.arange(n) - (n - 1) / 2.0
        f = self.elevationFocusDistance
        sec = (f - np.sqrt((idx * self.geometry.dz) ** 2 + f * f)) / self.soundSpeed
        return -np.rint(sec / self.dt).astype(np.int64)

    def delayMask(self, mode: DelayMaskMode = DelayMaskMode.both) -> np.ndarray:
        """Per-voxel transmit delay in samples, minimum active delay shifted to zero
        (KWaveTransducer.swift:185-208). Flat, C-order."""
        mask = np.zeros(self._indexed.size, np.int64)
        act = self._indexed_active()
        on = act != 0
        if (not np.isinf(self.focusDistance) or self.steeringAngle != 0) and mode != DelayMaskMode.elevationOnly:
            mask[on] = -self.beamformingDelays[act[on] - 1]
        if (not np.isinf(self.elevationFocusDistance) and self.geometry.elementLength > 1
                and mode != DelayMaskMode.azimuthOnly):
            mask[on] += self.elevationBeamformingDelays[self._voxel[on] - 1]
        if on.any():
            mn = int(mask[on].min())
            if mn != 0:
                mask[on] -= mn
        return mask

    # ------------------------------------------------------------------ apodization
    def _apodization(self, spec: Apodization) -> np.ndarray:
        n = self.numberActiveElements
        if isinstance(spec, str):
            if spec == "rectangular":
                return np.ones(n)
            if spec == "hanning":
                return np.ones(1) if n <= 1 else 0.5 - 0.5 * np.cos(2 * np.pi * np.arange(n) / (n - 1))
            raise ValueError(f"unknown apodization {spec!r}")
        v = np.asarray(spec, np.float64)
        if v.size != n:
            raise ValueError("apodization length must match active elements")
        return v

    def getTransmitApodization(self) -> np.ndarray:
        return self._apodization(self.transmitApodization)

    def getReceiveApodization(self) -> np.ndarray:
        return self._apodization(self.receiveApodization)

    @property
    def transmitApodizationMask(self) -> np.ndarray:
        act = self._indexed_active()
        apod = np.concatenate([[0.0], self.getTransmitApodization()])
        return apod[act].astype(np.float32).reshape(self.gridSize)

    # ------------------------------------------------------------------ signals
    def paddedInputSignal(self) -> np.ndarray:
        """Drive signal with zeros added so the largest delay fits (KWaveTransducer.swift:244-258)."""
        if self.inputSignal is None:
            raise ValueError("Transducer input signal is not defined")
        s = self.inputSignal
        dmax = int(self.delayMask().max(initial=0))
        nz = np.flatnonzero(s)
        lead = int(nz[0]) if nz.size else s.size
        trail = int(s.size - 1 - nz[-1]) if nz.size else s.size
        if lead < dmax + 1:
            s = np.concatenate([np.zeros(dmax - lead + 1, np.float32), s])
        if trail < dmax + 1:
            s = np.concatenate([s, np.zeros(dmax - trail + 1, np.float32)])
        return s

    def velocity_source(self):
        """The transmit as a delayed velocity source on the active-element voxels: source point k
        (ascending flat index) plays the padded signal delayed by delayMask()[k] samples, scaled
        by the transmit apodization of its element (k-Wave's transducer source). Returns
        (mask, signal, delays, scales) for KWaveSource(uMask=mask, ux=signal, uxDelay=delays,
        uxScale=scales)."""
        mask = self.activeElementsMask
        flat = np.flatnonzero(mask.ravel())
        delays = self.delayMask()[flat].astype(np.int32)
        scales = self.transmitApodizationMask.ravel()[flat].astype(np.float32)
        return mask, self.paddedInputSignal(), delays, scales

    def receive_combination(self):
        """Per sensor row (active voxels, ascending flat index): its element (0-based among the
        active ones) and its elevation delay, the inputs of combineSensorData."""
        act = self._indexed_active()
        flat = np.flatnonzero(act)
        return (act[flat] - 1).astype(np.int32), self.delayMask(DelayMaskMode.elevationOnly)[flat].astype(np.int32)

    # ------------------------------------------------------------------ receive
    def combineSensorData(self, sensorData: np.ndarray) -> np.ndarray:
        """Per-element series from per-voxel sensor data [nVoxels, nt], compensating the
        elevation delays (KWaveTransducer.swift:263-289); host reference of the device path."""
        if np.isinf(self.elevationFocusDistance):
            raise ValueError("combineSensorData requires elevation focusing")
        data = np.asarray(sensorData, np.float32)
        g = self.geometry
        expected = self.numberActiveElements * g.elementWidth * g.elementLength
        if data.ndim != 2 or data.shape[0] != expected:
            raise ValueError("sensorData rows must match the active transducer grid points")
        nt = data.shape[1]
        elem, dly = self.receive_combination()
        out = np.zeros((self.numberActiveElements, nt), np.float32)
        for row in range(expected):
            d = int(dly[row])
            out[elem[row], :nt - d] += data[row, d:]
        return out * np.float32(1.0 / (g.elementWidth * g.elementLength))

    def scanLine(self, sensorData: np.ndarray) -> np.ndarray:
        """Receive beamforming into one line (KWaveTransducer.swift:293-315)."""
        data = np.asarray(sensorData, np.float32)
        n = self.numberActiveElements
        if data.ndim != 2 or data.shape[0] != n:
            raise ValueError("sensorData must be [numberActiveElements, Nt]")
        nt = data.shape[1]
        apod = self.getReceiveApodization().astype(np.float32)
        delays = -self.beamformingDelays
        line = np.zeros(nt, np.float32)
        for e in range(n):
            d = int(delays[e])
            if d > 0:
                line[:nt - d] += data[e, d:] * apod[e]
            elif d < 0:
                line[-d:] += data[e, :nt + d] * apod[e]
            else:
                line += data[e] * apod[e]
        return line
