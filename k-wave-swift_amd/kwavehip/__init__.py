"""kwavehip -- MI355X (gfx950) HIP implementation of KWaveSwift's `kspaceFirstOrder` time loop.

The compute runs in libkwavehip.so through the C ABI of include/kwave_hip.h; this package is the
host-side mirror of the reference's Swift API (Sources/KWave/...) plus input helpers.
"""
from .api import (AbsorptionMode, DeviceKind, DirectivityPattern, DTypePrecision, KWaveGrid, KWaveMedium,
                  KWaveSensor, KWaveSource, RecordField, Simulation, SimulationOptions, SimulationOutput,
                  SourceMode, flatNonzeroIndices, kspaceFirstOrder)
from .inputs import (db2neper, gaussian, getWin, makeBall, makeCartCircle, makeCircle, makeDisc, neper2db,
                     random_field, toneBurst)

__all__ = [
    "AbsorptionMode", "DeviceKind", "DirectivityPattern", "DTypePrecision", "KWaveGrid", "KWaveMedium",
    "KWaveSensor", "KWaveSource", "RecordField", "Simulation", "SimulationOptions", "SimulationOutput",
    "SourceMode", "flatNonzeroIndices", "kspaceFirstOrder", "db2neper", "gaussian", "getWin", "makeBall",
    "makeCartCircle", "makeCircle", "makeDisc", "neper2db", "random_field", "toneBurst",
]
