"""kwavehip -- MI355X (gfx950) HIP implementation of KWaveSwift's `kspaceFirstOrder` time loop.

The compute runs in libkwavehip.so through the C ABI of include/kwave_hip.h; this package is the
host-side mirror of the reference's Swift API (Sources/KWave/...) plus input helpers.
"""
from .api import (AbsorptionMode, DeviceKind, DirectivityPattern, DTypePrecision, KWaveGrid, KWaveMedium,
                  KWaveSensor, KWaveSource, RecordField, Simulation, SimulationBatch, SimulationOptions,
                  SimulationOutput, SourceMode, flatNonzeroIndices, kspaceFirstOrder, kspaceFirstOrderAS,
                  timeReversal)
from .inputs import (db2neper, gaussian, getWin, makeBall, makeCartCircle, makeCircle, makeDisc, neper2db,
                     random_field, toneBurst)
from .transducer import DelayMaskMode, KWaveTransducer, KWaveTransducerGeometry
from ._lib import KWaveHIPError, KWaveNonFiniteError

__all__ = [
    "AbsorptionMode", "DeviceKind", "DirectivityPattern", "DTypePrecision", "KWaveGrid", "KWaveMedium",
    "KWaveSensor", "KWaveSource", "RecordField", "Simulation", "SimulationBatch", "SimulationOptions",
    "SimulationOutput", "timeReversal", "KWaveTransducer", "KWaveTransducerGeometry", "DelayMaskMode",
    "SourceMode", "flatNonzeroIndices", "kspaceFirstOrder", "kspaceFirstOrderAS", "db2neper", "gaussian", "getWin", "makeBall",
    "makeCartCircle", "makeCircle", "makeDisc", "neper2db", "random_field", "toneBurst",
    "KWaveHIPError", "KWaveNonFiniteError",
]
