"""Parity cases for the middle (outermost-axis) pass with its inner radix across lanes (DPP),
the opt-in KWH_MIDDLE_XLANE=1 variant. The library reads that switch once per process, so
test_gpu_xlane.py runs this file in a child process with the switch set; here each case is
checked against the float64 oracle at the float32 parity tolerances (relL2 <= 1e-4, relMax <=
1e-3; FP64 <= 1e-9 / 1e-8). Not collected by pytest itself (no test_ prefix).

Middle-axis lengths covered: 96 (4 x 24), 160 (4 x 40), 128 (8 x 16), 256 (16 x 16), 64 (8 x 8),
so every lane distance of the cross-lane network (1, 2, 4, 8) runs.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
for p in (HERE, ROOT, os.path.join(ROOT, "k-wave-swift_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

from kwavehip import (AbsorptionMode, KWaveGrid, KWaveMedium, KWaveSensor, KWaveSource, RecordField,  # noqa: E402
                      SourceMode, makeBall, makeDisc, random_field, toneBurst)
from test_gpu_parity import F32, F64, check  # noqa: E402


def case_2d(nx, ny, dtype):
    shape = (nx, ny)
    dx = 1e-4
    c = random_field(shape, 91, 4.0, 1400.0, 1600.0)
    rho = random_field(shape, 92, 4.0, 900.0, 1100.0)
    med = KWaveMedium(c, rho, alphaCoeff=0.75, alphaPower=1.5, alphaMode=AbsorptionMode.powerLaw, bOnA=6.0)
    grid = KWaveGrid(nx, dx, ny, dx)
    grid.setTime(80, 0.3 * dx / float(c.max()))
    umask = np.zeros(shape, np.float32)
    umask[nx // 2 - 3, ny // 2 - 4: ny // 2 + 4] = 1
    sig = (0.2 * toneBurst(1 / grid.dt, 3e6, 2)).astype(np.float32)
    sens = np.zeros(shape, np.float32)
    sens[nx // 2 + 6, :] = 1
    check(f"xlane_2d_{nx}x{ny}", grid, med,
          KWaveSource(p0=makeDisc(nx, ny, radius=5), uMask=umask, ux=sig, uMode=SourceMode.additive),
          KWaveSensor(mask=sens, record=RecordField.p | RecordField.ux | RecordField.uy | RecordField.pFinal),
          ["p", "ux", "uy", "pFinal"], dtype=dtype, pmlSize=10)


def case_3d(nx, ny, nz):
    shape = (nx, ny, nz)
    c = random_field(shape, 93, 4.0, 1400.0, 1600.0)
    rho = random_field(shape, 94, 4.0, 900.0, 1100.0)
    med = KWaveMedium(c, rho, alphaCoeff=0.5, alphaPower=1.1, alphaMode=AbsorptionMode.powerLaw, bOnA=5.0)
    grid = KWaveGrid(nx, 1e-4, ny, 1e-4, nz, 1e-4)
    grid.setTime(30, 0.3 * 1e-4 / float(c.max()))
    sens = np.zeros(shape, np.float32)
    sens[:, ny // 2, :] = 1
    check(f"xlane_3d_{nx}x{ny}x{nz}", grid, med, KWaveSource(p0=makeBall(nx, ny, nz, radius=4)),
          KWaveSensor(mask=sens, record=RecordField.p | RecordField.ux | RecordField.pFinal),
          ["p", "ux", "pFinal"], pmlSize=6, workers=8)


CASES = {
    "2d_96x160": lambda: case_2d(96, 160, F32),
    "2d_160x96": lambda: case_2d(160, 96, F32),
    "2d_128x64": lambda: case_2d(128, 64, F32),
    "2d_256x64": lambda: case_2d(256, 64, F32),
    "2d_64x96_f64": lambda: case_2d(64, 96, F64),
    "3d_96x32x64": lambda: case_3d(96, 32, 64),
}

if __name__ == "__main__":
    assert os.environ.get("KWH_MIDDLE_XLANE") == "1", "run with KWH_MIDDLE_XLANE=1"
    for name in sys.argv[1:] or list(CASES):
        CASES[name]()
        print(f"[ok] {name}", flush=True)
