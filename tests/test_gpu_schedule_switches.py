"""GPU checks of the step's schedule switches, read from the environment when a simulation is
created:
- KWH_STRIDED_REV (order in which each XCD walks the tiles of the register y passes): results
  bitwise equal either way (tile order changes no arithmetic);
- KWH_GRAD_DUAL (gradient as two x-pass outputs, the y and z derivatives applied after the x
  inverse, or three x-pass outputs): both against the float64 oracle, and within rounding of each
  other;
- KWH_CORR_FUSE (plane mode at 128 x 128: the velocity kernel inverts the source correction's
  spectrum itself, or a separate plane inverse writes it as a real field): bitwise equal.
"""
import os

import numpy as np
import pytest

from kwavehip import (KWaveGrid, KWaveMedium, KWaveSensor, KWaveSource, RecordField, SimulationOptions, SourceMode,
                      kspaceFirstOrder, makeBall, random_field, toneBurst)

from test_gpu_parity import F32, _device, check, metrics  # noqa: F401  (_device: module fixture)

pytestmark = pytest.mark.gpu


def _with_env(env, fn):
    old = {k: os.environ.get(k) for k in env}
    os.environ.update(env)
    try:
        return fn()
    finally:
        for k, v in old.items():
            if v is None:
                del os.environ[k]
            else:
                os.environ[k] = v


def _hetero_3d(shape, steps):
    c = random_field(shape, 31, 4.0, 1400.0, 1600.0)
    rho = random_field(shape, 32, 4.0, 900.0, 1100.0)
    grid = KWaveGrid(shape[0], 1e-4, shape[1], 1e-4, shape[2], 1e-4)
    grid.setTime(steps, 0.3 * 1e-4 / float(c.max()))
    sens = np.zeros(shape, np.float32)
    sens[shape[0] // 2 + 5, :, :] = 1
    return grid, KWaveMedium(c, rho), KWaveSource(p0=makeBall(*shape, radius=4)), \
        KWaveSensor(mask=sens, record=RecordField.p | RecordField.uz | RecordField.pFinal)


def test_strided_walk_order_is_bitwise_neutral():
    """48 x 64 x 64 with the split passes (register y passes of length 64), heterogeneous: p and uz
    series and pFinal bitwise equal with the y passes walking forwards and backwards."""
    grid, med, src, sens = _hetero_3d((48, 64, 64), 40)
    print("\n[case] strided walk order, 3D 48x64x64", flush=True)
    run = lambda: kspaceFirstOrder(grid, med, src, sens, SimulationOptions(pmlSize=8, dtype=F32))  # noqa: E731
    a = _with_env({"KWH_PLANE": "0", "KWH_STRIDED_REV": "1"}, run)
    b = _with_env({"KWH_PLANE": "0", "KWH_STRIDED_REV": "0"}, run)
    for f in ("p", "uz", "pFinal"):
        assert np.array_equal(np.asarray(getattr(a, f)), np.asarray(getattr(b, f))), f


@pytest.mark.parametrize("plane", ["0", "1"])
def test_gradient_two_outputs_and_three_outputs_against_oracle(plane):
    """The gradient with two x-pass outputs (default) and with three, with the split passes (uz
    recorded) and in plane mode (p only: plane mode records no collocated velocity), each against
    the float64 oracle at the float32 parity tolerances, and within 1e-6 relL2 of each other."""
    grid, med, src, sens = _hetero_3d((48, 64, 64), 40)
    fields = ["p", "uz", "pFinal"] if plane == "0" else ["p", "pFinal"]
    if plane == "1":
        sens = KWaveSensor(mask=sens.mask, record=RecordField.p | RecordField.pFinal)
    for v in ("1", "0"):
        _with_env({"KWH_PLANE": plane, "KWH_GRAD_DUAL": v},
                  lambda: check(f"grad_dual_{v}_plane_{plane}", grid, med, src, sens, fields, pmlSize=8))
    run = lambda: kspaceFirstOrder(grid, med, src, sens, SimulationOptions(pmlSize=8, dtype=F32))  # noqa: E731
    a = _with_env({"KWH_PLANE": plane, "KWH_GRAD_DUAL": "1"}, run)
    b = _with_env({"KWH_PLANE": plane, "KWH_GRAD_DUAL": "0"}, run)
    for f in fields:
        l2, _ = metrics(getattr(a, f), getattr(b, f))
        assert l2 <= 1e-6, (f, l2)


def test_fused_source_correction_is_bitwise_neutral():
    """Plane mode on a 24 x 128 x 128 grid with an additive (k-space corrected) ux source in one x
    slab: the velocity kernel inverting the correction spectrum itself gives the same bits as the
    separate plane inverse."""
    shape = (24, 128, 128)
    grid = KWaveGrid(shape[0], 1e-4, shape[1], 1e-4, shape[2], 1e-4)
    grid.setTime(40, 2e-8)
    um = np.zeros(shape, np.float32)
    um[6, 40:88, 52:76] = 1
    sig = (0.1 * toneBurst(1 / grid.dt, 2e6, 2)).astype(np.float32)
    src = KWaveSource(uMask=um, ux=sig, uMode=SourceMode.additive)
    sens = np.zeros(shape, np.float32)
    sens[12, :, 64] = 1
    sensor = KWaveSensor(mask=sens, record=RecordField.p | RecordField.pFinal)
    med = KWaveMedium(1500.0, 1000.0)
    print("\n[case] fused source correction, 3D 24x128x128 plane mode", flush=True)
    run = lambda: kspaceFirstOrder(grid, med, src, sensor, SimulationOptions(pmlSize=6, smoothP0=False,  # noqa: E731
                                                                             dtype=F32))
    a = _with_env({"KWH_CORR_FUSE": "1"}, run)
    b = _with_env({"KWH_CORR_FUSE": "0"}, run)
    for f in ("p", "pFinal"):
        assert np.array_equal(np.asarray(getattr(a, f)), np.asarray(getattr(b, f))), f
