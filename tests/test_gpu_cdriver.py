"""The plain C caller of the ABI (k-wave-swift_amd/cdriver/kwh_ivp.c, built next to the library):
its outputs, read back through kwh_sim_read_output from C, are bitwise those of the Python API on
the same descriptor, and they agree with the float64 oracle (the C program runs in a child process)."""
import os
import subprocess

import numpy as np
import pytest

from oracle import kwave_oracle as O
from kwavehip import (DTypePrecision, KWaveGrid, KWaveMedium, KWaveSensor, KWaveSource, RecordField,
                      SimulationOptions, kspaceFirstOrder)

from test_gpu_parity import _device  # noqa: F401  (module fixture: the library must load, a device must exist)

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "k-wave-swift_amd", "kwavehip", "kwh_ivp")


def _same_case(n, nt, f64):
    """The problem kwh_ivp builds, through the Python API (same inputs, same descriptor)."""
    grid = KWaveGrid(n, 1e-4, n, 1e-4, n, 1e-4)
    grid.setTime(nt, 0.3 * 1e-4 / 1500.0)
    i, j, k = np.ogrid[:n, :n, :n]
    h, r = n // 2, n // 8
    p0 = (((i - h) ** 2 + (j - h) ** 2 + (k - h) ** 2) <= r * r).astype(np.float32)
    mask = np.zeros((n, n, n), np.float32)
    mask[3 * n // 4] = 1
    opts = SimulationOptions(pmlSize=10, pmlAlpha=2.0, smoothP0=False,
                             dtype=DTypePrecision.float64 if f64 else DTypePrecision.float32)
    return (grid, KWaveMedium(1500.0, 1000.0), KWaveSource(p0=p0),
            KWaveSensor(mask=mask, record=RecordField.p | RecordField.pFinal), opts)


@pytest.mark.parametrize("n,nt,f64", [(64, 40, False), (48, 30, True)])
def test_c_driver_matches_python_api_and_oracle(tmp_path, n, nt, f64):
    print(f"\n[case] C driver kwh_ivp {n}^3 nt={nt} {'f64' if f64 else 'f32'}", flush=True)
    out = tmp_path / "ivp.bin"
    res = subprocess.run([DRIVER, str(n), str(nt), str(out), "f64" if f64 else "f32"], capture_output=True,
                         text=True, timeout=120)
    assert res.returncode == 0, res.stderr
    print(res.stdout.strip(), flush=True)
    assert '"callbacks": 2' in res.stdout
    dt = np.float64 if f64 else np.float32
    raw = out.read_bytes()
    ns, ntr = np.frombuffer(raw[:16], np.int64)
    assert (ns, ntr) == (n * n, nt)
    body = np.frombuffer(raw[16:], dt)
    series = body[:ns * ntr].reshape(ns, ntr)
    pfinal = body[ns * ntr:].reshape(n, n, n)
    case = _same_case(n, nt, f64)
    api = kspaceFirstOrder(*case)
    assert np.array_equal(series, api.p)
    assert np.array_equal(pfinal, api.pFinal)
    ref = O.kspace_first_order(*case)
    for a, b in ((series, ref["p"]), (pfinal, ref["pFinal"])):
        d = a.astype(np.float64) - b
        rl2 = np.sqrt(np.mean(d * d)) / np.sqrt(np.mean(b * b))
        assert rl2 <= (1e-9 if f64 else 1e-4), rl2
