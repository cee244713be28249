"""GPU: the autotuner's schedule cache persists across processes (kwh_sim_autotune with
KWH_AUTOTUNE_CACHE): the first process times the candidates and appends one entry keyed by device,
grid shape and precision; a second process takes the cached schedule without timing; another
precision is another entry. Each process runs in its own subprocess, one at a time."""
import ctypes as C
import os
import subprocess
import sys

import pytest

from kwavehip import _lib as L

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

SCRIPT = """
import sys
sys.path[:0] = [{root!r}, {pkg!r}]
import numpy as np
from kwavehip import (DTypePrecision, KWaveGrid, KWaveMedium, KWaveSensor, KWaveSource, Simulation,
                      SimulationOptions, makeBall)
n = (40, 36, 24)
grid = KWaveGrid(n[0], 1e-4, n[1], 1e-4, n[2], 1e-4)
grid.setTime(10, 2e-8)
dt = DTypePrecision.float64 if sys.argv[1] == "f64" else DTypePrecision.float32
sim = Simulation(grid, KWaveMedium(1500.0, 1000.0), KWaveSource(p0=makeBall(*n, radius=4)), KWaveSensor(),
                 SimulationOptions(pmlSize=4, dtype=dt))
print(sim.autotune(), flush=True)
sim.run()
print("pmax", float(np.abs(sim.outputs().pFinal).max()), flush=True)
"""


@pytest.fixture(scope="module", autouse=True)
def _device():
    lib = L.load()
    h = C.c_void_p()
    if lib.kwh_ctx_create(0, C.byref(h)) != 0:
        pytest.skip("no HIP device")
    lib.kwh_ctx_destroy(h)


def _run(cache, dtype):
    env = dict(os.environ, KWH_AUTOTUNE_CACHE=str(cache))
    code = SCRIPT.format(root=ROOT, pkg=os.path.join(ROOT, "k-wave-swift_amd"))
    r = subprocess.run([sys.executable, "-c", code, dtype], env=env, capture_output=True, text=True, timeout=180)
    assert r.returncode == 0, r.stderr[-2000:]
    return r.stdout


def test_schedule_cache_persists_across_processes(tmp_path):
    cache = tmp_path / "sub" / "autotune.txt"
    print("\n[case] autotune cache, first process", flush=True)
    out1 = _run(cache, "f32")
    assert "chosen:" in out1 and f"cached in {cache}" in out1
    lines = cache.read_text().strip().splitlines()
    assert len(lines) == 1
    dev, prob, ntr, ps, lds, mb = lines[0].split()
    assert dev.startswith("gfx") and prob.startswith("3:40:36:24:4:")
    print("[case] autotune cache, second process", flush=True)
    out2 = _run(cache, "f32")
    assert "cached schedule" in out2 and "chosen:" not in out2
    assert out1.splitlines()[-1] == out2.splitlines()[-1]          # same result either way
    print("[case] autotune cache, float64 entry", flush=True)
    out3 = _run(cache, "f64")
    assert "chosen:" in out3
    assert len(cache.read_text().strip().splitlines()) == 2
