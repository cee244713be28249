"""CPU tests of the C ABI: the library loads, exports every symbol include/kwave_hip.h declares, and
its host-side validation returns the documented error codes (no device compute)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from kwavehip import (AbsorptionMode, KWaveGrid, KWaveMedium, KWaveSensor, KWaveSource, RecordField,
                      SimulationOptions, SourceMode, makeDisc)
from kwavehip import _lib as L
from kwavehip.api import build_desc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "kwave_hip.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|void|const char\*|const kwh_sim_desc\*)\s+(kwh_\w+)\s*\(", text,
                                 flags=re.M)))


def test_library_loads_and_exports_every_declared_symbol():
    lib = L.load()
    assert lib.kwh_abi_version() == L.ABI_VERSION
    syms = declared_symbols()
    assert len(syms) >= 18
    for s in syms:
        assert hasattr(lib, s), s
    assert sorted(L.EXPORTED) == syms


def _case(**kw):
    n = kw.pop("n", 32)
    grid = KWaveGrid(n, 1e-4, n, 1e-4)
    grid.setTime(kw.pop("nt", 10), 2e-8)
    medium = kw.pop("medium", KWaveMedium(1500.0, 1000.0))
    source = kw.pop("source", KWaveSource(p0=makeDisc(n, n, radius=3)))
    sensor = kw.pop("sensor", KWaveSensor())
    return grid, medium, source, sensor, SimulationOptions(**kw)


def validate(*args):
    d, keep, _ = build_desc(*args)
    lib = L.load()
    code = lib.kwh_sim_validate(C.byref(d))
    return code, lib.kwh_last_error().decode()


def test_valid_descriptor_passes():
    assert validate(*_case())[0] == L.OK


def test_nt_zero_rejected():
    g, m, s, se, o = _case()
    g.nt = 0
    code, msg = validate(g, m, s, se, o)
    assert code == L.EINVAL and "makeTime" in msg


def test_nonpositive_sound_speed_rejected():
    code, msg = validate(*_case(medium=KWaveMedium(-1.0, 1000.0)))
    assert code == L.EINVAL and "soundSpeed" in msg


def test_nan_p0_rejected():
    p0 = makeDisc(32, 32, radius=3)
    p0[3, 3] = np.nan
    code, msg = validate(*_case(source=KWaveSource(p0=p0)))
    assert code == L.EINVAL and "p0" in msg


def test_mismatched_map_shape_rejected():
    code, msg = validate(*_case(medium=KWaveMedium(np.full((31, 32), 1500.0, np.float32), 1000.0)))
    assert code == L.EINVAL and "soundSpeed" in msg


def test_cartesian_point_outside_grid_rejected():
    pts = np.array([[1.0], [0.0]], dtype=np.float32)  # 1 m: far outside a 3.2 mm grid
    code, msg = validate(*_case(sensor=KWaveSensor(mask=pts)))
    assert code == L.EINVAL and "Cartesian" in msg


def test_missing_source_rejected():
    code, msg = validate(*_case(source=KWaveSource()))
    assert code == L.EINVAL and "source" in msg


def test_pml_larger_than_grid_rejected():
    code, _ = validate(*_case(n=16, pmlSize=20))
    assert code == L.EINVAL


def test_signal_rows_mismatch_rejected():
    mask = np.zeros((32, 32), np.float32)
    mask[4, 4:7] = 1
    sig = np.ones((2, 5), np.float32)   # 3 points, 2 rows
    code, msg = validate(*_case(source=KWaveSource(pMask=mask, p=sig)))
    assert code == L.EINVAL and "row" in msg


def test_bad_dimension_unsupported():
    d, keep, _ = build_desc(*_case())
    d.dim = 4
    lib = L.load()
    assert lib.kwh_sim_validate(C.byref(d)) == L.EUNSUPPORTED


def test_no_device_reports_edevice():
    if os.path.exists("/dev/kfd"):
        pytest.skip("a device is present")
    lib = L.load()
    h = C.c_void_p()
    assert lib.kwh_ctx_create(0, C.byref(h)) == L.EDEVICE
    assert "device" in lib.kwh_last_error().decode()


def test_python_api_errors_mirror_reference_preconditions():
    from kwavehip import kspaceFirstOrder
    g, m, s, se, o = _case()
    g.nt = 0
    with pytest.raises(ValueError, match="makeTime"):
        kspaceFirstOrder(g, m, s, se, o)
    g, m, s, se, o = _case(source=KWaveSource())
    with pytest.raises(ValueError, match="source"):
        kspaceFirstOrder(g, m, s, se, o)
    g, m, s, se, o = _case()
    o.device = "gpu"
    with pytest.raises(NotImplementedError):
        kspaceFirstOrder(g, m, s, se, o)


def test_c_driver_host_checks():
    # the plain C caller (cdriver/kwh_ivp.c) links against the library and runs its host-only
    # checks: ABI version, a valid descriptor, KWH_EINVAL with a message for nt = 0
    import subprocess
    exe = os.path.join(ROOT, "k-wave-swift_amd", "kwavehip", "kwh_ivp")
    res = subprocess.run([exe, "--validate"], capture_output=True, text=True, timeout=60)
    assert res.returncode == 0, res.stderr
    assert '"validate": "ok"' in res.stdout and "makeTime" in res.stderr
