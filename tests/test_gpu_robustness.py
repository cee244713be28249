"""GPU: error paths of the C ABI that must not corrupt device memory or hide a blow-up.

- kwh_sim_load_state rejects a blob of another simulation, a corrupted header or inconsistent
  counters with KWH_EINVAL before copying anything; the sim then still runs to the right answer.
- kwh_sim_set_nonfinite_check: a run above the stability limit (time step CFL 1.5 with power-law
  absorption, whose explicit absorption term is not k-space corrected) returns KWH_ENONFINITE; a
  stable run with the check on is bitwise the run without it.
- pMax / pMin keep a NaN once one appears, as the oracle's max over the series does.
- receive delays near INT32_MAX drop their terms instead of wrapping to a negative time index.
NaN arithmetic does not fault the GPU: the blow-up case is an ordinary run of finite length.
"""
import ctypes as C
import struct

import numpy as np
import pytest

from oracle import kwave_oracle as O
from kwavehip import (AbsorptionMode, KWaveGrid, KWaveMedium, KWaveNonFiniteError, KWaveSensor, KWaveSource,
                      RecordField, Simulation, SimulationOptions, kspaceFirstOrder, makeDisc)
from kwavehip import _lib as L

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _device():
    lib = L.load()
    h = C.c_void_p()
    if lib.kwh_ctx_create(0, C.byref(h)) != 0:
        pytest.skip("no HIP device")
    lib.kwh_ctx_destroy(h)


def _args(n=32, steps=60, record=RecordField.p | RecordField.pMax):
    grid = KWaveGrid(n, 1e-4, n, 1e-4)
    grid.setTime(steps, 2e-8)
    sens = np.zeros((n, n), np.float32)
    sens[5, :] = 1
    return (grid, KWaveMedium(1500.0, 1000.0), KWaveSource(p0=makeDisc(n, n, radius=4)),
            KWaveSensor(mask=sens, record=record))


# header layout: magic u32, elem_bytes/dim/nt i32, n[3] i32, record u32, nsens i64, payload u64,
# steps i32, rec_steps i32 (56 bytes, no padding)
_HDR = struct.Struct("<I3i3iIqQii")


def test_load_state_rejects_foreign_and_corrupt_blobs():
    print("\n[case] load_state validation", flush=True)
    args = _args()
    full = Simulation(*args)
    full.run()
    b = Simulation(*args)
    b.run(0, 25)
    blob = b.state()
    fields = list(_HDR.unpack_from(blob))
    assert fields[0] == 0x4B574853 and fields[-2] == 25 and fields[-1] == 25

    def patched(**kw):
        f = list(fields)
        names = ["magic", "elem", "dim", "nt", "n0", "n1", "n2", "record", "nsens", "payload", "steps", "rec"]
        for k, v in kw.items():
            f[names.index(k)] = v
        return _HDR.pack(*f) + blob[_HDR.size:]

    c = Simulation(*args)
    bad = [patched(magic=0), patched(rec=10**6), patched(rec=24), patched(steps=10**6, rec=10**6),
           patched(steps=-1, rec=-1), patched(nsens=7), patched(record=1), blob[:-1], blob[:8]]
    for i, blob_bad in enumerate(bad):
        with pytest.raises(ValueError):
            c.load_state(blob_bad)
        assert c.steps_done == 0, i
    # a blob of another simulation: other record set, other precision
    other = Simulation(*_args(record=RecordField.p))
    other.run(0, 25)
    with pytest.raises(ValueError):
        c.load_state(other.state())
    o64 = Simulation(*_args(), SimulationOptions(dtype=__import__("kwavehip").DTypePrecision.float64))
    o64.run(0, 25)
    with pytest.raises(ValueError):
        c.load_state(o64.state())
    # nothing was copied: the sim still runs from step 0 to the uninterrupted result
    c.run()
    assert np.array_equal(full.output(L.OUT_P_FINAL), c.output(L.OUT_P_FINAL))
    assert np.array_equal(full.output(L.OUT_P), c.output(L.OUT_P))
    # and the valid blob still restores
    d = Simulation(*args)
    d.load_state(blob)
    d.run(25, 60)
    assert np.array_equal(full.output(L.OUT_P_MAX), d.output(L.OUT_P_MAX))


def _unstable(n=64, steps=200, cfl=1.5):
    grid = KWaveGrid(n, 1e-4, n, 1e-4)
    grid.setTime(steps, cfl * 1e-4 / 1500.0)
    med = KWaveMedium(1500.0, 1000.0, alphaCoeff=5.0, alphaPower=1.5, alphaMode=AbsorptionMode.powerLaw)
    sens = np.zeros((n, n), np.float32)
    sens[n // 2, :] = 1
    return (grid, med, KWaveSource(p0=makeDisc(n, n, radius=5)),
            KWaveSensor(mask=sens, record=RecordField.p | RecordField.pMax | RecordField.pMin))


def test_over_cfl_run_returns_enonfinite():
    print("\n[case] CFL 1.5 power-law absorption, check every 16 steps", flush=True)
    args = _unstable()
    with np.errstate(all="ignore"):
        ref = O.kspace_first_order(*args, SimulationOptions(pmlSize=10), dtype=np.float32)
    assert not np.all(np.isfinite(ref["pFinal"]))      # the float32 oracle blows up as well
    sim = Simulation(*args, SimulationOptions(pmlSize=10))
    sim.set_nonfinite_check(16)
    with pytest.raises(KWaveNonFiniteError) as e:
        sim.run()
    assert e.value.code == L.ENONFINITE and "non-finite pressure" in str(e.value)
    done = sim.steps_done
    assert 0 < done < args[0].nt and done % 16 == 0       # stopped at the first flagged chunk
    assert not np.all(np.isfinite(sim.output(L.OUT_P_FINAL)))
    # through the one-call API with the option, and with a callback (progress)
    with pytest.raises(KWaveNonFiniteError):
        kspaceFirstOrder(*args, SimulationOptions(pmlSize=10, nonfiniteCheckEvery=32))
    seen = []
    with pytest.raises(KWaveNonFiniteError):
        kspaceFirstOrder(*args, SimulationOptions(pmlSize=10, nonfiniteCheckEvery=32,
                                                  progress=lambda t, nt: seen.append(t)))
    assert seen and (seen[-1] + 1) % 32 == 0


def test_unchecked_blowup_keeps_nan_in_aggregates():
    print("\n[case] CFL 1.5 without the check: pMax / pMin carry the NaN", flush=True)
    args = _unstable()
    gpu = kspaceFirstOrder(*args, SimulationOptions(pmlSize=10))
    bad = ~np.all(np.isfinite(gpu.p), axis=1)        # sensor points whose series went non-finite
    assert bad.any()
    assert np.all(np.isnan(gpu.pMax[bad]) | np.isinf(gpu.pMax[bad]))
    assert np.all(np.isnan(gpu.pMin[bad]) | np.isinf(gpu.pMin[bad]))
    with np.errstate(invalid="ignore"):
        assert np.array_equal(np.isnan(gpu.pMax), np.isnan(np.max(gpu.p, axis=1)))


def test_check_is_neutral_on_a_stable_run():
    print("\n[case] stable run with the check every 7 steps", flush=True)
    args = _args()
    a = Simulation(*args)
    a.run()
    b = Simulation(*args)
    b.set_nonfinite_check(7)
    b.run(0, 30)
    b.run(30, 60, callback=lambda t, nt, p: None, interval=5)
    for f in (L.OUT_P, L.OUT_P_MAX, L.OUT_P_FINAL):
        assert np.array_equal(a.output(f), b.output(f))
    with pytest.raises(ValueError):
        b.set_nonfinite_check(-1)


def test_huge_receive_delay_drops_its_terms():
    print("\n[case] element traces with a delay near INT32_MAX", flush=True)
    args = _args()
    sim = Simulation(*args)
    sim.run()
    ns = sim.n_sensor
    el = np.zeros(ns, np.int32)
    el[ns // 2:] = 1
    delay = np.zeros(ns, np.int32)
    delay[0] = 2**31 - 1
    delay[ns - 1] = 2**31 - 5
    got = sim.element_traces(L.OUT_P, el, delay, 2, 1.0)
    ser = sim.output(L.OUT_P).astype(np.float64)
    want = np.stack([ser[1:ns // 2].sum(0), ser[ns // 2:ns - 1].sum(0)])
    np.testing.assert_allclose(got, want, rtol=1e-6, atol=1e-6 * np.abs(want).max())
