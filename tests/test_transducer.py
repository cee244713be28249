"""Linear-array transducer model (KWaveTransducer.swift) on the CPU.

The reference pins this type against k-wave-python's NotATransducer through an HDF5 fixture
(Tests/KWaveTests/ParityTransducerTests.swift) whose generator needs k-wave-python; neither is
available here, so the model is pinned by properties its definition fixes, on the reference
test's own configuration (12x64x16 grid, 8 elements of 4x12 voxels, kerf 2, position (0, 4, 2),
elements 0-1 inactive, focus 30 mm, elevation focus 20 mm, steering 10 deg, Hanning).
"""
import numpy as np
import pytest

from kwavehip import DelayMaskMode, KWaveTransducer, KWaveTransducerGeometry


def make(signal=None, **kw):
    geom = KWaveTransducerGeometry(numberElements=8, elementWidth=4, elementLength=12, elementSpacing=2,
                                   position=(0, 4, 2), dy=1e-4, dz=1e-4)
    active = np.ones(8, bool)
    active[:2] = False
    args = dict(activeElements=active, focusDistance=30e-3, elevationFocusDistance=20e-3, steeringAngle=10.0,
                transmitApodization="hanning", receiveApodization="hanning", inputSignal=signal)
    args.update(kw)
    return KWaveTransducer(geom, [12, 64, 16], 2e-8, 1540.0, **args)


def test_masks_cover_the_elements():
    t = make()
    allm, act = t.allElementsMask, t.activeElementsMask
    assert allm.sum() == 8 * 4 * 12 and act.sum() == 6 * 4 * 12
    assert np.all(allm[1:] == 0)                          # only the plane x = 0
    ys = np.flatnonzero(allm[0].any(axis=1))
    zs = np.flatnonzero(allm[0].any(axis=0))
    assert zs.min() == 2 and zs.max() == 13                # elevation extent 12 from z = 2
    # element e occupies y = 4 + 6 e .. 4 + 6 e + 3 (width 4, kerf 2)
    expect = np.concatenate([4 + 6 * e + np.arange(4) for e in range(8)])
    assert np.array_equal(ys, expect)
    assert np.all(act[0, 4:16] == 0)                       # elements 0, 1 inactive
    idx = t._indexed_active().reshape(12, 64, 16)
    assert idx[0, 16, 2] == 1 and idx[0, 4 + 6 * 7, 2] == 6


def test_beamforming_delay_symmetries():
    unsteered = make(steeringAngle=0.0)
    d = unsteered.beamformingDelays
    assert np.array_equal(d, d[::-1])                      # focus only: symmetric about the centre
    assert np.all(d <= 0) and d[0] == d.min()              # outer elements fire first
    plane = make(focusDistance=np.inf, steeringAngle=10.0).beamformingDelays
    assert np.array_equal(plane, -plane[::-1])             # steering only: antisymmetric
    assert np.all(np.diff(plane) >= 0)
    assert np.all(make(focusDistance=np.inf, steeringAngle=0.0).beamformingDelays == 0)


def test_elevation_delays_and_delay_mask():
    t = make()
    el = t.elevationBeamformingDelays
    assert el.size == 12 and np.array_equal(el, el[::-1]) and np.all(el >= 0)
    dm = t.delayMask()
    act = t._indexed_active()
    on = act != 0
    assert dm[on].min() == 0 and np.all(dm[~on] == 0)
    # each active voxel: -beamforming(element) + elevation(voxel) - min, recomposed from the parts
    bf = -t.beamformingDelays[act[on] - 1]
    ev = el[t._voxel[on] - 1]
    raw = bf + ev
    assert np.array_equal(dm[on], raw - raw.min())
    assert np.array_equal(t.delayMask(DelayMaskMode.elevationOnly)[on], ev - ev.min())


def test_apodization_and_padded_signal():
    sig = np.concatenate([np.zeros(3), np.sin(np.linspace(0, 6 * np.pi, 40)), np.zeros(2)]).astype(np.float32)
    t = make(signal=sig)
    ap = t.getTransmitApodization()
    assert ap.size == 6 and ap[0] == 0 and ap[-1] == 0 and np.allclose(ap, ap[::-1])
    m = t.transmitApodizationMask.ravel()
    act = t._indexed_active()
    assert np.allclose(m[act != 0], ap[act[act != 0] - 1]) and np.all(m[act == 0] == 0)
    pad = t.paddedInputSignal()
    dmax = int(t.delayMask().max())
    nz = np.flatnonzero(pad)
    assert nz[0] >= dmax + 1 and pad.size - 1 - nz[-1] >= dmax + 1
    core = pad[nz[0]:nz[-1] + 1]
    assert np.array_equal(core, sig[np.flatnonzero(sig)[0]:np.flatnonzero(sig)[-1] + 1])


def test_velocity_source_form_matches_delay_mask():
    sig = np.hanning(30).astype(np.float32)
    t = make(signal=sig)
    mask, row, delays, scales = t.velocity_source()
    flat = np.flatnonzero(mask.ravel())
    assert np.array_equal(mask, t.activeElementsMask)
    assert np.array_equal(delays, t.delayMask()[flat])
    assert np.allclose(scales, t.transmitApodizationMask.ravel()[flat])
    assert np.array_equal(row, t.paddedInputSignal())


def test_combine_recovers_element_waveforms():
    """Voxel traces that are the element waveform delayed by the voxel's elevation delay combine
    back into the waveform (for samples every voxel covers)."""
    rng = np.random.default_rng(4)
    t = make()
    elem, dly = t.receive_combination()
    nt = 120
    w = rng.standard_normal((6, nt)).astype(np.float32)
    data = np.zeros((elem.size, nt), np.float32)
    for r in range(elem.size):
        d = int(dly[r])
        data[r, d:] = w[elem[r], :nt - d]
    out = t.combineSensorData(data)
    span = nt - int(dly.max())
    assert np.allclose(out[:, :span], w[:, :span], atol=1e-6)


def test_scan_line_coherent_sum():
    """Element traces carrying the transmit delays sum coherently: the line equals
    sum_e apod_e * w[t] where every element's shifted copy is defined."""
    t = make()
    n, nt = 6, 200
    w = np.hanning(50).astype(np.float32)
    w = np.concatenate([np.zeros(70, np.float32), w, np.zeros(nt - 120, np.float32)])
    delays = -t.beamformingDelays
    data = np.zeros((n, nt), np.float32)
    for e in range(n):
        data[e] = np.roll(w, delays[e])
    line = t.scanLine(data)
    apod = t.getReceiveApodization().astype(np.float32)
    assert np.allclose(line, apod.sum() * w, atol=1e-5)


def test_invalid_geometry_raises():
    with pytest.raises(ValueError):
        KWaveTransducerGeometry(numberElements=0, dy=1e-4, dz=1e-4)
    geom = KWaveTransducerGeometry(numberElements=20, elementWidth=4, dy=1e-4, dz=1e-4)
    with pytest.raises(ValueError):
        KWaveTransducer(geom, [12, 64, 16], 2e-8)
