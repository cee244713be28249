"""Pins for the CPU float64 oracle (oracle/kwave_oracle.py) against closed forms, known answers
from the reference's own tests, invariants and physics -- none retypes the oracle's formulas."""
import math

import numpy as np
import pytest

from oracle import kwave_oracle as O
from kwavehip import (AbsorptionMode, KWaveGrid, KWaveMedium, KWaveSensor, KWaveSource, RecordField,
                      SimulationOptions, SourceMode, makeCircle, makeDisc, toneBurst, db2neper, neper2db)


def run(grid, medium, source, sensor, **opt):
    return O.kspace_first_order(grid, medium, source, sensor, SimulationOptions(**opt))


# ---------------------------------------------------------------------------------- grid / time
def test_fft_freq_matches_numpy():
    for n in (1, 2, 7, 8, 64, 96, 129):
        assert np.allclose(O.fft_freq(n, 1e-4), np.fft.fftfreq(n, 1e-4), rtol=0, atol=1e-9)


def test_make_time_known_answers():
    # GridTests.swift:18-24
    dt, nt = O.make_time([16, 16], [1e-3, 1e-3], 1500)
    assert abs(dt - 0.3 * 1e-3 / 1500) < 1e-12 and nt > 0
    # ParityTests.swift:63: 64^2, dx 1e-4, c 1500, default tEnd -> nt == 302
    dt, nt = O.make_time([64, 64], [1e-4, 1e-4], 1500)
    assert nt == 302 and abs(dt - 2e-8) < 1e-20
    # 128^2 IVP, tEnd 4e-6 -> dt 2e-8, nt 201
    dt, nt = O.make_time([128, 128], [1e-4, 1e-4], 1500, t_end=4e-6)
    assert nt == 201
    g = KWaveGrid(64, 1e-4, 64, 1e-4)
    g.makeTime(1500)
    assert g.nt == 302


# ---------------------------------------------------------------------------------- operators
@pytest.mark.parametrize("n,m", [(64, 3), (64, 31), (45, 7)])
def test_staggered_derivative_plane_wave(n, m):
    """Re ifft(ddxPos fft sin(k x)) = k cos(k (x + dx/2)) exactly for a resolvable mode
    (extends FFTTests.swift:16-30 to the staggered operators)."""
    dx = 1e-3
    x = np.arange(n) * dx
    k = 2 * math.pi * m / (n * dx)
    kv = O.wavenumber_vector(n, dx)
    f = np.sin(k * x)
    gp = np.real(np.fft.ifft(O.derivative_operator(kv, dx, +1) * np.fft.fft(f)))
    gn = np.real(np.fft.ifft(O.derivative_operator(kv, dx, -1) * np.fft.fft(f)))
    assert np.allclose(gp, k * np.cos(k * (x + dx / 2)), atol=1e-9 * k)
    assert np.allclose(gn, k * np.cos(k * (x - dx / 2)), atol=1e-9 * k)
    c = np.real(np.fft.ifft(O.collocation_op(kv, dx) * np.fft.fft(f)))
    assert np.allclose(c, np.sin(k * (x - dx / 2)), atol=1e-12)
    # ddxNeg * ddxPos = -k^2 on every bin
    prod = O.derivative_operator(kv, dx, -1) * O.derivative_operator(kv, dx, +1)
    assert np.allclose(prod, -kv ** 2, atol=1e-9 * kv.max() ** 2)


def test_pml_profile_positions():
    n, m, dx, dt, c, a = 40, 10, 1e-4, 2e-8, 1500.0, 2.0
    reg = O.pml_profile(n, dx, dt, c, m, a, False)
    stg = O.pml_profile(n, dx, dt, c, m, a, True)
    coef = a * c / dx * dt / 2
    assert np.all(reg[m:n - m] == 1.0)
    # outermost regular cell: position 1 on both sides
    assert math.isclose(reg[0], math.exp(-coef), rel_tol=1e-14)
    assert math.isclose(reg[-1], math.exp(-coef), rel_tol=1e-14)
    # innermost regular cells sit at 1/m
    assert math.isclose(reg[m - 1], math.exp(-coef * (1 / m) ** 4), rel_tol=1e-14)
    # staggered: left outermost at (m - 0.5)/m, right outermost at (m + 0.5)/m
    assert math.isclose(stg[0], math.exp(-coef * ((m - 0.5) / m) ** 4), rel_tol=1e-14)
    assert math.isclose(stg[-1], math.exp(-coef * ((m + 0.5) / m) ** 4), rel_tol=1e-14)
    assert np.all(np.diff(reg[:m]) > 0)


def test_stagger_density():
    r = np.arange(12.0).reshape(3, 4)
    s0 = O.stagger_density(r, 0)
    assert np.allclose(s0[:-1], 0.5 * (r[:-1] + r[1:])) and np.allclose(s0[-1], r[-1])
    s1 = O.stagger_density(r, 1)
    assert np.allclose(s1[:, :-1], 0.5 * (r[:, :-1] + r[:, 1:])) and np.allclose(s1[:, -1], r[:, -1])


def test_db2neper_roundtrip():
    # UtilTests.swift:139-143
    assert math.isclose(neper2db(db2neper(0.75, 1.5), 1.5), 0.75, rel_tol=1e-12)
    assert math.isclose(O.db2neper(0.75, 1.5), db2neper(0.75, 1.5), rel_tol=1e-15)


# ---------------------------------------------------------------------------------- smoothing
@pytest.mark.parametrize("sizes", [[64], [64, 48], [24, 20, 16]])
def test_smoothing_window_properties(sizes):
    w = O.smoothing_window(sizes)
    assert w.shape == tuple(sizes)
    assert math.isclose(w.reshape(-1)[0], 1.0, rel_tol=1e-12)       # DC gain 1 (Blackman centre)
    assert np.all(w >= 0) and np.all(w <= 1 + 1e-12)
    # even sizes: the placed window is even (w(-k) == w(k))
    mirror = w
    for a, n in enumerate(sizes):
        mirror = np.take(mirror, (-np.arange(n)) % n, axis=a)
    assert np.allclose(mirror, w, atol=1e-15)


def test_smooth_restores_max():
    f = makeDisc(64, 64, radius=5).astype(np.float64)
    s = O.smooth(f)
    assert math.isclose(np.max(np.abs(s)), 1.0, rel_tol=1e-12)
    assert np.allclose(s, s[:, ::-1][:, np.r_[-1, :63]], atol=1e-12)  # circular-reflection symmetry


# ---------------------------------------------------------------------------------- whole loop
@pytest.mark.parametrize("dim", [1, 2, 3])
def test_lossless_loop_matches_closed_form(dim):
    """Homogeneous, lossless, PML-free (pmlAlpha = 0) p0 problem: the k-space corrected leapfrog
    is exact, p^(k, n dt) = p^0(k) cos(c |k| n dt)."""
    n = {1: 64, 2: 32, 3: 16}[dim]
    dx, c = 1e-4, 1500.0
    rng = np.random.default_rng(5)
    p0 = rng.standard_normal((n,) * dim).astype(np.float32)
    grid = KWaveGrid(*sum(([n, dx] for _ in range(dim)), []))
    grid.setTime(40, 0.3 * dx / c)
    mask = np.ones((n,) * dim, dtype=np.float32)
    out = run(grid, KWaveMedium(c, 1000.0), KWaveSource(p0=p0), KWaveSensor(mask=mask, record=RecordField.p),
              pmlAlpha=0.0, smoothP0=False, pmlSize=4)
    ref = O.closed_form_lossless(p0.astype(np.float64), [n] * dim, [dx] * dim, c, grid.dt, range(grid.nt))
    ser = out["p"]
    for t in (0, 1, 17, 39):
        assert np.allclose(ser[:, t], ref[t].reshape(-1), atol=1e-10 * np.abs(p0).max())
    assert np.allclose(out["pFinal"], ref[-1], atol=1e-10)


def test_disc_ivp_symmetric():
    """SolverTests.swift:78-103: 4-fold circular-reflection symmetry of a centred disc IVP."""
    n = 64
    grid = KWaveGrid(n, 1e-4, n, 1e-4)
    grid.makeTime(1500, cfl=0.3, tEnd=3e-7)
    for alpha, tol in ((2.0, 1e-3), (0.0, 1e-12)):   # reference tolerance; exact without the PML
        out = run(grid, KWaveMedium(1500.0, 1000.0), KWaveSource(p0=makeDisc(n, n, radius=5)), KWaveSensor(),
                  pmlAlpha=alpha)
        pf = out["pFinal"]
        cx = np.roll(pf[::-1, :], 1, axis=0)
        cy = np.roll(pf[:, ::-1], 1, axis=1)
        m = np.abs(pf).max()
        assert m > 0 and np.abs(pf - cx).max() / m < tol and np.abs(pf - cy).max() / m < tol


def _mode_run(alpha_mode, alpha, y=1.5, steps=800, m=16, n=128, dx=1e-4, c=1500.0):
    x = np.arange(n)
    p0 = np.cos(2 * math.pi * m * x / n).astype(np.float32)
    grid = KWaveGrid(n, dx)
    grid.setTime(steps, 0.1 * dx / c)
    med = KWaveMedium(c, 1000.0, alphaCoeff=alpha, alphaPower=y, alphaMode=alpha_mode)
    sens = np.zeros(n, dtype=np.float32)
    sens[0] = 1
    out = run(grid, med, KWaveSource(p0=p0), KWaveSensor(mask=sens, record=RecordField.p),
              pmlAlpha=0.0, smoothP0=False, pmlSize=2)
    return out["p"][0], grid.dt, 2 * math.pi * m / (n * dx)


def test_absorption_decay_rate_matches_power_law():
    """A standing mode decays at alpha_Np * omega^y * c (the power-law model's temporal rate)."""
    alpha_db, y, c = 5.0, 1.5, 1500.0
    s, dt, k = _mode_run(AbsorptionMode.noDispersion, alpha_db, y)
    w = c * k
    t = np.arange(s.size) * dt
    # envelope from the analytic signal peaks
    peaks = [i for i in range(1, s.size - 1) if s[i] > s[i - 1] and s[i] >= s[i + 1] and s[i] > 0]
    tp, ap = t[peaks], s[peaks]
    gamma = -np.polyfit(tp, np.log(ap), 1)[0]
    expect = db2neper(alpha_db, y) * w ** y * c
    assert abs(gamma / expect - 1) < 0.03, (gamma, expect)


def test_dispersion_shifts_frequency_and_alpha_zero_limit():
    alpha_db, y = 5.0, 1.5
    a, _, _ = _mode_run(AbsorptionMode.powerLaw, alpha_db, y)
    b, _, _ = _mode_run(AbsorptionMode.noDispersion, alpha_db, y)
    lossless, _, _ = _mode_run(AbsorptionMode.noAbsorption, alpha_db, y)
    z, _, _ = _mode_run(AbsorptionMode.powerLaw, 0.0, y)
    assert np.allclose(z, lossless, atol=1e-12)
    # tan(pi y / 2) < 0 for 1 < y < 2: dispersion raises the phase speed, so zero crossings come
    # earlier with dispersion than without
    zc = lambda s: np.flatnonzero(np.diff(np.sign(s)) != 0)
    assert zc(a)[-1] <= zc(b)[-1]
    assert not np.allclose(a, b)


def test_stokes_equals_power_law_y2_without_dispersion():
    s1, _, _ = _mode_run(AbsorptionMode.stokes, 3.0, 1.5)
    s2, _, _ = _mode_run(AbsorptionMode.noDispersion, 3.0, 2.0)
    assert np.allclose(s1, s2, atol=1e-10)


def test_nonlinearity_quadratic_scaling():
    """B/A enters quadratically: N(2A) = 4 N(A) to leading order, N(A) = p_nl(A) - p_lin(A)."""
    n = 64
    grid = KWaveGrid(n, 1e-4)
    grid.setTime(60, 2e-8)
    x = np.arange(n)
    sens = np.zeros(n, dtype=np.float32)
    sens[n // 2 + 5] = 1
    res = {}
    for amp in (1e5, 2e5):
        p0 = (amp * np.exp(-((x - n / 2) / 3.0) ** 2)).astype(np.float32)
        for b in (0.0, 6.0):
            o = run(grid, KWaveMedium(1500.0, 1000.0, bOnA=b), KWaveSource(p0=p0),
                    KWaveSensor(mask=sens, record=RecordField.p), smoothP0=False, pmlSize=10)
            res[(amp, b)] = o["p"][0]
    n1 = res[(1e5, 6.0)] - res[(1e5, 0.0)]
    n2 = res[(2e5, 6.0)] - res[(2e5, 0.0)]
    assert np.abs(n1).max() > 0
    ratio = np.abs(n2).max() / np.abs(n1).max()
    assert abs(ratio - 4.0) < 0.05


# ---------------------------------------------------------------------------------- sources
def test_dirichlet_source_sets_pressure():
    """A dirichlet p source pins p at its points: p = c^2 * sum_a v/(dim c^2) = v (linear, lossless)."""
    n = 48
    grid = KWaveGrid(n, 1e-4, n, 1e-4)
    grid.setTime(30, 2e-8)
    mask = np.zeros((n, n), dtype=np.float32)
    mask[20, 24] = 1
    sig = toneBurst(1 / grid.dt, 2e6, 2).astype(np.float32)
    out = run(grid, KWaveMedium(1500.0, 1000.0),
              KWaveSource(pMask=mask, p=sig, pMode=SourceMode.dirichlet),
              KWaveSensor(mask=mask, record=RecordField.p), pmlSize=10)
    m = min(sig.size, grid.nt)
    assert np.allclose(out["p"][0, :m], sig[:m], atol=1e-9 * np.abs(sig).max())


@pytest.mark.parametrize("mode", [SourceMode.additive, SourceMode.additiveNoCorrection])
def test_additive_source_linearity(mode):
    n = 48
    grid = KWaveGrid(n, 1e-4, n, 1e-4)
    grid.setTime(40, 2e-8)
    mask = np.zeros((n, n), dtype=np.float32)
    mask[24, 20:28] = 1
    sens = np.zeros((n, n), dtype=np.float32)
    sens[10, :] = 1
    sig = toneBurst(1 / grid.dt, 1e6, 2).astype(np.float32)
    o1 = run(grid, KWaveMedium(1500.0, 1000.0), KWaveSource(pMask=mask, p=sig, pMode=mode),
             KWaveSensor(mask=sens), pmlSize=10)
    o2 = run(grid, KWaveMedium(1500.0, 1000.0), KWaveSource(pMask=mask, p=2 * sig, pMode=mode),
             KWaveSensor(mask=sens), pmlSize=10)
    assert np.abs(o1["p"]).max() > 0
    assert np.allclose(o2["p"], 2 * o1["p"], atol=1e-12 * np.abs(o1["p"]).max())


def test_source_kappa_correction_preserves_injected_mass():
    """source kappa = cos(c_ref dt |k| / 2) is 1 at k = 0: the corrected field sums to the injected sum."""
    shape = (32, 32)
    idx = np.array([5 * 32 + 7, 20 * 32 + 3])
    op = O.SourceOp(idx, np.array([[1.5, 2.0]]), np.array([1.0, 3.0]), "additive", shape)
    kmag = O.k_magnitude(list(shape), [1e-4, 1e-4])
    sk = np.cos(kmag * 1500 * 2e-8 / 2)
    out = op.apply(np.zeros(shape), 0, sk, O._NumpyFFT())
    assert math.isclose(out.sum(), 1.5 * 1.0 + 1.5 * 3.0, rel_tol=1e-12)


# ---------------------------------------------------------------------------------- recording
def test_intensity_half_sample_shift_closed_form():
    """u recorded half a sample behind p; after the +dt/2 Fourier shift an in-phase pair gives
    mean(cos^2) = 1/2 exactly (Recording.swift:113-124)."""
    nt, f = 64, 5
    n = np.arange(nt)
    p = np.cos(2 * math.pi * f * n / nt).reshape(1, -1)
    u = np.cos(2 * math.pi * f * (n - 0.5) / nt).reshape(1, -1)
    assert math.isclose(O.intensity_avg(p, u)[0], 0.5, rel_tol=1e-12)


def test_aggregates_and_record_plan():
    n = 48
    grid = KWaveGrid(n, 1e-4, n, 1e-4)
    grid.makeTime(1500, cfl=0.3, tEnd=1.2e-6)
    sens = np.zeros((n, n), dtype=np.float32)
    sens[8, :] = 1
    rec = RecordField.p | RecordField.pMax | RecordField.pMin | RecordField.pRms | RecordField.ux | RecordField.uMax
    out = run(grid, KWaveMedium(1500.0, 1000.0), KWaveSource(p0=makeDisc(n, n, radius=4)),
              KWaveSensor(mask=sens, record=rec))
    p = out["p"]
    assert p.shape == (n, grid.nt)
    assert np.all(out["pMax"] >= out["pRms"]) and np.all(out["pRms"] >= 0)
    assert np.allclose(out["pMax"], p.max(1)) and np.allclose(out["pMin"], p.min(1))
    assert "uy" not in out and "ux" in out and "uxMax" in out and "uyMax" in out


def test_cartesian_sampler_on_grid_nodes_equals_binary():
    n, dx = 32, 1e-4
    pts = np.array([[(-3) * dx, 5 * dx], [2 * dx, 0.0]], dtype=np.float32)   # x row, y row
    s = O.make_sampler(pts, [n, n], [dx, dx])
    f = np.random.default_rng(1).standard_normal((n, n))
    got = s.sample(f)
    want = [f[n // 2 - 3, n // 2 + 2], f[n // 2 + 5, n // 2]]
    assert np.allclose(got, want, atol=1e-6)
    mid = O.make_sampler(np.array([[0.5 * dx], [0.0]], dtype=np.float32), [n, n], [dx, dx])
    assert math.isclose(mid.sample(f)[0], 0.5 * (f[n // 2, n // 2] + f[n // 2 + 1, n // 2]), rel_tol=1e-6)


def test_cartesian_outside_raises():
    with pytest.raises(ValueError):
        O.make_sampler(np.array([[1.0], [0.0]], dtype=np.float32), [16, 16], [1e-4, 1e-4])


def test_directivity_isotropic_limit():
    """Pressure pattern with a vanishing element size is the identity filter: equals plain sampling."""
    n = 48
    grid = KWaveGrid(n, 1e-4, n, 1e-4)
    grid.setTime(50, 2e-8)
    sens = makeCircle(n, n, radius=15)
    ang = np.zeros((n, n), dtype=np.float32)
    ang[sens > 0] = np.linspace(0, 3, int(sens.sum()), dtype=np.float32)
    src = KWaveSource(p0=makeDisc(n, n, radius=4))
    plain = run(grid, KWaveMedium(1500.0, 1000.0), src, KWaveSensor(mask=sens))
    direc = run(grid, KWaveMedium(1500.0, 1000.0), src,
                KWaveSensor(mask=sens, directivityAngle=ang, directivitySize=1e-12))
    assert np.allclose(direc["p"], plain["p"], atol=1e-10)


def test_make_circle_known_answer():
    """UtilTests.swift:8-16."""
    c = makeCircle(11, 11, radius=3)
    assert int(c.sum()) == 16
    for x, y in [(2, 4), (3, 3), (4, 2), (5, 8), (8, 6)]:
        assert c[x, y] == 1


def test_chunked_run_equals_full_run():
    n = 32
    grid = KWaveGrid(n, 1e-4, n, 1e-4)
    grid.setTime(30, 2e-8)
    src = KWaveSource(p0=makeDisc(n, n, radius=4))
    full = O.kspace_first_order(grid, KWaveMedium(1500.0, 1000.0), src, KWaveSensor(), SimulationOptions())
    a = O.kspace_first_order(grid, KWaveMedium(1500.0, 1000.0), src, KWaveSensor(), SimulationOptions(),
                             t_end=12, return_state=True)
    b = O.kspace_first_order(grid, KWaveMedium(1500.0, 1000.0), src, KWaveSensor(), SimulationOptions(),
                             t_begin=12, state=a["_state"])
    assert np.array_equal(full["pFinal"], b["pFinal"])


def test_float32_mode_stays_single_precision():
    # the timing-only float32 mode: every state array is float32 and the result is the float64
    # run to single-precision rounding (a stray float64 operator would not change the result,
    # a dropped cast of the operators would leave the state in float64)
    n = 32
    grid = KWaveGrid(n, 1e-4, n, 1e-4)
    grid.setTime(40, 2e-8)
    med = KWaveMedium(1500.0, 1000.0, alphaCoeff=0.5, alphaPower=1.1, alphaMode=AbsorptionMode.powerLaw)
    src = KWaveSource(p0=makeDisc(n, n, radius=4))
    sens = KWaveSensor(mask=makeCircle(n, n, radius=10), record=RecordField.p | RecordField.ux)
    r64 = O.kspace_first_order(grid, med, src, sens, SimulationOptions(pmlSize=6), return_state=True)
    r32 = O.kspace_first_order(grid, med, src, sens, SimulationOptions(pmlSize=6), dtype=np.float32,
                               return_state=True)
    st = r32["_state"]
    assert st["p"].dtype == np.float32
    assert all(x.dtype == np.float32 for x in st["u"] + st["rho"])
    for k in ("p", "ux"):
        err = np.abs(r32[k] - r64[k]).max() / np.abs(r64[k]).max()
        assert 1e-9 < err < 1e-5, (k, err)


def test_streamed_aggregates_and_several_sensors_equal_separate_runs():
    """Aggregates folded step by step (no series stored) equal those of the stored series, and one
    run recording several sensors equals one run per sensor."""
    n = 40
    grid = KWaveGrid(n, 1e-4, n, 1e-4)
    grid.setTime(50, 2e-8)
    med = KWaveMedium(1500.0, 1000.0, alphaCoeff=0.5, alphaPower=1.5, alphaMode=AbsorptionMode.powerLaw)
    src = KWaveSource(p0=makeDisc(n, n, radius=4))
    full = np.ones((n, n), np.float32)
    line = np.zeros((n, n), np.float32)
    line[6, 4:30] = 1
    agg = RecordField.pMax | RecordField.pMin | RecordField.pRms | RecordField.uMax | RecordField.uRms
    s_agg = KWaveSensor(mask=full, record=agg)
    s_ser = KWaveSensor(mask=full, record=agg | RecordField.p | RecordField.ux | RecordField.uy)
    s_line = KWaveSensor(mask=line, record=RecordField.p)
    both = O.kspace_first_order(grid, med, src, [s_agg, s_line], SimulationOptions(pmlSize=8))
    ser = O.kspace_first_order(grid, med, src, s_ser, SimulationOptions(pmlSize=8))
    one = O.kspace_first_order(grid, med, src, s_line, SimulationOptions(pmlSize=8))
    assert "p" not in both[0]
    for k in ("pMax", "pMin", "uxMax", "uyMax"):
        np.testing.assert_array_equal(both[0][k], ser[k])
    for k in ("pRms", "uxRms", "uyRms"):
        np.testing.assert_allclose(both[0][k], ser[k], rtol=1e-13)
    np.testing.assert_array_equal(both[1]["p"], one["p"])
    np.testing.assert_array_equal(both[1]["pFinal"], one["pFinal"])
