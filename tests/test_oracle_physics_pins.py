"""Physics pins for three parts of the float64 oracle that the reference's own tests leave open:

- the Blackman smoothing profile (Filters.swift:203-347) against numpy.blackman / numpy.interp and
  exact radial distances (3-4-5 and 2-3-6 triangles);
- the heterogeneous-medium coupling (c0^2 in the equation of state, dt / rho0 on the staggered
  grid; KSpaceFirstOrder.swift:610-614, :799-806) against the normal-incidence reflection and
  transmission coefficients of an impedance step, (Z2 - Z1)/(Z2 + Z1) and 2 Z2/(Z2 + Z1);
- the B/A nonlinearity (KSpaceFirstOrder.swift:889-907) against the Fubini solution of a
  progressive plane wave, second harmonic A J2(2 sigma)/sigma with sigma = beta omega A t /
  (rho0 c0^2) and beta = 1 + B/(2A).

Each pin is followed by planted mistakes (monkeypatched into the oracle on the CPU) that the pin
must reject, so a dropped factor, a wrong power or a wrong window cannot pass it.
"""
import math

import numpy as np
import pytest
from scipy.special import jv

from oracle import kwave_oracle as O
from kwavehip import KWaveGrid, KWaveMedium, KWaveSensor, KWaveSource, RecordField, SimulationOptions


# ------------------------------------------------------------------------------------ smoothing
def _centred(sizes):
    """The oracle's window moved back from FFT order to peak-at-centre order.

    1D / 3D place w[i] at (i + (n+1)//2) % n; 2D reads w[(i + (n+1)//2) % n] (Filters.swift:288,
    312-314, 331-337), so the inverse roll differs by direction."""
    placed = O.smoothing_window(list(sizes))
    shift = [(n + 1) // 2 for n in sizes]
    if len(sizes) == 2:
        return np.roll(placed, shift, axis=(0, 1))
    return np.roll(placed, [-s for s in shift], axis=tuple(range(len(sizes))))


def _smoothing_errors():
    """Largest deviation of the oracle's smoothing from the known answers, per check."""
    err = {}
    # symmetric Blackman of every small length
    err["blackman"] = max(float(np.max(np.abs(O._blackman(n) - np.blackman(n)), initial=0.0))
                          for n in range(0, 70))
    # 1D: even sizes take the first n samples of the enlarged (n+1)-point window
    e = 0.0
    for n in (16, 17, 32, 33, 64, 65):
        want = np.fft.ifftshift(np.blackman(n if n % 2 else n + 1)[:n])
        e = max(e, float(np.max(np.abs(O.smoothing_window([n]) - want))))
    err["1d"] = e
    # 2D odd sizes: the profile through the centre is the Blackman window of each axis length,
    # whatever the other axis (the shorter axis is stretched onto the same radius)
    e = 0.0
    for nx, ny in ((33, 17), (17, 33), (33, 33), (65, 9)):
        w = _centred((nx, ny))
        e = max(e, float(np.max(np.abs(w[nx // 2, :] - np.blackman(ny)))),
                float(np.max(np.abs(w[:, ny // 2] - np.blackman(nx)))))
    err["2d_axes"] = e
    # exact integer radii: 3-4-5, 6-8-10, 5-12-13 in 2D; 1-2-2, 2-3-6, 4-4-7 in 3D
    b33 = np.blackman(33)
    w2 = _centred((33, 33))
    pts2 = [((3, 4), 5), ((4, 3), 5), ((6, 8), 10), ((-5, 12), 13), ((12, -5), 13)]
    e = max(abs(w2[16 + a, 16 + b] - b33[16 + r]) for (a, b), r in pts2)
    w3 = _centred((33, 33, 33))
    pts3 = [((1, 2, 2), 3), ((2, -3, 6), 7), ((-4, 4, 7), 9), ((0, 0, 16), 16)]
    e = max(e, max(abs(w3[16 + a, 16 + b, 16 + c] - b33[16 + r]) for (a, b, c), r in pts3))
    # beyond the radius the window is the (zero) end sample: the corners
    e = max(e, abs(w2[0, 0]), abs(w3[0, 0, 0]))
    err["radii"] = float(e)
    # a stretched axis off the unit grid: linear interpolation of the 33-point window (numpy.interp)
    w = _centred((33, 12))                       # 12 -> enlarged 13 samples over radius 16
    ys = -16 + 32.0 * np.arange(12) / 12
    want = np.abs(np.interp(np.abs(ys) + 16, np.arange(33), b33))
    err["interp"] = float(np.max(np.abs(w[16, :] - want)))
    # the whole smooth(): two cosine modes keep their amplitude ratio w(m2)/w(m1), peak restored
    n = 65
    x = np.arange(n)
    f = np.cos(2 * np.pi * 3 * x / n) + 0.5 * np.cos(2 * np.pi * 11 * x / n)
    s = O.smooth(f)
    spec = np.abs(np.fft.fft(s))
    b65 = np.blackman(65)
    ratio = (spec[11] / spec[3]) / (0.5 * b65[32 + 11] / b65[32 + 3])
    err["smooth"] = max(abs(ratio - 1.0), abs(np.max(np.abs(s)) - np.max(np.abs(f))))
    return err


def test_smoothing_profile_known_answers():
    err = _smoothing_errors()
    assert all(v < 1e-12 for v in err.values()), err


def _hann(n):
    return np.hanning(n) if n > 1 else np.ones(max(n, 0))


def _blackman_n_slip(n):
    if n <= 1:
        return np.ones(max(n, 0))
    x = 2.0 * math.pi * np.arange(n) / n
    return 0.42 - 0.5 * np.cos(x) + 0.08 * np.cos(2 * x)


def _nearest_lower(r, win, radius):
    pos = np.minimum(r, radius) + radius
    return np.abs(win[np.minimum(np.floor(pos).astype(np.int64), len(win) - 1)])


def _unshifted_window(sizes, _orig=O.smoothing_window):
    w = _orig(sizes)
    return np.roll(w, [-(n + 1) // 2 for n in sizes], axis=tuple(range(len(sizes))))


@pytest.mark.parametrize("name,planted", [("_blackman", _hann), ("_blackman", _blackman_n_slip),
                                          ("_interp_radial", _nearest_lower),
                                          ("smoothing_window", _unshifted_window)])
def test_planted_smoothing_mistakes_fail(monkeypatch, name, planted):
    monkeypatch.setattr(O, name, planted)
    err = _smoothing_errors()
    if name == "_blackman":
        err.pop("blackman")            # the direct comparison fails trivially; the others must too
    assert max(err.values()) > 1e-3, err


# ---------------------------------------------------------------- progressive-wave initial states
def _run_from_state(shape, spacing, c, rho, dt, steps, state, record_mask, bona=None):
    """Run the oracle loop from an explicit state (p^0, u^{-1/2}, split rho^0) for `steps` steps,
    starting at loop index 1 (no t = 0 override), PML-free and periodic (pml_alpha = 0)."""
    dim = len(shape)
    grid = KWaveGrid(*sum(([n, spacing] for n in shape), []))
    grid.setTime(steps + 1, dt)
    med = KWaveMedium(np.asarray(c, np.float32), np.asarray(rho, np.float32), bOnA=bona)
    src = KWaveSource(p0=np.zeros(shape, np.float32))
    sens = KWaveSensor(mask=record_mask.astype(np.float32), record=RecordField.p)
    opts = SimulationOptions(pmlSize=2, pmlAlpha=0.0, smoothP0=False)
    assert dim == len(state["u"])
    return O.kspace_first_order(grid, med, src, sens, opts, t_begin=1, state=state)


def _travelling_state(f, axis, shape, dx, dt, c, rho, rho_split_eos=lambda p: p):
    """State of a wave f(s - c t) moving towards +axis: p = f(s), the split density along `axis`
    carries all of rho (the other splits are 0), u_axis = p/(rho c) at s + dx/2, t = -dt/2."""
    dim = len(shape)
    s = np.arange(shape[axis]) * dx
    bshape = [1] * dim
    bshape[axis] = shape[axis]
    p = np.broadcast_to(f(s).reshape(bshape), shape).astype(np.float64).copy()
    u_ax = np.broadcast_to((f(s + dx / 2 + c * dt / 2) / (rho * c)).reshape(bshape), shape).copy()
    u = [np.zeros(shape) for _ in range(dim)]
    u[axis] = u_ax
    rh = [np.zeros(shape) for _ in range(dim)]
    rh[axis] = rho_split_eos(p) / (c * c)
    return {"p": p, "u": u, "rho": rh}


# --------------------------------------------------------------------- heterogeneous coupling
N_STEP = 1024          # cells along the propagation axis; interface at N_STEP/2, pulse at N_STEP/4


def _impedance_step(c1, c2, rho1, rho2, axis=0, dim=1):
    """Reflected and transmitted peak amplitudes of a unit Gaussian pulse hitting an impedance step
    at normal incidence, measured once the reflection is clear of the interface."""
    dx = 1e-4
    shape = [N_STEP if a == axis else 4 for a in range(dim)]
    s_idx = np.arange(N_STEP)
    prof = lambda v1, v2: np.where(s_idx < N_STEP // 2, v1, v2)
    bshape = [1] * dim
    bshape[axis] = N_STEP
    c = np.broadcast_to(prof(c1, c2).reshape(bshape), shape).astype(np.float32)
    rho = np.broadcast_to(prof(rho1, rho2).reshape(bshape), shape).astype(np.float32)
    dt = 0.3 * dx / max(c1, c2)
    s0, width = N_STEP // 4 * dx, 10 * dx
    f = lambda s: np.exp(-0.5 * ((s - s0) / width) ** 2)
    state = _travelling_state(f, axis, shape, dx, dt, c1, rho1)
    # the pulse reaches the interface after N/4 cells at c1; stop when it has run N/6 further in
    # medium 1 (reflection at N/2 - N/6 > N/4 + 5 widths of the incident tail)
    steps = int(round((N_STEP / 4 + N_STEP / 6) * dx / c1 / dt))
    mask = np.zeros(shape, bool)
    mask[tuple(slice(None) if a == axis else 0 for a in range(dim))] = True
    out = _run_from_state(shape, dx, c, rho, dt, steps, state, mask)
    line = out["p"][:, -1]
    refl = line[: N_STEP // 2 - 8]
    trans = line[N_STEP // 2 + 8:]
    r_meas = refl[np.argmax(np.abs(refl))]
    t_meas = trans[np.argmax(np.abs(trans))]
    # the transmitted pulse's centre travelled at c2 since it crossed the interface
    t_cross = (N_STEP / 4) * dx / c1
    centre = N_STEP // 2 + 8 + int(np.argmax(np.abs(trans)))
    travelled = (centre - N_STEP // 2) * dx
    c_meas = travelled / (steps * dt - t_cross)
    return float(r_meas), float(t_meas), float(c_meas)


STEP_CASES = [  # (c1, c2, rho1, rho2): a sound-speed step, a density step, both, and a drop
    (1500.0, 2000.0, 1000.0, 1000.0),
    (1500.0, 1500.0, 1000.0, 1600.0),
    (1500.0, 1800.0, 1000.0, 1300.0),
    (1800.0, 1400.0, 1200.0, 1000.0),
]


def _coeffs(c1, c2, rho1, rho2):
    z1, z2 = rho1 * c1, rho2 * c2
    return (z2 - z1) / (z2 + z1), 2 * z2 / (z2 + z1)


@pytest.mark.parametrize("case", STEP_CASES)
def test_impedance_step_reflection_transmission(case):
    r, t, c_meas = _impedance_step(*case)
    r_want, t_want = _coeffs(*case)
    assert abs(r - r_want) < 2e-3 and abs(t - t_want) < 2e-3, (r, r_want, t, t_want)
    assert abs(c_meas / case[1] - 1) < 0.02, (c_meas, case[1])


@pytest.mark.parametrize("dim,axis", [(2, 1), (3, 2), (3, 0)])
def test_impedance_step_along_other_axes(dim, axis):
    case = STEP_CASES[2]
    r, t, _ = _impedance_step(*case, axis=axis, dim=dim)
    r_want, t_want = _coeffs(*case)
    assert abs(r - r_want) < 2e-3 and abs(t - t_want) < 2e-3, (r, r_want, t, t_want)


@pytest.mark.parametrize("name,planted", [
    ("eos_sound_speed_squared", lambda c0: c0 * 1500.0),                # c instead of c^2
    ("dt_over_staggered_density", lambda dt, rho0, axis: dt * O.stagger_density(rho0, axis) / 1e6),
    ("dt_over_staggered_density", lambda dt, rho0, axis: dt / np.mean(rho0)),   # contrast dropped
])
def test_planted_coupling_mistakes_fail(monkeypatch, name, planted):
    monkeypatch.setattr(O, name, planted)
    worst = 0.0
    for case in STEP_CASES[:2]:
        r, t, _ = _impedance_step(*case)
        r_want, t_want = _coeffs(*case)
        worst = max(worst, abs(r - r_want), abs(t - t_want))
    assert worst > 1e-2, worst


# --------------------------------------------------------------------------------- nonlinearity
def _fubini_run(bona, amp=2e5, m=8, n=256, sigma_end=0.15, beta_for_time=4.0):
    """A progressive cosine wave of amplitude amp in a periodic lossless medium; returns the
    measured harmonic amplitudes against the Fubini solution at a few times."""
    dx, c0, rho0 = 1e-4, 1500.0, 1000.0
    dt = 0.3 * dx / c0
    k = 2 * np.pi * m / (n * dx)
    omega = c0 * k
    steps = int(sigma_end * rho0 * c0 ** 2 / (beta_for_time * omega * amp) / dt)
    f = lambda s: amp * np.cos(k * s)
    # density from the equation of state inverted to second order in p
    eos_inv = lambda p: p - 0.5 * (bona or 0.0) * p * p / (rho0 * c0 * c0)
    state = _travelling_state(f, 0, [n], dx, dt, c0, rho0, eos_inv)
    out = _run_from_state([n], dx, np.float32(c0), np.float32(rho0), dt, steps, state, np.ones(n, bool),
                          bona=np.float32(bona) if bona is not None else None)
    series = out["p"]                   # [n, steps], column t = after loop index t + 1
    beta = 1.0 + 0.5 * (bona or 0.0)
    rows = []
    for j in (steps // 3, 2 * steps // 3, steps - 1):
        t = (j + 1) * dt
        spec = np.abs(np.fft.rfft(series[:, j])) * 2 / n
        sigma = beta * omega * amp * t / (rho0 * c0 ** 2)
        rows.append((spec[m], spec[2 * m], amp * 2 * jv(1, sigma) / sigma, amp * jv(2, 2 * sigma) / sigma))
    return rows


@pytest.mark.parametrize("bona", [1e-9, 2.0, 6.0])
def test_second_harmonic_growth_matches_fubini(bona):
    """beta = 1 + B/2A: from the mass-term factor alone (B/A -> 0, still enabled) and with the
    equation-of-state term (B/A = 2, 6)."""
    for a1, a2, f1, f2 in _fubini_run(bona):
        assert abs(a2 / f2 - 1) < 0.01, (bona, a2, f2)
        assert abs(a1 / f1 - 1) < 1e-3, (bona, a1, f1)


def test_linear_medium_generates_no_harmonic():
    for _, a2, _, _ in _fubini_run(None):
        assert a2 < 1e-9 * 2e5


@pytest.mark.parametrize("name,planted", [
    ("nonlinear_mass_factor", lambda rp, rho0: np.ones_like(rp)),                   # factor dropped
    ("nonlinear_mass_factor", lambda rp, rho0: (rp + rho0) / rho0),                 # 2 rho -> rho
    ("nonlinear_eos_term", lambda b, r, rho0: b * r * r / rho0),                    # 1/(2 rho0) -> 1/rho0
    ("nonlinear_eos_term", lambda b, r, rho0: 0.0 * r),                              # term dropped
])
def test_planted_nonlinearity_mistakes_fail(monkeypatch, name, planted):
    monkeypatch.setattr(O, name, planted)
    worst = 0.0
    for bona in (1e-9, 6.0):
        for _, a2, _, f2 in _fubini_run(bona):
            worst = max(worst, abs(a2 / f2 - 1))
    assert worst > 0.05, worst
