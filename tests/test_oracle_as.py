"""Pins of the axisymmetric oracle (kspace_first_order_as, KSpaceFirstOrderAS.swift). The
reference's AS parity fixture (Scripts/parity/reference_as.h5) is gitignored and its generator
needs k-wave-python + h5py, so the restatement is pinned by what the mathematics fixes:

  * a radially symmetric initial pressure in 3D Cartesian coordinates (the Cartesian oracle,
    itself pinned by the closed form in test_oracle.py) evolves like the axisymmetric solution
    p(x, r): agreement to discretisation error, with planted symmetry mistakes failing by far;
  * the expansions' symmetries (zeros of WSWA at ny and 3ny, odd / even reflections);
  * the one-sided radial PML profile in closed form;
  * limits: Stokes alpha -> 0 and B/A -> 0 reduce to the lossless linear loop; a dirichlet
    pressure source pins p at its points to the signal (linear, lossless).
"""
import math

import numpy as np
import pytest

from oracle import kwave_oracle as O
from kwavehip import (AbsorptionMode, KWaveGrid, KWaveMedium, KWaveSensor, KWaveSource, SimulationOptions,
                      SourceMode)


def _rl2(a, b):
    d = np.asarray(a, np.float64) - b
    return math.sqrt(np.mean(d * d)) / math.sqrt(np.mean(np.asarray(b, np.float64) ** 2))


def _gaussian_cases(n=64, ny=32, s2=9.0):
    x = (np.arange(n) - n // 2)
    p02 = np.exp(-(x[:, None] ** 2 + np.arange(ny)[None, :] ** 2) / s2).astype(np.float32)
    c = np.arange(n) - n // 2
    p03 = np.exp(-(c[:, None, None] ** 2 + c[None, :, None] ** 2 + c[None, None, :] ** 2) / s2).astype(np.float32)
    return p02, p03


@pytest.fixture(scope="module")
def cartesian_3d():
    """A radially symmetric Gaussian in 3D (64^3, 40 steps, PML off, nothing reaches the edge)."""
    n, steps, dx = 64, 40, 1e-4
    dt = 0.3 * dx / 1500.0
    _, p03 = _gaussian_cases(n)
    g3 = KWaveGrid(n, dx, n, dx, n, dx)
    g3.setTime(steps, dt)
    opts = SimulationOptions(pmlSize=8, pmlAlpha=0.0, smoothP0=False)
    out = O.kspace_first_order(g3, KWaveMedium(1500.0, 1000.0), KWaveSource(p0=p03), KWaveSensor(), opts,
                               workers=8)
    return out["pFinal"], dt, steps


def _as_run(dt, steps, n=64, ny=32):
    p02, _ = _gaussian_cases(n, ny)
    g2 = KWaveGrid(n, 1e-4, ny, 1e-4)
    g2.setTime(steps, dt)
    opts = SimulationOptions(pmlSize=8, pmlAlpha=0.0, smoothP0=False)
    return O.kspace_first_order_as(g2, KWaveMedium(1500.0, 1000.0), KWaveSource(p0=p02), KWaveSensor(), opts)["pFinal"]


def test_axisymmetric_matches_3d_cartesian(cartesian_3d):
    p3, dt, steps = cartesian_3d
    a = _as_run(dt, steps)[:, :24]
    for ref in (p3[:, 32:56, 32], p3[:, 32, 32:56]):     # along y and along z from the axis
        assert _rl2(a, ref) < 1e-2


@pytest.mark.parametrize("name,repl", [("mirror_hsha", "mirror_hahs"), ("mirror_wswa", "mirror_hahs"),
                                       ("mirror_hahs", "mirror_hsha")])
def test_planted_symmetry_mistakes_fail_the_3d_pin(cartesian_3d, monkeypatch, name, repl):
    p3, dt, steps = cartesian_3d
    monkeypatch.setattr(O, name, getattr(O, repl))
    a = _as_run(dt, steps)[:, :24]
    assert _rl2(a, p3[:, 32:56, 32]) > 2e-2   # the correct expansion agrees to 1e-2 (0.4% here)


def test_expansion_symmetries():
    rng = np.random.default_rng(3)
    w = rng.standard_normal((5, 7))
    ny = 7
    e = O.mirror_wswa(w)
    assert e.shape == (5, 28)
    assert np.all(e[:, ny] == 0) and np.all(e[:, 3 * ny] == 0)
    for k in range(1, ny):
        np.testing.assert_array_equal(e[:, ny + k], -e[:, ny - k])      # odd about the point ny
        np.testing.assert_array_equal(e[:, 4 * ny - k], e[:, k])        # even about the axis
    np.testing.assert_array_equal(e[:, 2 * ny:], -e[:, :2 * ny])
    h = O.mirror_hahs(w)
    s = O.mirror_hsha(w)
    np.testing.assert_array_equal(h[:, ny:2 * ny], w[:, ::-1])     # even about the half point ny - 1/2
    np.testing.assert_array_equal(s[:, ny:2 * ny], -w[:, ::-1])    # odd about it
    np.testing.assert_array_equal(h[:, 2 * ny:], -h[:, :2 * ny])
    np.testing.assert_array_equal(s[:, 2 * ny:], -s[:, :2 * ny])


def test_one_sided_radial_pml_profile():
    n, dx, dt, c, m, alpha = 40, 1e-4, 2e-8, 1500.0, 10, 2.0
    for staggered in (False, True):
        p = O.pml_profile(n, dx, dt, c, m, alpha, staggered, axisymmetric=True)
        full = O.pml_profile(n, dx, dt, c, m, alpha, staggered)
        assert np.all(p[:n - m] == 1.0)                      # no ramp next to the axis
        np.testing.assert_array_equal(p[n - m:], full[n - m:])
        k = np.arange(1, m + 1, dtype=np.float64)
        pos = (k + 0.5) / m if staggered else k / m
        np.testing.assert_allclose(p[n - m:], np.exp(-alpha * (c / dx) * dt / 2 * pos ** 4), rtol=1e-14)


def _small_as(medium, source, sensor=None, steps=60, n=48, ny=24, **opt):
    g = KWaveGrid(n, 1e-4, ny, 1e-4)
    g.setTime(steps, 0.3 * 1e-4 / 1500.0)
    opts = SimulationOptions(pmlSize=8, smoothP0=False, **opt)
    return O.kspace_first_order_as(g, medium, source, sensor or KWaveSensor(), opts)


def test_stokes_and_nonlinear_limits():
    p02, _ = _gaussian_cases(48, 24)
    src = KWaveSource(p0=p02)
    base = _small_as(KWaveMedium(1500.0, 1000.0), src)["pFinal"]
    tiny = _small_as(KWaveMedium(1500.0, 1000.0, alphaCoeff=1e-12, alphaPower=2.0, alphaMode=AbsorptionMode.stokes),
                     src)["pFinal"]
    lin = _small_as(KWaveMedium(1500.0, 1000.0, bOnA=1e-12), src)["pFinal"]
    assert _rl2(tiny, base) < 1e-9 and _rl2(lin, base) < 1e-9
    damped = _small_as(KWaveMedium(1500.0, 1000.0, alphaCoeff=5.0, alphaPower=2.0, alphaMode=AbsorptionMode.stokes),
                       src)["pFinal"]
    assert np.sum(damped ** 2) < np.sum(base ** 2)
    with pytest.raises(ValueError):
        _small_as(KWaveMedium(1500.0, 1000.0, alphaCoeff=0.5, alphaPower=1.5, alphaMode=AbsorptionMode.powerLaw), src)


def test_dirichlet_pressure_source_pins_the_points():
    n, ny, steps = 48, 24, 50
    mask = np.zeros((n, ny), np.float32)
    mask[12, 0:3] = 1
    t = np.arange(steps)
    sig = np.sin(2 * np.pi * 2e6 * t * 0.3 * 1e-4 / 1500.0) * np.exp(-((t - 20.0) ** 2) / 100.0)
    out = _small_as(KWaveMedium(1500.0, 1000.0), KWaveSource(pMask=mask, p=sig.astype(np.float32),
                                                             pMode=SourceMode.dirichlet),
                    KWaveSensor(mask=mask), steps=steps)
    np.testing.assert_allclose(out["p"], np.tile(sig.astype(np.float32).astype(np.float64), (3, 1)), rtol=1e-12,
                               atol=1e-15)
