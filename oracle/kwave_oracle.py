"""CPU float64 restatement of KWaveSwift's `kspaceFirstOrder` time loop (1D/2D/3D).

TEST INFRASTRUCTURE ONLY. This module is the parity checker for the HIP library. Only `tests/`,
`__graft_entry__.smoke()` and the `cpu_baseline` leg of `bench.py` may import it, and only to
check or to time a reported CPU baseline -- never as the thing measured or shipped.

Each function restates the reference algorithm with full complex FFTs and `Re(ifftn(...))`,
exactly as the reference composes them, so that the R2C/C2R + Hermitian-operator formulation of
the GPU path is checked against the literal formulation. Arithmetic is float64 by default
(`dtype=np.float32` gives a float32 port of the same loop for CPU timing).

Reference files (paths relative to the reference checkout):
  Sources/KWave/Solver/KSpaceFirstOrder.swift   solver, operators, sources, absorption, B/A
  Sources/KWave/Grid/KWaveGrid.swift            wavenumbers, makeTime
  Sources/KWave/Internal/Support.swift          fftFreq
  Sources/KWave/PML/PML.swift                   pmlProfile
  Sources/KWave/Filter/Filters.swift            smooth / smooth1D / smooth3D
  Sources/KWave/Sensor/Recording.swift          RecordPlan, samplers, collocation, finalize
  Sources/KWave/Sensor/Directivity.swift        2D directivity filter
  Sources/KWave/Material/Conversion.swift       db2neper

How each part is pinned (tests/test_oracle.py, tests/test_oracle_physics_pins.py):
  fft_freq / wavenumbers     numpy.fft.fftfreq, GridTests.swift:8-16
  make_time                  GridTests.swift:18-24, ParityTests.swift:63 (nt == 302)
  derivative/colloc ops      plane-wave identity (exact spectral derivative of sin)
  whole loop (lossless)      closed form p^(k,t) = p^0(k) cos(c|k|t) with pml_alpha = 0
  pml_profile                closed form exp(-a dt/2 x^4) at the documented positions
  smoothing                  numpy.blackman, exact integer radii (3-4-5, 2-3-6), numpy.interp on a
                             stretched axis, two-mode amplitude ratio; DC gain 1, max restore
  heterogeneous coupling     impedance step: reflection (Z2-Z1)/(Z2+Z1), transmission 2Z2/(Z2+Z1)
                             for c and rho contrasts, along each axis; transmitted speed c2
  absorption / dispersion    mode-wise damping rate vs alpha0 * omega^y * c; alpha -> 0 limit
  B/A                        Fubini solution of a progressive wave (second harmonic growth with
                             beta = 1 + B/2A, B/A -> 0, 2, 6); quadratic scaling; linear limit
  sources                    dirichlet sets values exactly; additive-no-correction linearity
  recording                  pMax >= pRms >= 0, iAvg of an in-phase plane wave
  directivity                isotropic limit (size -> 0) equals plain sampling
Planted mistakes (a Hann window, an n for n-1, c for c^2, dt*rho for dt/rho, a dropped or halved
nonlinear factor) are monkeypatched into the module-level functions below and shown to fail.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Callable, Optional

import numpy as np

# RecordField bits (KWaveSensor.swift:8-20).
P, P_MAX, P_MIN, P_RMS, P_FINAL = 1 << 0, 1 << 1, 1 << 2, 1 << 3, 1 << 4
UX, UY, UZ, U_MAX, U_RMS, U_FINAL = 1 << 5, 1 << 6, 1 << 7, 1 << 8, 1 << 9, 1 << 10
I_AVG, I_MAX = 1 << 11, 1 << 12


class _NumpyFFT:
    """Complex n-D FFT backend (numpy pocketfft). `workers` is accepted for scipy parity."""

    def __init__(self, workers: int = 1):
        self.workers = workers
        try:
            import scipy.fft as sfft  # multithreaded pocketfft, same algorithm
            self._s = sfft
        except Exception:  # pragma: no cover
            self._s = None

    def fftn(self, a):
        if self._s is not None:
            return self._s.fftn(a, workers=self.workers)
        return np.fft.fftn(a)

    def ifftn(self, a):
        if self._s is not None:
            return self._s.ifftn(a, workers=self.workers)
        return np.fft.ifftn(a)


# ---------------------------------------------------------------------------------------------
# Grid (KWaveGrid.swift, Support.swift)
# ---------------------------------------------------------------------------------------------

def fft_freq(n: int, d: float) -> np.ndarray:
    """numpy-style fftfreq (Support.swift:6-14)."""
    val = 1.0 / (n * d)
    out = np.zeros(n, dtype=np.float64)
    cutoff = (n - 1) // 2 + 1
    out[:cutoff] = np.arange(cutoff) * val
    out[cutoff:] = np.arange(-(n // 2), 0) * val
    return out


def wavenumber_vector(n: int, d: float) -> np.ndarray:
    """Angular wavenumbers in FFT order, 2*pi*fftfreq (Support.swift:17-19)."""
    return 2.0 * math.pi * fft_freq(n, d)


def make_time(sizes, spacings, sound_speed_max, sound_speed_min=None, cfl=0.3, t_end=None):
    """dt = cfl*min(d)/cmax; nt = floor(tEnd/dt)+1 (KWaveGrid.swift:112-135)."""
    c_min = sound_speed_min if sound_speed_min is not None else sound_speed_max
    step = cfl * min(spacings) / sound_speed_max
    if t_end is None:
        diag = math.sqrt(sum((n * d) ** 2 for n, d in zip(sizes, spacings)))
        t_end = diag / c_min
    return step, int(math.floor(t_end / step)) + 1


def k_vectors(sizes, spacings):
    return [wavenumber_vector(n, d) for n, d in zip(sizes, spacings)]


def k_magnitude(sizes, spacings):
    """|k| on the full grid in FFT order (KWaveGrid.swift:79-100)."""
    kv = k_vectors(sizes, spacings)
    dim = len(sizes)
    k2 = np.zeros(tuple(sizes))
    for a in range(dim):
        shape = [1] * dim
        shape[a] = sizes[a]
        k2 = k2 + kv[a].reshape(shape) ** 2
    return np.sqrt(k2)


def _bcast(v, a, dim):
    shape = [1] * dim
    shape[a] = v.shape[0]
    return v.reshape(shape)


# ---------------------------------------------------------------------------------------------
# Operators (KSpaceFirstOrder.swift, PML.swift, Recording.swift)
# ---------------------------------------------------------------------------------------------

def pml_profile(n, dx, dt, c, pml_size, pml_alpha, staggered, axisymmetric=False):
    """Split-field PML exponential profile, power 4, independent sides (PML.swift:17-43);
    axisymmetric: no left-side ramp (the radial axis of the axisymmetric solver, :36-38)."""
    a = pml_alpha * (c / dx)
    m = float(pml_size)
    pml = np.ones(n)
    for k in range(min(pml_size, n)):
        x = float(k + 1)
        if staggered:
            pos_l, pos_r = (m + 0.5 - x) / m, (x + 0.5) / m
        else:
            pos_l, pos_r = (m + 1 - x) / m, x / m
        if not axisymmetric:
            pml[k] = math.exp(-a * dt / 2.0 * pos_l ** 4)
        pml[n - pml_size + k] = math.exp(-a * dt / 2.0 * pos_r ** 4)
    return pml


def derivative_operator(kvec, spacing, shift):
    """i*k*exp(+-i*k*d/2) staggered derivative multiplier (KSpaceFirstOrder.swift:764-769)."""
    theta = kvec * (spacing / 2.0)
    return 1j * kvec * (np.cos(theta) + shift * 1j * np.sin(theta))


def collocation_op(kvec, spacing):
    """exp(-i*k*d/2) unstagger multiplier (Recording.swift:101-104)."""
    theta = kvec * (spacing / 2.0)
    return np.cos(theta) - 1j * np.sin(theta)


def stagger_density(rho0, axis):
    """rho0 at +d/2: neighbour mean, trailing cell replicated (KSpaceFirstOrder.swift:799-806)."""
    if np.size(rho0) == 1:
        return rho0
    m = np.moveaxis(rho0, axis, 0)
    st = np.concatenate([0.5 * (m[:-1] + m[1:]), m[-1:]], axis=0)
    return np.moveaxis(st, 0, axis)


def dt_over_staggered_density(dt, rho0, axis):
    """dt / rho0 on the +d/2 grid of `axis`, the velocity update's coefficient
    (KSpaceFirstOrder.swift:612-614; 2D :401-402, 1D :245)."""
    return dt / stagger_density(rho0, axis)


def eos_sound_speed_squared(c0):
    """c0^2, the linear equation-of-state coefficient p = c0^2 rho (KSpaceFirstOrder.swift:610; 2D :399, 1D :243)."""
    return c0 * c0


def nonlinear_mass_factor(rho_total_prev, rho0):
    """(2 sum(rho) + rho0) / rho0 scaling the mass-conservation term, from the previous step's
    split densities (Nonlinearity.factor, KSpaceFirstOrder.swift:894)."""
    return (2.0 * rho_total_prev + rho0) / rho0


def nonlinear_eos_term(bona, rho_total, rho0):
    """B/A * rho^2 / (2 rho0), the quadratic equation-of-state term (Nonlinearity.term,
    KSpaceFirstOrder.swift:897)."""
    return bona * (rho_total * rho_total) / (2.0 * rho0)


def db2neper(alpha, y=1.0):
    """dB/(MHz^y cm) -> Np/((rad/s)^y m) (Conversion.swift:6-8)."""
    return 100.0 * alpha * (1e-6 / (2.0 * math.pi)) ** y / (20.0 * math.log10(math.e))


def fractional_laplacian(k, power):
    """|k|^power with the DC bin forced to 0 (KSpaceFirstOrder.swift:812-816)."""
    out = np.zeros_like(k)
    nz = k != 0
    out[nz] = k[nz] ** power
    return out


# ---------------------------------------------------------------------------------------------
# Smoothing (Filters.swift:203-347)
# ---------------------------------------------------------------------------------------------

def _blackman(n):
    """Symmetric Blackman window (Filters.swift:203-210)."""
    if n <= 1:
        return np.ones(max(n, 0))
    x = 2.0 * math.pi * np.arange(n) / (n - 1)
    return 0.42 - 0.5 * np.cos(x) + 0.08 * np.cos(2 * x)


def _rotation_profile(sizes):
    """Radial Blackman profile on odd-enlarged axes (Filters.swift:252-262)."""
    enlarged = [s if s % 2 else s + 1 for s in sizes]
    L = max(enlarged)
    win = _blackman(L)
    radius = (L - 1) / 2.0
    axes = []
    for count in enlarged:
        if count > 1:
            axes.append(-radius + 2.0 * radius * np.arange(count) / (count - 1))
        else:
            axes.append(np.zeros(1))
    return win, radius, axes


def _interp_radial(r, win, radius):
    """Linear lookup on the unit-spaced radial axis, then abs (Filters.swift:264-269)."""
    pos = np.minimum(r, radius) + radius
    i0 = np.minimum(np.floor(pos).astype(np.int64), len(win) - 2)
    frac = pos - i0
    return np.abs(win[i0] * (1 - frac) + win[i0 + 1] * frac)


def smoothing_window(sizes):
    """The k-space smoothing window in FFT order, exactly as the reference places it.

    1D/3D (Filters.swift:273-321): w[i] lands at (i + (n+1)//2) % n (an ifftshift).
    2D (Filters.swift:325-347): flat[i] = w[(i + (n+1)//2) % n], which equals the 1D/3D
    placement for even n and is a one-bin fftshift for odd n; kept as the reference has it.
    """
    dim = len(sizes)
    win, radius, axes = _rotation_profile(sizes)
    grids = np.meshgrid(*[axes[a][: sizes[a]] for a in range(dim)], indexing="ij")
    r = np.sqrt(sum(g * g for g in grids))
    w = _interp_radial(r, win, radius)
    out = np.zeros(tuple(sizes))
    if dim == 2:
        nx, ny = sizes
        sx, sy = (nx + 1) // 2, (ny + 1) // 2
        ii = (np.arange(nx) + sx) % nx
        jj = (np.arange(ny) + sy) % ny
        out = w[np.ix_(ii, jj)]
    else:
        idx = [(np.arange(n) + (n + 1) // 2) % n for n in sizes]
        out[np.ix_(*idx)] = w
    return out


def smooth(field_, fft=None, restore_max=True):
    """Blackman k-space smoothing with max restore (Filters.swift:273-347)."""
    fft = fft or _NumpyFFT()
    f = np.asarray(field_, dtype=np.float64)
    w = smoothing_window(f.shape)
    sm = np.real(fft.ifftn(w * fft.fftn(f)))
    src_max = np.max(np.abs(f))
    out_max = np.max(np.abs(sm))
    scale = src_max / out_max if out_max > 0 else 1.0
    return sm * scale


# ---------------------------------------------------------------------------------------------
# Sources (KSpaceFirstOrder.swift:129-201, scales :405-415)
# ---------------------------------------------------------------------------------------------

@dataclass
class SourceOp:
    indices: np.ndarray          # flat C-order indices, ascending
    signal: np.ndarray           # [rows, len]
    scale: np.ndarray            # [nSrc]
    mode: str                    # "dirichlet" | "additive" | "additiveNoCorrection"
    shape: tuple

    @property
    def signal_len(self):
        return self.signal.shape[1]

    def values_at(self, t):
        """Scaled values at t, or None once the signal is exhausted (:141-151)."""
        if t >= self.signal_len:
            return None
        if self.signal.shape[0] == 1:
            return self.signal[0, t] * self.scale
        return self.signal[:, t] * self.scale

    def apply(self, fld, t, source_kappa, fft):
        """Inject at time t (:155-174)."""
        v = self.values_at(t)
        if v is None:
            return fld
        if self.mode == "dirichlet":
            out = fld.copy().reshape(-1)
            out[self.indices] = v
            return out.reshape(self.shape)
        buf = np.zeros(int(np.prod(self.shape)))
        buf[self.indices] = v
        buf = buf.reshape(self.shape)
        if self.mode == "additiveNoCorrection":
            return fld + buf
        return fld + np.real(fft.ifftn(source_kappa * fft.fftn(buf)))


def flat_nonzero(mask):
    """Flat C-order indices of nonzero mask entries (KSpaceFirstOrder.swift:778-782)."""
    return np.flatnonzero(np.asarray(mask).reshape(-1) != 0)


def signal_rows(sig):
    """1-D -> one broadcast row; 2-D -> one row per source point (:179-187)."""
    a = np.asarray(sig, dtype=np.float64)
    if a.ndim <= 1:
        return a.reshape(1, -1)
    return a.reshape(a.shape[0], -1)


def make_source_op(mask, signal, mode, c0, scale_fn, shape):
    """(:192-201). Returns None if mask/signal absent or mask empty."""
    if mask is None or signal is None:
        return None
    idx = flat_nonzero(mask)
    if idx.size == 0:
        return None
    c = np.asarray(c0, dtype=np.float64)
    cs = np.full(idx.size, float(c.reshape(-1)[0])) if c.size == 1 else c.reshape(-1)[idx]
    rows = signal_rows(signal)
    if rows.shape[0] not in (1, idx.size):
        raise ValueError("source signal must have 1 row or one row per source point")
    return SourceOp(idx, rows, scale_fn(cs), mode, tuple(shape))


# ---------------------------------------------------------------------------------------------
# Sensors (Recording.swift, Directivity.swift)
# ---------------------------------------------------------------------------------------------

@dataclass
class Sampler:
    n_points: int
    idx: np.ndarray
    weight: Optional[np.ndarray] = None
    n_corners: int = 1

    def sample(self, f):
        flat = f.reshape(-1)
        if self.weight is None:
            return flat[self.idx]
        return (flat[self.idx] * self.weight).reshape(self.n_points, self.n_corners).sum(axis=1)


def make_sampler(mask, sizes, spacings):
    """Binary vs Cartesian dispatch (Recording.swift:50-56)."""
    m = np.asarray(mask)
    dim = len(sizes)
    if m.ndim == 2 and m.shape[0] == dim:
        return cartesian_sampler(m, sizes, spacings)
    idx = flat_nonzero(m)
    return Sampler(idx.size, idx)


def cartesian_sampler(mask, sizes, spacings):
    """Multilinear corners + weights, coordinate (i - n//2)*d (Recording.swift:60-97)."""
    dim = len(sizes)
    pts = np.asarray(mask, dtype=np.float32).astype(np.float64)
    npts = pts.shape[1]
    nc = 1 << dim
    idx = np.zeros(npts * nc, dtype=np.int64)
    w = np.zeros(npts * nc)
    for p in range(npts):
        i0, fr = [], []
        for a in range(dim):
            g = pts[a, p] / spacings[a] + float(sizes[a] // 2)
            if not (g >= 0 and g <= sizes[a] - 1):
                raise ValueError(f"Cartesian sensor point outside grid bounds (axis {a})")
            f = math.floor(g)
            i0.append(int(f))
            fr.append(g - f)
        for c in range(nc):
            flat, ww = 0, 1.0
            for a in range(dim):
                bit = (c >> a) & 1
                ai = min(i0[a] + bit, sizes[a] - 1)
                ww *= fr[a] if bit else (1 - fr[a])
                flat = flat * sizes[a] + ai
            idx[p * nc + c] = flat
            w[p * nc + c] = ww
    return Sampler(npts, idx, w, nc)


def directivity_filter(sensor, sizes, spacings):
    """Per unique angle: weight, rows, gather (Directivity.swift:40-79). None if not configured."""
    if len(sizes) != 2 or sensor.mask is None or getattr(sensor, "directivityAngle", None) is None:
        return None
    flat = flat_nonzero(sensor.mask)
    if flat.size == 0:
        return None
    ang = np.asarray(sensor.directivityAngle, dtype=np.float32).reshape(-1).astype(np.float64)
    row_angle = ang[flat]
    uniq = sorted(set(row_angle.tolist()))
    size = sensor.directivitySize if sensor.directivitySize is not None else 10 * max(spacings)
    kv = k_vectors(sizes, spacings)
    kx = np.broadcast_to(kv[0].reshape(-1, 1), tuple(sizes))
    ky = np.broadcast_to(kv[1].reshape(1, -1), tuple(sizes))
    kmag = np.sqrt(kx * kx + ky * ky)
    entries = []
    for th in uniq:
        if _mode(getattr(sensor, "directivityPattern", None), "pressure") == "gradient":
            kn = math.sin(th) * ky + math.cos(th) * kx
            wgt = np.where(kmag == 0, 0.0, kn / np.where(kmag == 0, 1.0, kmag))
        else:
            kt = math.cos(th) * ky - math.sin(th) * kx
            x = kt * (size / 2)
            wgt = np.where(x == 0, 1.0, np.sin(math.pi * x) / np.where(x == 0, 1.0, math.pi * x))
        rows = np.array([i for i in range(flat.size) if row_angle[i] == th], dtype=np.int64)
        entries.append((wgt, rows, flat[rows]))
    return entries, flat.size


def intensity_avg(p, u):
    """mean_t(p * u shifted by +dt/2 via a Fourier interpolant) (Recording.swift:113-124)."""
    nt = p.shape[1]
    half = nt // 2
    kk = np.array([k if k <= half else k - nt for k in range(nt)], dtype=np.float64)
    theta = 2.0 * math.pi * kk / nt * 0.5
    op = (np.cos(theta) + 1j * np.sin(theta)).reshape(1, nt)
    us = np.real(np.fft.ifft(op * np.fft.fft(u, axis=1), axis=1))
    return np.mean(p * us, axis=1)


def finalize_recording(record, p, us, p_final, u_finals):
    """Aggregates from the series; pFinal always (Recording.swift:129-168)."""
    out = {"pFinal": p_final}
    if record & P:
        out["p"] = p
    if p is not None:
        if record & P_MAX:
            out["pMax"] = p.max(axis=1)
        if record & P_MIN:
            out["pMin"] = p.min(axis=1)
        if record & P_RMS:
            out["pRms"] = np.sqrt(np.mean(p * p, axis=1))
    names = ["x", "y", "z"]
    for a, u in enumerate(us):
        if u is None:
            continue
        nm = names[a]
        if record & (UX << a):
            out["u" + nm] = u
        if record & U_MAX:
            out[f"u{nm}Max"] = u.max(axis=1)
        if record & U_RMS:
            out[f"u{nm}Rms"] = np.sqrt(np.mean(u * u, axis=1))
        if record & U_FINAL and u_finals is not None:
            out[f"u{nm}Final"] = u_finals[a]
        if record & I_AVG and p is not None:
            out[f"i{nm}Avg"] = intensity_avg(p, u)
    return out


class _Recorder:
    """The recording of one sensor (RecordPlan / SensorSampler, Recording.swift:8-97, and
    finalizeRecording, :129-168). When the record needs no series (no p / u* / iAvg) the samples
    are folded into running max / min / sum of squares instead of being stored, which is the same
    max / min / rms over the recorded steps without the [nSensor, nt] array."""

    def __init__(self, sensor, sizes, spac, dim, kv):
        self.dim = dim
        self.rec = rec = int(sensor.record) if sensor is not None else P
        self.rec_p = bool(rec & (P | P_MAX | P_MIN | P_RMS | I_AVG))
        self.rec_u = bool(rec & (UX | UY | UZ | U_MAX | U_RMS | I_AVG))
        self.sampler = (make_sampler(sensor.mask, sizes, spac)
                        if (sensor is not None and sensor.mask is not None) else None)
        self.direct = directivity_filter(sensor, sizes, spac) if sensor is not None else None
        self.colloc = [_bcast(collocation_op(kv[a], spac[a]), a, dim) for a in range(dim)] if self.rec_u else None
        self.keep = bool(rec & (P | UX | UY | UZ | I_AVG))
        self.p_rec, self.u_rec = [], [[] for _ in range(dim)]
        self.p_agg, self.u_agg = None, [None] * dim

    @staticmethod
    def _fold(agg, v):
        if agg is None:
            return [v.copy(), v.copy(), v * v, 1]
        np.maximum(agg[0], v, out=agg[0])          # NaN propagates, as in max over the series
        np.minimum(agg[1], v, out=agg[1])
        agg[2] += v * v
        agg[3] += 1
        return agg

    def record(self, p, u, fft):
        if self.sampler is None:
            return
        if self.rec_p:
            if self.direct is not None:
                entries, nsens = self.direct
                pkd = fft.fftn(p)
                v = np.zeros(nsens)
                for wgt, rows, gather in entries:
                    v[rows] = np.real(fft.ifftn(pkd * wgt)).reshape(-1)[gather]
            else:
                v = self.sampler.sample(p)
            if self.keep:
                self.p_rec.append(v)
            else:
                self.p_agg = self._fold(self.p_agg, v)
        if self.rec_u:
            for a in range(self.dim):
                uc = np.real(fft.ifftn(self.colloc[a] * fft.fftn(u[a])))
                v = self.sampler.sample(uc)
                if self.keep:
                    self.u_rec[a].append(v)
                else:
                    self.u_agg[a] = self._fold(self.u_agg[a], v)

    def finish(self, p, u):
        rec = self.rec
        if self.keep or (self.p_agg is None and all(a is None for a in self.u_agg)):
            return finalize_recording(
                rec,
                np.stack(self.p_rec, axis=1) if self.p_rec else None,
                [np.stack(self.u_rec[a], axis=1) if self.u_rec[a] else None for a in range(self.dim)],
                p,
                [x for x in u] if rec & U_FINAL else None,
            )
        out = {"pFinal": p}
        if self.p_agg is not None:
            mx, mn, s2, n = self.p_agg
            if rec & P_MAX:
                out["pMax"] = mx
            if rec & P_MIN:
                out["pMin"] = mn
            if rec & P_RMS:
                out["pRms"] = np.sqrt(s2 / n)
        for a, agg in enumerate(self.u_agg):
            if agg is None:
                continue
            nm = "xyz"[a]
            mx, _, s2, n = agg
            if rec & U_MAX:
                out[f"u{nm}Max"] = mx
            if rec & U_RMS:
                out[f"u{nm}Rms"] = np.sqrt(s2 / n)
            if rec & U_FINAL:
                out[f"u{nm}Final"] = u[a]
        return out


# ---------------------------------------------------------------------------------------------
# Solver (KSpaceFirstOrder.swift:203-761)
# ---------------------------------------------------------------------------------------------

def _mode(x, default):
    """Enum-or-string option value."""
    if x is None:
        return default
    return str(getattr(x, "value", x))


def kspace_first_order(grid, medium, source, sensor, options, dtype=np.float64, workers=1,
                       t_begin=0, t_end=None, state=None, return_state=False, progress=None):
    """Run the k-space first-order loop and return a dict of recorded outputs.

    `grid` exposes dim, size (list), spacing (list), dt, nt. `medium` exposes soundSpeed,
    density, alphaCoeff, alphaPower, alphaMode, bOnA. `source` exposes p0, pMask, p, pMode,
    uMask, ux, uy, uz, uMode. `sensor` exposes mask, record (RecordField bits),
    directivityAngle, directivitySize, directivityPattern; a list of sensors records each of them
    in the same run (a list of dicts is returned). `options` exposes pmlSize, pmlAlpha, smoothP0,
    smoothC0, smoothRho0. `progress(t)` is called after every step.

    Follows kspaceFirstOrder1D/2D/3D (KSpaceFirstOrder.swift:203-349, 351-555, 557-761); the
    2D no-source fused step (:475-483) is the same arithmetic and is not separated here.
    """
    fft = _NumpyFFT(workers)
    R = np.dtype(dtype).type
    dim = grid.dim
    sizes = list(grid.size)
    spac = list(grid.spacing)
    shape = tuple(sizes)
    dt = float(grid.dt)
    nt = int(grid.nt)
    if nt <= 0:
        raise ValueError("call grid.makeTime(...) before running the solver")
    if source.p0 is None and source.pMask is None and source.uMask is None:
        raise ValueError(f"{dim}D solver requires source.p0 or a time-varying p/u source")

    c0 = np.asarray(medium.soundSpeed, dtype=np.float32).astype(np.float64)
    rho0 = np.asarray(medium.density, dtype=np.float32).astype(np.float64)
    if not (c0.size == 1 or c0.shape == shape):
        raise ValueError("soundSpeed must be scalar or grid-shaped")
    if not (rho0.size == 1 or rho0.shape == shape):
        raise ValueError("density must be scalar or grid-shaped")
    if c0.size == 1:
        c0 = c0.reshape(())
    if rho0.size == 1:
        rho0 = rho0.reshape(())
    if options.smoothC0 and c0.size > 1:
        c0 = smooth(c0, fft)
    if options.smoothRho0 and rho0.size > 1:
        rho0 = smooth(rho0, fft)
    c_ref = float(np.max(c0))                                          # :786-788

    kmag = k_magnitude(sizes, spac)
    arg = kmag * (c_ref * dt / 2.0)                                      # :581-583
    kappa = np.where(arg == 0, 1.0, np.sin(arg) / np.where(arg == 0, 1.0, arg))
    source_kappa = np.cos(arg)

    kv = k_vectors(sizes, spac)
    ddx_pos = [_bcast(derivative_operator(kv[a], spac[a], +1), a, dim) for a in range(dim)]
    ddx_neg = [_bcast(derivative_operator(kv[a], spac[a], -1), a, dim) for a in range(dim)]
    pml_n = int(options.pmlSize)
    pml = [_bcast(pml_profile(sizes[a], spac[a], dt, c_ref, pml_n, options.pmlAlpha, False), a, dim)
           for a in range(dim)]
    pml_sg = [_bcast(pml_profile(sizes[a], spac[a], dt, c_ref, pml_n, options.pmlAlpha, True), a, dim)
              for a in range(dim)]

    c2 = eos_sound_speed_squared(c0)
    dt_rho0 = dt * rho0
    dt_over_rho = [dt_over_staggered_density(dt, rho0, a) for a in range(dim)]

    n = float(dim)
    p_mode = _mode(source.pMode, "additive")
    u_mode = _mode(source.uMode, "additive")

    def p_scale(d):
        if p_mode == "dirichlet":
            return lambda c: 1.0 / (n * c * c)
        return lambda c: 2.0 * dt / (n * c * d)

    def u_scale(d):
        if u_mode == "dirichlet":
            return lambda c: np.ones_like(c)
        return lambda c: 2.0 * c * dt / d

    p_ops = [make_source_op(source.pMask, source.p, p_mode, c0, p_scale(spac[a]), shape)
             for a in range(dim)]
    u_sig = [source.ux, getattr(source, "uy", None), getattr(source, "uz", None)]
    u_ops = [make_source_op(source.uMask, u_sig[a], u_mode, c0, u_scale(spac[a]), shape)
             for a in range(dim)]

    p0f = None
    if source.p0 is not None:
        p0f = np.asarray(source.p0, dtype=np.float32).astype(np.float64).reshape(shape)
        if options.smoothP0:
            p0f = smooth(p0f, fft)

    # Absorption (:851-882) and nonlinearity (:902-907).
    absorb = None
    alpha_mode = _mode(medium.alphaMode, "noAbsorption")
    if medium.alphaCoeff is not None and alpha_mode != "noAbsorption":
        y = medium.alphaPower if medium.alphaPower is not None else 1.5
        if alpha_mode == "stokes":
            y = 2.0
        a_np = np.asarray(medium.alphaCoeff, dtype=np.float32).astype(np.float64)
        a_np = (a_np.reshape(()) if a_np.size == 1 else a_np) * db2neper(1.0, y)
        stokes = alpha_mode == "stokes" or abs(y - 2.0) < 1e-10
        if stokes:
            tau, nab1 = -2.0 * a_np * c0, None
        else:
            tau, nab1 = -2.0 * a_np * c0 ** (y - 1.0), fractional_laplacian(kmag, y - 2.0)
        eta = nab2 = None
        if not stokes and alpha_mode != "noDispersion":
            eta = 2.0 * a_np * c0 ** y * math.tan(math.pi * y / 2.0)
            nab2 = fractional_laplacian(kmag, y - 1.0)
        absorb = (tau, nab1, eta, nab2)
    bona = None
    if medium.bOnA is not None:
        b = np.asarray(medium.bOnA, dtype=np.float32).astype(np.float64)
        if np.max(np.abs(b)) != 0:
            bona = b.reshape(()) if b.size == 1 else b

    def eos_term(div_u, rho_t):
        tau, nab1, eta, nab2 = absorb
        if nab1 is not None:
            ab = tau * np.real(fft.ifftn(nab1 * fft.fftn(rho0 * div_u)))
        else:
            ab = tau * rho0 * div_u
        if eta is None:
            return ab
        return ab - eta * np.real(fft.ifftn(nab2 * fft.fftn(rho_t)))

    several = isinstance(sensor, (list, tuple))
    recorders = [_Recorder(sn, sizes, spac, dim, kv) for sn in (sensor if several else [sensor])]

    cast = (lambda x: x) if R is np.float64 else (lambda x: np.asarray(x, dtype=R))
    if R is not np.float64:
        # single precision end to end: the operators in float32 / complex64 keep every product and
        # every pocketfft call (complex64 in, complex64 out) at the working precision
        C = np.result_type(R, np.complex64).type
        kappa = kappa.astype(R)
        ddx_pos = [d.astype(C) for d in ddx_pos]
        ddx_neg = [d.astype(C) for d in ddx_neg]
        pml = [x.astype(R) for x in pml]
        pml_sg = [x.astype(R) for x in pml_sg]
        dt_over_rho = [np.asarray(x, dtype=R) for x in dt_over_rho]
        dt_rho0 = np.asarray(dt_rho0, dtype=R)
        c2 = np.asarray(c2, dtype=R)
        rho0 = np.asarray(rho0, dtype=R)          # eos_term and the nonlinear factor read it
        if absorb is not None:
            absorb = tuple(None if x is None else np.asarray(x, dtype=R) for x in absorb)
        if bona is not None:
            bona = np.asarray(bona, dtype=R)
    if state is None:
        p = np.zeros(shape, dtype=R)
        u = [np.zeros(shape, dtype=R) for _ in range(dim)]
        rho = [np.zeros(shape, dtype=R) for _ in range(dim)]
    else:
        p = state["p"].copy()
        u = [x.copy() for x in state["u"]]
        rho = [x.copy() for x in state["rho"]]
    t_stop = nt if t_end is None else t_end
    import time as _time
    loop_start = _time.perf_counter()
    step_seconds = []

    for t in range(t_begin, t_stop):
        step_start = _time.perf_counter()
        pk = fft.fftn(p)
        for a in range(dim):                                              # velStep
            g = np.real(fft.ifftn(ddx_pos[a] * kappa * pk))
            u[a] = cast(pml_sg[a] * (pml_sg[a] * u[a] - dt_over_rho[a] * g))
        for a in range(dim):
            if u_ops[a] is not None:
                u[a] = cast(u_ops[a].apply(u[a], t, source_kappa, fft))
        rho_prev = sum(rho)                                               # rhoStep
        nl = nonlinear_mass_factor(rho_prev, rho0) if bona is not None else None
        du = []
        for a in range(dim):
            d = cast(np.real(fft.ifftn(ddx_neg[a] * kappa * fft.fftn(u[a]))))
            mass = dt_rho0 * d * nl if nl is not None else dt_rho0 * d
            rho[a] = cast(pml[a] * (pml[a] * rho[a] - mass))
            du.append(d)
        for a in range(dim):
            if p_ops[a] is not None:
                rho[a] = cast(p_ops[a].apply(rho[a], t, source_kappa, fft))
        rho_t = sum(rho)                                                  # eosStep
        eos = rho_t
        if absorb is not None:
            eos = eos + eos_term(sum(du), rho_t)
        if bona is not None:
            eos = eos + nonlinear_eos_term(bona, rho_t, rho0)
        p = cast(c2 * eos)
        if t == 0 and p0f is not None:                                    # :724-733
            p = cast(p0f.copy())
            for a in range(dim):
                rho[a] = cast(p0f / (n * c2))
            pk0 = fft.fftn(p)
            for a in range(dim):
                u[a] = cast((dt_over_rho[a] / 2.0) * np.real(fft.ifftn(ddx_pos[a] * kappa * pk0)))
        for r in recorders:
            r.record(p, u, fft)
        step_seconds.append(_time.perf_counter() - step_start)
        if progress is not None:
            progress(t)

    loop_seconds = _time.perf_counter() - loop_start
    outs = [r.finish(p, u) for r in recorders]
    for out in outs:
        if return_state:
            out["_state"] = {"p": p, "u": u, "rho": rho}
        out["_loop_seconds"] = loop_seconds
        out["_step_seconds"] = step_seconds
    return outs if several else outs[0]


def closed_form_lossless(p0, sizes, spacings, c, dt, steps):
    """p(t_n) for the homogeneous, lossless, PML-free p0 problem: IFFT(P0(k) cos(c|k| n dt)).

    The k-space corrected leapfrog is exact for a homogeneous medium; with the t=0 override the
    recorded pressure at loop index n is this closed form. Used to pin the oracle's loop.
    """
    kmag = k_magnitude(list(sizes), list(spacings))
    P0 = np.fft.fftn(p0)
    return [np.real(np.fft.ifftn(P0 * np.cos(c * kmag * n * dt))) for n in steps]


# ---------------------------------------------------------------------------------------------
# Axisymmetric solver (Sources/KWave/Solver/KSpaceFirstOrderAS.swift)
# ---------------------------------------------------------------------------------------------

def mirror_wswa(w):
    """[w, 0, -w[:, ny-1..1], -w, 0, w[:, ny-1..1]] along the radial axis (:201-204)."""
    nx = w.shape[0]
    tail = w[:, 1:][:, ::-1]
    z = np.zeros((nx, 1))
    return np.concatenate([w, z, -tail, -w, z, tail], axis=1)


def mirror_hahs(w):
    """[w, flip w, -w, -flip w] (:205-208)."""
    f = w[:, ::-1]
    return np.concatenate([w, f, -w, -f], axis=1)


def mirror_hsha(w):
    """[w, -flip w, -w, flip w] (:209-212)."""
    f = w[:, ::-1]
    return np.concatenate([w, -f, -w, f], axis=1)


def kspace_first_order_as(grid, medium, source, sensor, options, workers=1):
    """Axisymmetric k-space solver with WSWA-FFT radial symmetry (KSpaceFirstOrderAS.swift:17-309).

    The grid is [nx, ny] with x axial (periodic) and y radial (index 0 on the axis). Every field
    is mirrored into a 4x radially expanded grid with its own symmetry (p, ux: WSWA; uy: HAHS;
    uy / r: HSHA) and transformed there with full complex FFTs; the radial PML is one-sided and
    the divergence gains uy / r on the staggered radial grid r = (j + 1/2) dy. Returns
    {"p": [nSensor, nt] or None, "pFinal": [nx, ny]}.
    """
    fft = _NumpyFFT(workers)
    if grid.dim != 2:
        raise ValueError("axisymmetric solver requires a 2D grid (x axial, y radial)")
    nx, ny = (int(v) for v in grid.size)
    dx, dy = (float(v) for v in grid.spacing)
    shape = (nx, ny)
    dt = float(grid.dt)
    nt = int(grid.nt)
    if nt <= 0:
        raise ValueError("call grid.makeTime(...) before running the solver")
    if source.p0 is None and source.pMask is None and source.uMask is None:
        raise ValueError("axisymmetric solver requires source.p0 or a time-varying p/u source")

    c0 = np.asarray(medium.soundSpeed, dtype=np.float32).astype(np.float64)
    rho0 = np.asarray(medium.density, dtype=np.float32).astype(np.float64)
    if c0.size == 1:
        c0 = c0.reshape(())
    if rho0.size == 1:
        rho0 = rho0.reshape(())
    if options.smoothC0 and c0.size > 1:
        c0 = smooth(c0, fft)
    if options.smoothRho0 and rho0.size > 1:
        rho0 = smooth(rho0, fft)
    c_ref = float(np.max(c0))

    # expanded grid (:52-67): kappa from the expanded |k|, axial / radial staggered operators
    ny_e = 4 * ny
    kx = wavenumber_vector(nx, dx)
    ky_e = wavenumber_vector(ny_e, dy).reshape(1, ny_e)
    k_exp = np.sqrt(kx.reshape(nx, 1) ** 2 + ky_e ** 2)
    arg = k_exp * (c_ref * dt / 2.0)
    kappa = np.where(arg == 0, 1.0, np.sin(arg) / np.where(arg == 0, 1.0, arg))
    ddx_pos = derivative_operator(kx, dx, +1).reshape(nx, 1)
    ddx_neg = derivative_operator(kx, dx, -1).reshape(nx, 1)
    ddy_k = 1j * ky_e
    theta = ky_e * (dy / 2.0)
    y_shift_pos = np.cos(theta) + 1j * np.sin(theta)
    y_shift_neg = np.cos(theta) - 1j * np.sin(theta)

    # PMLs (:69-81): axial two-sided, radial outer edge only
    pml_n = int(options.pmlSize)
    pml_x = pml_profile(nx, dx, dt, c_ref, pml_n, options.pmlAlpha, False).reshape(nx, 1)
    pml_x_sg = pml_profile(nx, dx, dt, c_ref, pml_n, options.pmlAlpha, True).reshape(nx, 1)
    pml_y = pml_profile(ny, dy, dt, c_ref, pml_n, options.pmlAlpha, False, True).reshape(1, ny)
    pml_y_sg = pml_profile(ny, dy, dt, c_ref, pml_n, options.pmlAlpha, True, True).reshape(1, ny)
    inv_y_sg = (1.0 / ((np.arange(ny) + 0.5) * dy)).reshape(1, ny)      # (:84-86)

    c2 = eos_sound_speed_squared(c0)
    dt_rho0 = dt * rho0
    dt_over_rho_x = dt_over_staggered_density(dt, rho0, 0)
    dt_over_rho_y = dt_over_staggered_density(dt, rho0, 1)

    absorb_tau = None                                                     # (:93-101)
    alpha_mode = _mode(medium.alphaMode, "noAbsorption")
    if medium.alphaCoeff is not None and alpha_mode != "noAbsorption":
        y = 2.0 if alpha_mode == "stokes" else float(medium.alphaPower if medium.alphaPower is not None else 2.0)
        if abs(y - 2.0) >= 1e-10:
            raise ValueError("axisymmetric absorption requires the Stokes model (alphaPower = 2)")
        alpha = np.asarray(medium.alphaCoeff, dtype=np.float32).astype(np.float64)
        absorb_tau = -2.0 * alpha * db2neper(1.0, 2.0) * c0
    bona = None                                                           # (:103-108)
    if medium.bOnA is not None:
        b = np.asarray(medium.bOnA, dtype=np.float32).astype(np.float64)
        if np.max(np.abs(b)) != 0:
            bona = b

    p0_field = None
    if source.p0 is not None:
        p0_field = np.asarray(source.p0, dtype=np.float32).astype(np.float64)
        if options.smoothP0:
            p0_field = smooth(p0_field, fft)

    # time-varying sources, scaled as in kspaceFirstOrder_scaleSourceTerms (:152-163)
    p_mode = _mode(source.pMode, "additive")
    u_mode = _mode(source.uMode, "additive")
    p_op = make_source_op(source.pMask, source.p, p_mode, c0,
                          (lambda c: 1.0 / (2.0 * c * c)) if p_mode == "dirichlet"
                          else (lambda c: 2.0 * dt / (2.0 * c * dx)), shape)
    ux_op = make_source_op(source.uMask, source.ux, u_mode, c0,
                           (lambda c: np.ones_like(c)) if u_mode == "dirichlet" else (lambda c: 2.0 * c * dt / dx),
                           shape)
    uy_op = make_source_op(source.uMask, source.uy, u_mode, c0,
                           (lambda c: np.ones_like(c)) if u_mode == "dirichlet" else (lambda c: 2.0 * c * dt / dy),
                           shape)
    source_kappa = np.cos(arg)

    def crop(spec):
        return np.real(fft.ifftn(spec))[:, :ny]

    def correct(mat, mirror):                                             # (:170-173)
        return crop(source_kappa * fft.fftn(mirror(mat)))

    def matrix_at(op, t):
        v = op.values_at(t)
        if v is None:
            return None
        buf = np.zeros(nx * ny)
        buf[op.indices] = v
        return buf.reshape(shape)

    def apply_velocity_source(op, fld, t, mirror):                        # (:177-194)
        if op is None:
            return fld
        if u_mode == "dirichlet":
            v = op.values_at(t)
            if v is None:
                return fld
            out = fld.copy().reshape(-1)
            out[op.indices] = v
            return out.reshape(shape)
        mat = matrix_at(op, t)
        if mat is None:
            return fld
        if u_mode == "additiveNoCorrection":
            return fld + mat
        return fld + correct(mat, mirror)

    def pressure_gradients(pf):                                           # (:216-221)
        pk = kappa * fft.fftn(mirror_wswa(pf))
        return crop(ddx_pos * pk), crop(ddy_k * y_shift_pos * pk)

    p = np.zeros(shape)
    ux = np.zeros(shape)
    uy = np.zeros(shape)
    rhox = np.zeros(shape)
    rhoy = np.zeros(shape)
    sampler = make_sampler(sensor.mask, [nx, ny], [dx, dy]) if sensor.mask is not None else None
    p_rec = []
    for t in range(nt):
        dpdx, dpdy = pressure_gradients(p)
        ux = pml_x_sg * (pml_x_sg * ux - dt_over_rho_x * dpdx)
        ux = apply_velocity_source(ux_op, ux, t, mirror_wswa)
        uy = pml_y_sg * (pml_y_sg * uy - dt_over_rho_y * dpdy)
        uy = apply_velocity_source(uy_op, uy, t, mirror_hahs)

        ux_k = kappa * fft.fftn(mirror_wswa(ux))                           # (:243-251)
        duxdx = crop(ddx_neg * ux_k)
        uy_k = ddy_k * fft.fftn(mirror_hahs(uy)) + fft.fftn(mirror_hsha(inv_y_sg * uy))
        duydy = crop(kappa * y_shift_neg * uy_k)

        if bona is not None:                                              # (:253-261)
            dt_rho_total = dt * rho0 * nonlinear_mass_factor(rhox + rhoy, rho0)
            rhox = pml_x * (pml_x * rhox - dt_rho_total * duxdx)
            rhoy = pml_y * (pml_y * rhoy - dt_rho_total * duydy)
        else:
            rhox = pml_x * (pml_x * rhox - dt_rho0 * duxdx)
            rhoy = pml_y * (pml_y * rhoy - dt_rho0 * duydy)

        if p_op is not None:                                              # (:263-279)
            if p_mode == "dirichlet":
                v = p_op.values_at(t)
                if v is not None:
                    rhox = rhox.copy().reshape(-1)
                    rhoy = rhoy.copy().reshape(-1)
                    rhox[p_op.indices] = v
                    rhoy[p_op.indices] = v
                    rhox = rhox.reshape(shape)
                    rhoy = rhoy.reshape(shape)
            else:
                mat = matrix_at(p_op, t)
                if mat is not None:
                    if p_mode == "additive":
                        mat = correct(mat, mirror_wswa)
                    rhox = rhox + mat
                    rhoy = rhoy + mat

        rho_total = rhox + rhoy                                           # (:281-287)
        eos = rho_total
        if absorb_tau is not None:
            eos = eos + absorb_tau * rho0 * (duxdx + duydy)
        if bona is not None:
            eos = eos + nonlinear_eos_term(bona, rho_total, rho0)
        p = c2 * eos

        if t == 0 and p0_field is not None:                               # (:289-297)
            p = p0_field.copy()
            rhox = p0_field / (2.0 * c2)
            rhoy = p0_field / (2.0 * c2)
            dpdx0, dpdy0 = pressure_gradients(p)
            ux = (dt_over_rho_x / 2.0) * dpdx0
            uy = (dt_over_rho_y / 2.0) * dpdy0

        if sampler is not None:
            p_rec.append(sampler.sample(p))
    return {"p": np.stack(p_rec, axis=1) if p_rec else None, "pFinal": p}
