"""Input construction helpers (host-side, numpy), as the reference provides them.

  makeDisc / makeCircle / makeBall   Sources/KWave/Geometry/Shapes.swift:14-79
  makeCartCircle                     Sources/KWave/Geometry/CartesianShapes.swift:53-68
  toneBurst / gaussian               Sources/KWave/Signal/Generation.swift:10-49
  getWin                             Sources/KWave/Filter/Filters.swift:12-24
  db2neper / neper2db                Sources/KWave/Material/Conversion.swift:6-15
plus the seeded random-medium recipe used by the benchmark workloads (`random_field`).
"""
from __future__ import annotations

import math

import numpy as np


def _center(n: int) -> int:
    """Default centre index n//2 (Shapes.swift:4-6)."""
    return n // 2


def makeDisc(nx: int, ny: int, cx=None, cy=None, radius: float = 1.0) -> np.ndarray:
    """Filled disc mask, float32 (Shapes.swift:14-23)."""
    cx = _center(nx) if cx is None else cx
    cy = _center(ny) if cy is None else cy
    xs = np.arange(nx, dtype=np.float32).reshape(nx, 1) - np.float32(cx)
    ys = np.arange(ny, dtype=np.float32).reshape(1, ny) - np.float32(cy)
    return ((xs * xs + ys * ys) <= np.float32(radius * radius)).astype(np.float32)


def makeBall(nx: int, ny: int, nz: int, cx=None, cy=None, cz=None, radius: float = 1.0) -> np.ndarray:
    """Filled ball mask, float32 (Shapes.swift:64-79)."""
    cx = _center(nx) if cx is None else cx
    cy = _center(ny) if cy is None else cy
    cz = _center(nz) if cz is None else cz
    xs = (np.arange(nx, dtype=np.float32) - np.float32(cx)).reshape(nx, 1, 1)
    ys = (np.arange(ny, dtype=np.float32) - np.float32(cy)).reshape(1, ny, 1)
    zs = (np.arange(nz, dtype=np.float32) - np.float32(cz)).reshape(1, 1, nz)
    return ((xs * xs + ys * ys + zs * zs) <= np.float32(radius * radius)).astype(np.float32)


def makeCircle(nx: int, ny: int, cx=None, cy=None, radius: float = 1.0, arcAngle=None) -> np.ndarray:
    """Midpoint-circle outline mask (Shapes.swift:32-61), 1-based internally as the reference."""
    arc = min(max(arcAngle if arcAngle is not None else 2 * math.pi, 0.0), 2 * math.pi)
    r = int(round(radius))
    cx1 = (_center(nx) if cx is None else cx) + 1
    cy1 = (_center(ny) if cy is None else cy) + 1
    circle = np.zeros(nx * ny, dtype=np.float32)

    def put(px, py):
        if math.atan2(px - cx1, py - cy1) + math.pi > arc:
            return
        if px < 1 or px > nx or py < 1 or py > ny:
            return
        circle[(px - 1) * ny + (py - 1)] = 1

    x, y, d = 0, r, 1 - r
    put(cx1, cy1 - y)
    put(cx1, cy1 + y)
    put(cx1 + y, cy1)
    put(cx1 - y, cy1)
    while x < y - 1:
        x += 1
        if d < 0:
            d += 2 * x + 1
        else:
            y -= 1
            d += 2 * (x - y + 1)
        px = [x + cx1, y + cx1, y + cx1, x + cx1, -x + cx1, -y + cx1, -y + cx1, -x + cx1]
        py = [y + cy1, x + cy1, -x + cy1, -y + cy1, -y + cy1, -x + cy1, x + cy1, y + cy1]
        for i in range(8):
            put(px[i], py[i])
    return circle.reshape(nx, ny)


def makeCartCircle(radius: float, numPoints: int, center=(0.0, 0.0), arcAngle: float = 2 * math.pi) -> np.ndarray:
    """[2, numPoints] points on a circle / arc, float32 (CartesianShapes.swift:53-68)."""
    if numPoints <= 0:
        raise ValueError("numPoints must be positive")
    full = arcAngle == 2 * math.pi
    nsteps = float(numPoints if full else numPoints - 1)
    out = np.zeros((2, numPoints), dtype=np.float32)
    for i in range(numPoints):
        ang = i * arcAngle / nsteps + math.pi / 2
        out[0, i] = np.float32(radius * math.cos(ang) + center[0])
        out[1, i] = np.float32(radius * math.sin(-ang) + center[1])
    return out


def gaussian(x, magnitude=None, mean: float = 0.0, variance: float = 1.0) -> np.ndarray:
    """Sampled Gaussian (Generation.swift:10-13)."""
    x = np.asarray(x, dtype=np.float64)
    mag = 1.0 / math.sqrt(2 * math.pi * variance) if magnitude is None else magnitude
    return mag * np.exp(-(x - mean) ** 2 / (2 * variance))


def getWin(n: int, type: str = "hann") -> np.ndarray:
    """Symmetric window (Filters.swift:12-24)."""
    if n <= 1:
        return np.ones(max(n, 0))
    x = np.arange(n) / (n - 1)
    if type == "hann":
        return 0.5 - 0.5 * np.cos(2 * math.pi * x)
    if type == "blackman":
        return 0.42 - 0.5 * np.cos(2 * math.pi * x) + 0.08 * np.cos(4 * math.pi * x)
    raise ValueError(f"unknown window {type}")


def toneBurst(sampleFreq: float, signalFreq: float, numCycles: float, envelope: str = "gaussian") -> np.ndarray:
    """Enveloped tone burst (Generation.swift:29-49). `envelope="hann"` uses getWin(.hann) as the
    envelope, the one extension the Hann-windowed workload needs (not in the reference)."""
    dt = 1.0 / sampleFreq
    count = int(math.floor((numCycles / signalFreq) / dt)) + 1
    burst = np.sin(2 * math.pi * signalFreq * np.arange(count) * dt)
    if envelope == "rectangular":
        win = np.ones(count)
    elif envelope == "gaussian":
        step = 2 * 3.0 / (count - 1) if count > 1 else 0.0
        xs = -3.0 + np.arange(count) * step
        win = gaussian(xs, magnitude=1.0, mean=0.0, variance=1.0)
    elif envelope == "hann":
        win = getWin(count, "hann")
    else:
        raise ValueError(f"unknown envelope {envelope}")
    return burst * win


def db2neper(alpha: float, y: float = 1.0) -> float:
    """dB/(MHz^y cm) -> Np/((rad/s)^y m) (Conversion.swift:6-8)."""
    return 100 * alpha * (1e-6 / (2 * math.pi)) ** y / (20 * math.log10(math.e))


def neper2db(alpha: float, y: float = 1.0) -> float:
    """Np/((rad/s)^y m) -> dB/(MHz^y cm) (Conversion.swift:13-15)."""
    return 20 * math.log10(math.e) * alpha * (2 * math.pi * 1e6) ** y / 100


# ---------------------------------------------------------------------------------------------
# Seeded synthetic media
# ---------------------------------------------------------------------------------------------

_GAMMA = np.uint64(0x9E3779B97F4A7C15)


def splitmix64(x: np.ndarray) -> np.ndarray:
    """SplitMix64 output mix of a uint64 array (Steele, Lea & Flood 2014)."""
    z = x.astype(np.uint64)
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def uniform01(seed: int, n: int) -> np.ndarray:
    """Counter-based uniforms in (0, 1]: u_i = ((splitmix64(seed + (i+1)*gamma) >> 11) + 1) * 2^-53."""
    with np.errstate(over="ignore"):
        i = np.arange(1, n + 1, dtype=np.uint64)
        x = np.uint64(seed) + i * _GAMMA
        return ((splitmix64(x) >> np.uint64(11)).astype(np.float64) + 1.0) * 2.0 ** -53


def random_field(shape, seed: int, sigma_cells: float, lo: float, hi: float) -> np.ndarray:
    """Smooth Gaussian random field min-max mapped to [lo, hi], float32.

    Box-Muller normals from counter-based SplitMix64 uniforms, a periodic Gaussian k-space filter
    exp(-|k|^2 sigma^2 / 2) (k in rad/cell) in float64, min-max rescale, round to float32.
    """
    shape = tuple(int(s) for s in shape)
    n = int(np.prod(shape))
    m = n + (n & 1)
    u = uniform01(seed, m).reshape(-1, 2)
    rad = np.sqrt(-2.0 * np.log(u[:, 0]))
    z = np.empty(m)
    z[0::2] = rad * np.cos(2 * math.pi * u[:, 1])
    z[1::2] = rad * np.sin(2 * math.pi * u[:, 1])
    f = z[:n].reshape(shape)
    k2 = np.zeros(shape)
    for a, s in enumerate(shape):
        kk = 2 * math.pi * np.fft.fftfreq(s)
        sh = [1] * len(shape)
        sh[a] = s
        k2 = k2 + (kk ** 2).reshape(sh)
    g = np.real(np.fft.ifftn(np.fft.fftn(f) * np.exp(-k2 * sigma_cells ** 2 / 2)))
    g = (g - g.min()) / (g.max() - g.min())
    return (lo + (hi - lo) * g).astype(np.float32)
