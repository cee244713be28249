"""Float64 oracle of the discrete adjoint (VJP) of the lossless, linear k-space first-order loop.

TEST INFRASTRUCTURE ONLY: used by tests/ and nowhere in the product path.

The reference has no adjoint. This module restates its forward loop for the subset the
gradient is defined on -- p0 initial value problem, lossless, linear, no time-varying sources,
sensor recording p (binary or Cartesian) -- in float64 on float64 inputs
(KSpaceFirstOrder.swift:557-761 with the t = 0 override :724-733; the operators are those of
kwave_oracle.kspace_first_order), and derives the exact transpose of that discrete map:

forward, n >= 1:  u_a^n = s_a (s_a u_a^{n-1} - r_a D+_a p^{n-1})      s_a: staggered PML, r_a = dt/rho0_a
                  rho_a^n = q_a (q_a rho_a^{n-1} - m D-_a u_a^n)       q_a: PML, m = dt rho0
                  p^n = c^2 sum_a rho_a^n,   y^n = S p^n
          n = 0:  p^0 = p0, rho_a^0 = p0 / (d c^2), u_a^0 = (r_a / 2) D+_a p0,   y^0 = S p0
with D+-_a f = Re F^-1(i k_a e^{+-i k_a d_a/2} kappa F f). For real fields (D+_a)^T = -D-_a and
(D-_a)^T = -D+_a (the conjugate operator), so the reverse recursion (lambda^nt = 0) is

          lp^n   = g^n + sum_a D-_a[r_a s_a lu_a^{n+1}]
          lr_a^n = c^2 lp^n + q_a^2 lr_a^{n+1}                     (n >= 1)
          lu_a^n = s_a^2 lu_a^{n+1} + D+_a[m q_a lr_a^n]           (n >= 1)
          dL/dc += 2 c lp^n sum_a rho_a^n = 2 lp^n p^n / c         (n >= 1)
          dL/dp0 = lp^0 + sum_a q_a^2 lr_a^1 / (d c^2) - sum_a D-_a[(r_a/2) s_a^2 lu_a^1]
          dL/dc += -2 p0 / (d c^3) sum_a q_a^2 lr_a^1

with g^n = S^T dL/dy^n. c_ref = max(c) (kappa, PML) is held fixed: the gradient is the partial
derivative at fixed c_ref. Pinned by tests/test_adjoint_oracle.py: the dot-product identity
<J v, w> = <v, J^T w> for p0, central finite differences for c, and equality of this forward
with kwave_oracle.kspace_first_order.
"""
import numpy as np

from . import kwave_oracle as O


class LosslessOperators:
    """Operators of the lossless loop for one grid / medium / sensor (c_ref from c_ref_from)."""

    def __init__(self, grid, rho0, sensor_mask, options, c_ref, workers=1):
        self.fft = O._NumpyFFT(workers)
        self.dim = grid.dim
        self.sizes = list(grid.size)
        self.spac = list(grid.spacing)
        self.shape = tuple(self.sizes)
        self.dt = float(grid.dt)
        self.nt = int(grid.nt)
        dim, sizes, spac, dt = self.dim, self.sizes, self.spac, self.dt
        kmag = O.k_magnitude(sizes, spac)
        arg = kmag * (c_ref * dt / 2.0)
        kappa = np.where(arg == 0, 1.0, np.sin(arg) / np.where(arg == 0, 1.0, arg))
        kv = O.k_vectors(sizes, spac)
        self.dpos = [O._bcast(O.derivative_operator(kv[a], spac[a], +1), a, dim) * kappa for a in range(dim)]
        self.dneg = [O._bcast(O.derivative_operator(kv[a], spac[a], -1), a, dim) * kappa for a in range(dim)]
        n = int(options.pmlSize)
        self.q = [O._bcast(O.pml_profile(sizes[a], spac[a], dt, c_ref, n, options.pmlAlpha, False), a, dim)
                  for a in range(dim)]
        self.s = [O._bcast(O.pml_profile(sizes[a], spac[a], dt, c_ref, n, options.pmlAlpha, True), a, dim)
                  for a in range(dim)]
        rho0 = np.asarray(rho0, np.float64)
        rho0 = rho0.reshape(()) if rho0.size == 1 else rho0
        self.r = [dt / O.stagger_density(rho0, a) for a in range(dim)]
        self.m = dt * rho0
        self.sampler = O.make_sampler(sensor_mask, sizes, spac)

    def D(self, op, f):
        return np.real(self.fft.ifftn(op * self.fft.fftn(f)))

    def sample(self, p):
        return self.sampler.sample(p)

    def sample_T(self, g):
        out = np.zeros(int(np.prod(self.shape)))
        sm = self.sampler
        if sm.weight is None:
            np.add.at(out, sm.idx, g)
        else:
            np.add.at(out, sm.idx, (g[:, None] * sm.weight.reshape(sm.n_points, sm.n_corners)).reshape(-1))
        return out.reshape(self.shape)


def forward(ops: LosslessOperators, c, p0, keep_p=False):
    """y [nSensor, nt] (and the pressure history when keep_p)."""
    d, c2 = ops.dim, np.asarray(c, np.float64) ** 2
    p = np.asarray(p0, np.float64).reshape(ops.shape)
    rho = [p / (d * c2) for _ in range(d)]
    u = [(ops.r[a] / 2.0) * ops.D(ops.dpos[a], p) for a in range(d)]
    ys, hist = [ops.sample(p)], [p]
    for _ in range(1, ops.nt):
        u = [ops.s[a] * (ops.s[a] * u[a] - ops.r[a] * ops.D(ops.dpos[a], p)) for a in range(d)]
        rho = [ops.q[a] * (ops.q[a] * rho[a] - ops.m * ops.D(ops.dneg[a], u[a])) for a in range(d)]
        p = c2 * sum(rho)
        ys.append(ops.sample(p))
        if keep_p:
            hist.append(p)
    y = np.stack(ys, axis=1)
    return (y, hist) if keep_p else y


def vjp(ops: LosslessOperators, c, p0, dLdy):
    """(dL/dp0, dL/dc) for the cotangent dLdy [nSensor, nt] of the recorded pressure."""
    d = ops.dim
    c = np.asarray(c, np.float64) * np.ones(ops.shape)
    c2 = c * c
    p0 = np.asarray(p0, np.float64).reshape(ops.shape)
    _, hist = forward(ops, c, p0, keep_p=True)
    dLdy = np.asarray(dLdy, np.float64)
    lu = [np.zeros(ops.shape) for _ in range(d)]
    lr = [np.zeros(ops.shape) for _ in range(d)]
    gc = np.zeros(ops.shape)
    for n in range(ops.nt - 1, 0, -1):
        lp = ops.sample_T(dLdy[:, n]) + sum(ops.D(ops.dneg[a], ops.r[a] * ops.s[a] * lu[a]) for a in range(d))
        lr = [c2 * lp + ops.q[a] ** 2 * lr[a] for a in range(d)]
        lu = [ops.s[a] ** 2 * lu[a] + ops.D(ops.dpos[a], ops.m * ops.q[a] * lr[a]) for a in range(d)]
        gc += 2.0 * lp * hist[n] / c
    lp0 = ops.sample_T(dLdy[:, 0]) + sum(ops.D(ops.dneg[a], ops.r[a] * ops.s[a] * lu[a]) for a in range(d))
    lr0 = sum(ops.q[a] ** 2 * lr[a] for a in range(d))
    gp0 = lp0 + lr0 / (d * c2) - sum(ops.D(ops.dneg[a], (ops.r[a] / 2.0) * ops.s[a] ** 2 * lu[a]) for a in range(d))
    gc += -2.0 * p0 / (d * c2 * c) * lr0
    return gp0, gc
